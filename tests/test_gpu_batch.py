"""GPU: batched wakefield runs (one launch per step over many grids).

Every grid of a batch must end bit-identical to its own single-grid
wakesolve (fields and wake potentials), for each tile shape, with and
without the captured HIP graph, and with grids of different step counts.
"""

import numpy as np
import pytest

from tests import gpu_cases as gc

pytestmark = pytest.mark.gpu


def _pairs(dtype, n=5, seed=3):
    from wakis_amd import batch
    specs = batch.random_box_cavities(seed, n=n, cell=2.0e-3)
    return ([batch.cavity_solver(sp, dtype=dtype) for sp in specs],
            [batch.cavity_solver(sp, dtype=dtype) for sp in specs])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(1, 16), (0, 32), (0, 16, "rows"), (0, 32, "rows"), None])
@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("codes", [True, False])
def test_batch_matches_single_grid_runs(dtype, shape, use_graph, codes):
    from wakis_amd import batch
    print("batch", np.dtype(dtype).name, shape, use_graph, codes, flush=True)
    singles, batched = _pairs(dtype)
    for s in singles + batched:
        s.use_material_code = codes
    wl = [0.02 + 0.01 * i for i in range(len(singles))]  # uneven step counts
    for s, w in zip(singles, wl):
        s.wakesolve(wakelength=w)
    kw = dict(autotune=True, tune_steps=8) if shape is None else dict(
        cells_per_lane=shape[0], kz=shape[1], full_rows=len(shape) > 2)
    info = batch.wakesolve_batch(batched, wl, use_graph=use_graph, **kw)
    assert len({s.Nt for s in singles}) > 1
    assert info["steps"] == max(s.Nt for s in singles)
    for a, b in zip(singles, batched):
        assert a.last_path == "fused" and b.last_path == "fused"
        assert a.Nt == b.Nt
        for name in ("E", "H", "J"):
            u, v = getattr(a, name).to_numpy(), getattr(b, name).to_numpy()
            assert np.array_equal(u, v), f"{name}: rel-L2 {gc.rel_l2(v, u):.3e}"
        assert np.array_equal(a.wake.WP_3d, b.wake.WP_3d)
        assert np.array_equal(a.wake.WPx, b.wake.WPx)


def test_batch_refuses_grids_off_the_fast_path():
    from wakis_amd import batch
    singles, _ = _pairs(np.float32, n=2)
    singles[1].E[0, 3, 3, "y"] = 1.0  # tangential E on the x-low PEC face: masked value != 0
    with pytest.raises(batch.BatchNotApplicable):
        batch.wakesolve_batch(singles, 0.02)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("per_cell", [False, True])
@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("full_rows", [None, False])
def test_adjoint_batch_matches_single_problems(dtype, per_cell, use_graph, full_rows):
    """Batched forward (with the block trajectories captured in the step loop)
    and batched reverse steps give each problem's own loss and gradient, bit for
    bit. per_cell: random eps_r per block cell (the batch reads coefficient
    arrays for those grids) instead of one eps_r per block."""
    import torch
    from wakis_amd import batch
    from wakis_amd.adjoint import WakeAdjoint
    print("adjoint batch", np.dtype(dtype).name, per_cell, use_graph, full_rows, flush=True)
    specs = batch.random_box_cavities(5, n=4, cell=2.0e-3)

    def problems():
        return [WakeAdjoint(batch.cavity_solver(sp, dtype=dtype), batch.dielectric_region(sp),
                            0.02 + 0.01 * i, w_loss=1.0, w_band=1e-13)
                for i, sp in enumerate(specs)]

    singles, batched = problems(), problems()
    rng = np.random.default_rng(2)
    ers = [rng.uniform(1.0, 6.0, adj.shape) if per_cell else
           np.full(adj.shape, sp["eps_r"]) for adj, sp in zip(singles, specs)]
    ref = []
    for adj, er in zip(singles, ers):
        L = adj.forward(er)
        g = adj.backward()
        assert adj.used_kernel == "fused" and adj.used_mode == "trajectory"
        ref.append((L, g))
    assert len({adj.s.Nt for adj in singles}) > 1
    s0 = singles[0].s  # per-cell eps_r: escape codes (or arrays) in the single runs
    assert (s0._code is None or s0._code_escape) == per_cell
    losses, grads = batch.wake_adjoint_batch(batched, ers, use_graph=use_graph,
                                             full_rows=full_rows)
    for (L, g), Lb, gb in zip(ref, losses, grads):
        assert Lb == L
        assert gb[1] is None
        assert torch.equal(g[0], gb[0])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_batch_matches_the_oracle(dtype):
    """Each grid of a batch (uneven sizes and step counts, the in-run tuner
    choosing the tile shape) against the oracle's wakesolve of the same cavity:
    E, H, J and WP_3d, WP_x, WP_y bit-identical, Z within 1e-12."""
    from oracle import fit_oracle as fo
    from oracle import wake_oracle as wo
    from wakis_amd import batch
    print("batch vs oracle", np.dtype(dtype).name, flush=True)
    specs = batch.random_box_cavities(11, n=3, cell=2.0e-3)
    solvers = [batch.cavity_solver(sp, dtype=dtype) for sp in specs]
    wl = [0.015 + 0.01 * i for i in range(len(specs))]
    batch.wakesolve_batch(solvers, wl, autotune=True, tune_steps=8)
    q, sigmaz, beta = 1e-9, 5e-3, 1.0
    ti = wo.wake_ti(sigmaz, beta)
    for sp, s, w in zip(specs, solvers, wl):
        so = fo.OracleSolver(sp["x"], sp["y"], sp["z"], ["pec", "pec", "abc"],
                             ["pec", "pec", "abc"], bg="pec",
                             solids=[(sp["vacuum"], "vacuum"), (sp["dielectric"],
                                                                [sp["eps_r"], 1.0])],
                             dtype=dtype)
        cap = fo.wakesolve_fields(so, w, q, sigmaz, beta, 0.0, 0.0, 0.0, 0.0, ti)
        ref = wo.wake_solve(cap, q, sigmaz, beta, w, skip_cells=4)
        assert s.Nt == cap["Nt"] and s.dt == so.dt
        for name in ("E", "H", "J"):
            got, r = getattr(s, name).to_numpy(), getattr(so, name)
            assert np.array_equal(got, r), f"{name}: rel-L2 {gc.rel_l2(got, r):.3e}"
        assert np.array_equal(s.wake.WP_3d, ref["WP_3d"])
        assert np.array_equal(s.wake.WPx, ref["WPx"]) and np.array_equal(s.wake.WPy, ref["WPy"])
        assert np.max(np.abs(s.wake.Z - ref["Z"])) <= 1e-12 * np.max(np.abs(ref["Z"]))
