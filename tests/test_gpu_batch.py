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
@pytest.mark.parametrize("shape", [(1, 16), (0, 32), None])
@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_matches_single_grid_runs(dtype, shape, use_graph):
    from wakis_amd import batch
    print("batch", np.dtype(dtype).name, shape, use_graph, flush=True)
    singles, batched = _pairs(dtype)
    wl = [0.02 + 0.01 * i for i in range(len(singles))]  # uneven step counts
    for s, w in zip(singles, wl):
        s.wakesolve(wakelength=w)
    kw = dict(autotune=True, tune_steps=8) if shape is None else dict(
        cells_per_lane=shape[0], kz=shape[1])
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
