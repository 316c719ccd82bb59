"""GPU: two leapfrog steps per launch (k_fused4_t2) against one step per launch,
FP32 (8 or 4 rows per workgroup) and FP64 (2 rows, two waves per row).

The two-step kernel forms every value with the single-step kernel's
expressions, so E, H, J and the captured Ez window must be bit-identical to
one step per launch (which the other tests tie bitwise to the oracle), for
both row counts, z chunks that split the grid unevenly, rows that do not fill
the last tile, odd and even step counts, and Nx below and at the 256-cell
row. The wakefield run of the cavity is also compared with the oracle
directly.
"""

import numpy as np
import pytest
import torch

from oracle import cases
from oracle import fit_oracle as fo
from oracle import wake_oracle as wo

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _cavity_solver(dims, two_step, kz=0, dtype=np.float32):
    import wakis_amd as wa
    from wakis_amd.solids import Box, Cylinder
    solid = Box(-25e-3, 25e-3, -25e-3, 25e-3, -25e-3, 25e-3) | Cylinder(5e-3)
    grid = wa.GridFIT3D(-30e-3, 30e-3, -30e-3, 30e-3, -80e-3, 80e-3, *dims,
                        stl_solids={"cavity": solid}, stl_materials={"cavity": "vacuum"},
                        verbose=0)
    s = wa.SolverFIT3D(grid, bc_low=["pec", "pec", "abc"], bc_high=["pec", "pec", "abc"],
                       use_stl=True, bg="pec", dtype=dtype, verbose=0)
    s.fused_two_step = two_step
    s.fused_kz = kz
    return s


def _pillbox_solver(dims, two_step, kz=0, dtype=np.float32):
    """The lossy pillbox: a 3-cell conductive wall (sigma 10 S/m), 10 PML layers
    on z; conduction everywhere (J = sigma E)."""
    import wakis_amd as wa
    x, y, z, vac, wall = cases.pillbox_geometry(*dims, r_cav=20e-3, half_len=15e-3,
                                                r_pipe=5e-3, wall_cells=3, half_x=25e-3,
                                                half_z=45e-3)
    grid = wa.GridFIT3D(x=x, y=y, z=z, stl_solids={"wall": wall, "vac": vac},
                        stl_materials={"wall": [1.0, 1.0, 10.0], "vac": "vacuum"}, verbose=0)
    s = wa.SolverFIT3D(grid, bc_low=["pec", "pec", "pml"], bc_high=["pec", "pec", "pml"],
                       use_stl=True, bg="pec", dtype=dtype, verbose=0, n_pml=10,
                       use_sibc=False)
    assert s.use_conductivity
    s.fused_two_step = two_step
    s.fused_kz = kz
    return s


def _beam_setup(s, tot):
    """Beam-table steps with a 3x3 Ez capture: returns make(n0, k), the
    arguments of steps n0 .. n0+k-1, and the capture buffer."""
    import wakis_amd as wa
    from wakis_amd import _lib
    beam = wa.Beam(q=1e-9, sigmaz=10e-3, beta=1.0)
    dev = s.device
    table = torch.as_tensor(beam.table(s, tot), device=dev).contiguous()
    cap = torch.zeros((9, tot + 1, s.Nz), dtype=table.dtype, device=dev)
    ixt, iyt = int(np.abs(s.x).argmin()), int(np.abs(s.y).argmin())
    s._set_ghosts_to_0()
    s.step_0 = False
    s._beam_line = (int(beam.ixs), int(beam.iys))

    def make(n0, k):
        a = s._step_args(k)
        a.ixs, a.iys = int(beam.ixs), int(beam.iys)
        if a.jmode == _lib.WF_J_NONE:
            a.jmode = _lib.WF_J_LINE
        a.beam_table, a.beam_row0, a.beam_rows = _lib.ptr(table), n0, tot
        a.capture = _lib.ptr(cap)
        a.cap_ix0, a.cap_iy0, a.cap_nx, a.cap_ny = ixt - 1, iyt - 1, 3, 3
        a.cap_samples, a.cap_sample0 = tot + 1, n0 + 1
        return a
    return make, cap


def _beam_run(dims, two_step, chunks, kz=0, make=_cavity_solver, tune=False,
              dtype=np.float32):
    """Steps with the beam table and a 3x3 Ez capture, in calls of `chunks` steps."""
    s = make(dims, two_step, kz, dtype)
    args, cap = _beam_setup(s, sum(chunks))
    if tune:
        s.tune_fused(lambda k: args(0, k), steps=4)
    n0 = 0
    for k in chunks:
        s._run(args(n0, k))
        assert s.last_path == "fused"
        if not tune:  # the default (1) applies from 2^20 cells (FP64: and Nx >= 224)
            wide = np.dtype(s.dtype) == np.float32 or s.Nx >= 224
            big = two_step != 1 or (s.N >= 1 << 20 and wide)
            assert s.last_two_step == (two_step != 0 and big and k >= 2 and s.Nx <= 256)
        n0 += k
    torch.cuda.synchronize()
    return (s.E.to_numpy(), s.H.to_numpy(), s.J.to_numpy(), cap.cpu().numpy()), s


@pytest.mark.parametrize("dims,chunks,kz", [
    ((256, 24, 64), (12,), 0),
    ((256, 24, 100), (5, 8), 3),
    ((64, 40, 48), (11,), 0),
    ((20, 13, 30), (7, 6), 4),
    ((36, 9, 7), (9,), 2),
    ((256, 256, 40), (10,), 0),
    ((64, 16, 40), (9,), 5),  # the 3x3 capture (rows 7..9) straddles two row tiles
])
@pytest.mark.parametrize("dtype,two_step", [(np.float32, 4), (np.float32, 8),
                                            (np.float64, 8)])
def test_two_step_equals_single_steps(dims, chunks, kz, dtype, two_step):
    _say("two-step", dims, chunks, kz, np.dtype(dtype).name, two_step)
    ref = _beam_run(dims, 0, chunks, kz, dtype=dtype)[0]
    got = _beam_run(dims, two_step, chunks, kz, dtype=dtype)[0]
    assert np.any(ref[0] != 0) and np.any(ref[3] != 0)
    for name, r, g in zip(("E", "H", "J", "capture"), ref, got):
        assert np.array_equal(g, r), f"{name} differs ({np.count_nonzero(g != r)} values)"


@pytest.mark.parametrize("dims,chunks,kz", [
    ((40, 40, 80), (12,), 0),
    ((200, 24, 60), (5, 8), 4),
    ((100, 100, 50), (7,), 0),
    ((256, 20, 70), (6, 5), 9),  # full 256-cell rows, a partial row tile, uneven chunks
])
@pytest.mark.parametrize("dtype,two_step", [(np.float32, 4), (np.float32, 8),
                                            (np.float64, 8)])
def test_two_step_conductive_equals_single_steps(dims, chunks, kz, dtype, two_step):
    """With conduction (wall and PML): J = sigma E recomputed from E, stored by
    the last launch of each call."""
    _say("two-step conductive", dims, chunks, kz, np.dtype(dtype).name, two_step)
    ref = _beam_run(dims, 0, chunks, kz, make=_pillbox_solver, dtype=dtype)[0]
    got = _beam_run(dims, two_step, chunks, kz, make=_pillbox_solver, dtype=dtype)[0]
    assert np.any(ref[0] != 0) and np.any(ref[2] != 0)
    for name, r, g in zip(("E", "H", "J", "capture"), ref, got):
        assert np.array_equal(g, r), f"{name} differs ({np.count_nonzero(g != r)} values)"


def test_two_step_applies_only_where_built():
    """Runs outside the kernel's case (FP64, coefficient arrays, wide rows) fall
    back to one step per launch with the same results."""
    _say("two-step fallback")
    ref = _beam_run((260, 12, 20), 0, (6,))[0]
    got = _beam_run((260, 12, 20), 8, (6,))[0]  # Nx > 256: single steps
    for r, g in zip(ref, got):
        assert np.array_equal(g, r)


def test_two_step_default_from_2_20_cells():
    """The default (1) runs two steps per launch from 2^20 cells on (FP64 also
    from 224 cells per row), one step per launch below; the results are the
    same."""
    _say("two-step default by size")
    for dims, dtype, two in (((64, 40, 48), np.float32, False), ((256, 64, 64), np.float32, True),
                             ((160, 80, 82), np.float64, False),
                             ((256, 64, 64), np.float64, True)):
        ref = _beam_run(dims, 0, (6,), dtype=dtype)[0]
        got, s = _beam_run(dims, 1, (6,), dtype=dtype)
        assert s.last_two_step == two
        for r, g in zip(ref, got):
            assert np.array_equal(g, r)


def test_tune_fused_times_two_step_launches():
    """tune_fused() times two steps per launch (-2) beside the one-step shapes,
    sets fused_two_step to 0 when one step of row segments wins, restores the
    fields, and the run that follows equals one step per launch bitwise."""
    _say("two-step tuner")
    dims, chunks = (256, 64, 64), (5, 6)
    ref = _beam_run(dims, 0, chunks)[0]
    got, s = _beam_run(dims, 1, chunks, tune=True)
    t = s.fused_tuning
    assert -2 in t and -1 in t and len(t) >= 4
    best = min(t, key=t.get)
    assert s.fused_full_rows == max(best, -1)
    assert s.fused_two_step == (0 if best == -1 else 1)
    assert s.last_two_step == (best == -2)
    for name, r, g in zip(("E", "H", "J", "capture"), ref, got):
        assert np.array_equal(g, r), f"{name} differs ({np.count_nonzero(g != r)} values)"


@pytest.mark.parametrize("dtype,two_step", [(np.float32, 4), (np.float32, 8),
                                            (np.float64, 8)])
def test_two_step_wakesolve_matches_oracle(dtype, two_step):
    import wakis_amd as wa
    _say("two-step wakesolve vs oracle", np.dtype(dtype).name, two_step)
    x, y, z, mask = cases.cavity_pipe_geometry(20, 20, 40)
    so = fo.OracleSolver(x, y, z, ["pec", "pec", "abc"], ["pec", "pec", "abc"], bg="pec",
                         solids=[(mask, "vacuum")], dtype=dtype)
    q, sigmaz, beta, wl = 1e-9, 10e-3, 1.0, 0.05
    cap = fo.wakesolve_fields(so, wl, q, sigmaz, beta, 0.0, 0.0, 0.0, 0.0,
                              wo.wake_ti(sigmaz, beta))
    ref = wo.wake_solve(cap, q, sigmaz, beta, wl, skip_cells=2)
    grid = wa.GridFIT3D(x=x, y=y, z=z, stl_solids={"cav": mask},
                        stl_materials={"cav": "vacuum"}, verbose=0)
    wake = wa.WakeSolver(q=q, sigmaz=sigmaz, beta=beta, skip_cells=2, save=False, verbose=0)
    sa = wa.SolverFIT3D(grid, wake=wake, bc_low=["pec", "pec", "abc"],
                        bc_high=["pec", "pec", "abc"], use_stl=True, bg="pec",
                        dtype=dtype, verbose=0)
    sa.fused_two_step = two_step
    sa.wakesolve(wakelength=wl)
    for name in ("E", "H", "J"):
        assert np.array_equal(getattr(sa, name).to_numpy(), getattr(so, name)), name
    assert np.array_equal(wake.WP_3d, ref["WP_3d"])


def test_energy_after_two_step_launches():
    """energy() needs the fields before the last step: after an odd count the
    last launch is a single step and the energy equals that of single steps;
    after an even count (the last launch a pair) it is refused."""
    import wakis_amd as wa
    from wakis_amd import _lib
    _say("two-step energy")
    out = {}
    for two_step, chunks in ((0, (7,)), (8, (7,)), (8, (6,))):
        s = _cavity_solver((64, 40, 48), two_step)
        beam = wa.Beam(q=1e-9, sigmaz=10e-3, beta=1.0)
        table = torch.as_tensor(beam.table(s, 16), device=s.device).contiguous()
        s._set_ghosts_to_0()
        s.step_0 = False
        s._beam_line = (int(beam.ixs), int(beam.iys))
        a = s._step_args(chunks[0])
        a.ixs, a.iys = int(beam.ixs), int(beam.iys)
        a.jmode = _lib.WF_J_LINE
        a.beam_table, a.beam_row0, a.beam_rows = _lib.ptr(table), 0, 16
        s._run(a)
        assert s.last_two_step == (two_step != 0)
        if chunks[0] % 2 == 0 and two_step:
            with pytest.raises(RuntimeError):
                s.energy()
        else:
            out[two_step] = s.energy()
    assert out[0] == out[8] and out[0] > 0
