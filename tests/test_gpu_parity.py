"""GPU parity: the HIP engine against the oracle (bit-exact where the
arithmetic is the same sequence of IEEE operations, tolerances stated where
not: 1e-12 relative L2 on FP64 fields, 1e-12 on the DFT bins)."""

import json
import math

import numpy as np
import pytest

from oracle import cases
from oracle import fit_oracle as fo
from oracle import wake_oracle as wo
from tests import gpu_cases as gc
from tests.test_stencil_rules import BCS

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-12


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("bci", range(len(BCS)))
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("conductive", [False, True])
def test_step_matches_oracle(bci, dtype, conductive):
    _say("case", bci, np.dtype(dtype).name, conductive)
    case = gc.random_case(BCS[bci], dtype, conductive, seed=11 + bci)
    so = gc.oracle_solver(case)
    sa = gc.amd_solver(case, path="general")
    assert sa.dt == so.dt and type(sa.dt) is type(so.dt)
    for _ in range(5):
        so.one_step()
        sa.one_step()
    for name in ("E", "H", "J"):
        got = getattr(sa, name).to_numpy()
        ref = getattr(so, name)
        assert gc.rel_l2(got, ref) <= FIELD_TOL, name
        assert np.array_equal(got, ref), f"{name} not bit-identical to the oracle"


def test_beam_line_mode_matches_full_mode():
    """J only on the beam line (WF_J_LINE) equals the general J read (WF_J_FULL)."""
    import wakis_amd as wa
    x, y, z, mask = cases.cavity_pipe_geometry(16, 16, 24)
    res = []
    for mode in ("line", "full"):
        grid = wa.GridFIT3D(x=x, y=y, z=z, stl_solids={"cav": mask},
                            stl_materials={"cav": "vacuum"}, verbose=0)
        s = wa.SolverFIT3D(grid, bc_low=["pec", "pec", "abc"], bc_high=["pec", "pec", "abc"],
                           use_stl=True, bg="pec", dtype=np.float64, verbose=0)
        beam = wa.Beam(q=1e-9, sigmaz=10e-3, beta=1.0)
        for n in range(40):
            beam.update(s, n * s.dt)
            if mode == "full":
                s.J.version += 1  # force the general J read
            s.one_step()
        res.append((s.E.to_numpy(), s.H.to_numpy()))
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])


def _run_both_wakes(Nx, Ny, Nz, dtype, wakelength, skip, plane="both"):
    import wakis_amd as wa
    x, y, z, mask = cases.cavity_pipe_geometry(Nx, Ny, Nz)
    so = fo.OracleSolver(x, y, z, ["pec", "pec", "abc"], ["pec", "pec", "abc"], bg="pec",
                         solids=[(mask, "vacuum")], dtype=dtype)
    q, sigmaz, beta = 1e-9, 10e-3, 1.0
    ti = wo.wake_ti(sigmaz, beta)
    cap = fo.wakesolve_fields(so, wakelength, q, sigmaz, beta, 0.0, 0.0, 0.0, 0.0, ti,
                              compute_plane=plane)
    ref = wo.wake_solve(cap, q, sigmaz, beta, wakelength, skip_cells=skip, compute_plane=plane)
    grid = wa.GridFIT3D(x=x, y=y, z=z, stl_solids={"cav": mask},
                        stl_materials={"cav": "vacuum"}, verbose=0)
    wake = wa.WakeSolver(q=q, sigmaz=sigmaz, beta=beta, skip_cells=skip, save=False, verbose=0)
    sa = wa.SolverFIT3D(grid, wake=wake, bc_low=["pec", "pec", "abc"],
                        bc_high=["pec", "pec", "abc"], use_stl=True, bg="pec", dtype=dtype,
                        verbose=0)
    sa.wakesolve(wakelength=wakelength, compute_plane=plane)
    assert sa.last_path == "fused", "the cavity/pipe wakefield run should use the fused kernel"
    return so, sa, ref, wake


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_wakesolve_matches_oracle(dtype):
    _say("wakesolve 20x20x40", np.dtype(dtype).name)
    so, sa, ref, wake = _run_both_wakes(20, 20, 40, dtype, wakelength=0.05, skip=2)
    assert np.array_equal(sa.E.to_numpy(), so.E) and np.array_equal(sa.H.to_numpy(), so.H)
    assert np.array_equal(sa.J.to_numpy(), so.J)
    assert np.array_equal(wake.s, ref["s"])
    assert np.array_equal(wake.WP_3d, ref["WP_3d"]), "WP_3d not bit-identical"
    assert np.array_equal(wake.WPx, ref["WPx"]) and np.array_equal(wake.WPy, ref["WPy"])
    assert np.array_equal(wake.f, ref["f"])
    for k in ("Z", "Zx", "Zy", "lambdaf"):
        assert gc.rel_l2(np.abs(getattr(wake, k)), np.abs(ref[k])) < 1e-12, k
        assert np.max(np.abs(getattr(wake, k) - ref[k])) <= 1e-12 * np.max(np.abs(ref[k])), k


def test_wakesolve_longitudinal_matches_oracle():
    _say("wakesolve longitudinal")
    so, sa, ref, wake = _run_both_wakes(16, 16, 32, np.float64, wakelength=0.04, skip=0,
                                        plane="longitudinal")
    assert np.array_equal(wake.WP, ref["WP"])
    assert np.max(np.abs(wake.Z - ref["Z"])) <= 1e-12 * np.max(np.abs(ref["Z"]))


def test_dft_kernel_vs_numpy_fft():
    import torch
    from wakis_amd import _lib
    rng = np.random.default_rng(5)
    L, N, nb = 1623, 166166, 1001
    x = rng.standard_normal((2, L))
    xt = torch.as_tensor(x, device="cuda")
    out = torch.zeros((2, nb, 2), dtype=torch.float64, device="cuda")
    _lib.check(_lib.lib().wf_dft(_lib.ptr(xt), 2, L, N, nb, _lib.ptr(out), _lib.stream_ptr()))
    got = out.cpu().numpy()
    got = got[..., 0] + 1j * got[..., 1]
    ref = np.fft.fft(x, n=N, axis=1)[:, :nb]
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    assert err < 1e-13, err


def test_test001_golden_on_gpu():
    """The reference's test_001 golden vectors (FP32) through the device wakesolve."""
    import wakis_amd as wa
    _say("test_001 golden, 50x50x150 FP32, 5184 steps")
    x, y, z, mask, g = cases.test001_geometry()
    b = g["beam"]
    grid = wa.GridFIT3D(x=x, y=y, z=z, stl_solids={"cavity": mask},
                        stl_materials={"cavity": "vacuum"}, verbose=0)
    wake = wa.WakeSolver(q=b["q"], sigmaz=b["sigmaz"], beta=b["beta"], xsource=b["xs"],
                         ysource=b["ys"], xtest=b["xt"], ytest=b["yt"], save=False,
                         skip_cells=b["skip_cells"], verbose=0)
    s = wa.SolverFIT3D(grid, wake, bc_low=g["bc"]["low"], bc_high=g["bc"]["high"], use_stl=True,
                       bg="pec", dtype=np.float32, verbose=0)
    s.wakesolve(wakelength=b["wakelength"])
    assert s.Nt == 5184
    WPg = np.array(g["WP_every50"])
    Zg = np.array(g["Z_every20_re_im"])
    Zg = Zg[:, 0] + 1j * Zg[:, 1]
    tw, tz = g["tol_WP"], g["tol_Z"]
    assert np.allclose(wake.WP[::50], WPg, **tw)
    assert np.allclose(np.abs(wake.Z)[::20], np.abs(Zg), **tz)
    assert np.allclose(np.real(wake.Z)[::20], np.real(Zg), **tz)
    assert np.allclose(np.imag(wake.Z)[::20], np.imag(Zg), **tz)
    assert np.cumsum(np.abs(wake.WP))[-1] == pytest.approx(g["cumsum_abs_WP"], g["cumsum_rel"])
    assert np.cumsum(np.abs(wake.Z))[-1] == pytest.approx(g["cumsum_abs_Z"], g["cumsum_rel"])


@pytest.mark.parametrize("dims", [(1, 4, 6), (2, 1, 5), (3, 3, 1), (4, 2, 2), (5, 7, 3),
                                  (2, 3, 4), (2, 2, 2), (8, 8, 2), (2, 9, 70)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("conductive", [False, True])
def test_degenerate_grids_match_oracle(dims, dtype, conductive):
    """Grids two cells thin along an axis (the step's neighbour truncation,
    tile edges and z-march limits at their smallest) against the oracle, on
    whichever path the engine picks; one-cell axes are refused with a
    ValueError (the engine needs two cells per axis)."""
    _say("degenerate", dims, np.dtype(dtype).name, conductive)
    case = gc.random_case(BCS[0], dtype, conductive, seed=sum(dims) + 5, dims=dims)
    if min(dims) < 2:
        with pytest.raises(ValueError, match="at least 2 cells"):
            gc.amd_solver(case)
        return
    so = gc.oracle_solver(case)
    sa = gc.amd_solver(case)
    for _ in range(3):
        so.one_step()
        sa.one_step()
    for name in ("E", "H", "J"):
        got, ref = getattr(sa, name).to_numpy(), getattr(so, name)
        assert np.array_equal(got, ref), f"{name}: rel-L2 {gc.rel_l2(got, ref):.3e}"
