"""Failure handling of a wakefield run on the device: checkpoint / resume and the
non-finite check.

A run is killed from the host (a callback raises after a number of steps),
rebuilt, and resumed from its last checkpoint; WP, Z and the final fields must
be bitwise those of the uninterrupted run. A NaN written into a field on the
host must trip the device's non-finite count. The reference has neither
(its manual save_state/load_state, solverFIT3D.py:886-1013, does not carry the
wake capture).
"""

import numpy as np
import pytest

from oracle import cases

pytestmark = pytest.mark.gpu


class _Kill(Exception):
    pass


def _solver(dtype, conductive):
    import wakis_amd as wa
    x, y, z, mask = cases.cavity_pipe_geometry(20, 20, 40)
    solids, mats = {"cav": mask}, {"cav": "vacuum"}
    if conductive:
        block = np.zeros_like(mask)
        block[7:13, 7:13, 16:24] = True
        solids["block"], mats["block"] = block & mask, [2.0, 1.0, 1.0]
    wake = wa.WakeSolver(q=1e-9, sigmaz=10e-3, beta=1.0, skip_cells=2, save=False, verbose=0)
    grid = wa.GridFIT3D(x=x, y=y, z=z, stl_solids=solids, stl_materials=mats, verbose=0)
    return wa.SolverFIT3D(grid, wake=wake, bc_low=["pec", "pec", "abc"],
                          bc_high=["pec", "pec", "abc"], use_stl=True, bg="pec",
                          dtype=dtype, verbose=0)


def _result(s):
    w = s.wake
    return dict(E=s.E.to_numpy(), H=s.H.to_numpy(), J=s.J.to_numpy(), WP_3d=w.WP_3d,
                WP=w.WP, WPx=w.WPx, WPy=w.WPy, Z=w.Z, Zx=w.Zx, Zy=w.Zy, s=w.s,
                nstep=s.nstep)


def _assert_same(a, b):
    for k in a:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("conductive", [False, True])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_killed_run_resumes_bitwise(tmp_path, dtype, conductive):
    print("resume", np.dtype(dtype).name, "conductive" if conductive else "lossless",
          flush=True)
    wl = 0.05
    a = _solver(dtype, conductive)
    a.wakesolve(wakelength=wl)
    ref = _result(a)
    assert a.fields_finite and a.use_conductivity == conductive

    path = str(tmp_path / "run.npz")
    calls = []

    def kill(solver, t):
        calls.append(t)
        if len(calls) == 137:
            raise _Kill()

    b = _solver(dtype, conductive)
    with pytest.raises(_Kill):
        b.wakesolve(wakelength=wl, checkpoint=path, checkpoint_every=50, callback=kill)
    with np.load(path) as st:
        assert int(st["wake_n"]) == 100
    c = _solver(dtype, conductive)
    c.wakesolve(wakelength=wl, resume=path)
    _assert_same(ref, _result(c))


def test_checkpointed_chunked_run_resumes_bitwise(tmp_path):
    print("resume chunked", flush=True)
    path = str(tmp_path / "run.npz")
    a = _solver(np.float32, True)
    a.wakesolve(wakelength=0.05, checkpoint=path, checkpoint_every=120)
    ref = _result(a)
    with np.load(path) as st:
        n = int(st["wake_n"])
        assert n % 120 == 0 and a.Nt - 120 <= n < a.Nt
        # the file also carries save_state's keys
        for k in ("E", "H", "J"):
            assert st[k].shape == ref[k].shape
    c = _solver(np.float32, True)
    c.wakesolve(wakelength=0.05, resume=path)
    _assert_same(ref, _result(c))


def test_resume_rejects_another_run(tmp_path):
    print("resume mismatch", flush=True)
    path = str(tmp_path / "run.npz")
    a = _solver(np.float64, False)
    a.wakesolve(wakelength=0.05, checkpoint=path, checkpoint_every=100)
    for dtype, wl in ((np.float64, 0.06), (np.float32, 0.05)):
        c = _solver(dtype, False)
        with pytest.raises(ValueError):
            c.wakesolve(wakelength=wl, resume=path)
    with pytest.raises(ValueError):
        _solver(np.float64, False).wakesolve(wakelength=0.05, checkpoint_every=10)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_nonfinite_values_trip_the_device_count(dtype):
    print("nonfinite", np.dtype(dtype).name, flush=True)
    s = _solver(dtype, False)
    for _ in range(3):
        s.one_step()
    assert s.count_nonfinite() == {"E": 0, "H": 0, "J": 0} and s.check_finite()
    s.E[10, 10, 20, "z"] = np.nan
    s.H[3, 4, 5, "x"] = np.inf
    s.J[1, 2, 3, "y"] = -np.inf
    assert s.count_nonfinite() == {"E": 1, "H": 1, "J": 1}
    assert not s.check_finite(("E",)) and s.check_finite(())
    s.one_step()
    got = s.count_nonfinite(("E", "H"))
    assert got["E"] > 1 and got["H"] > 1
    assert got["E"] == int(np.count_nonzero(~np.isfinite(s.E.to_numpy())))
    assert got["H"] == int(np.count_nonzero(~np.isfinite(s.H.to_numpy())))
