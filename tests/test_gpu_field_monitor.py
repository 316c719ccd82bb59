"""GPU: FieldMonitor on the device, with the reference's test_012 cases
(tests/test_012_field_monitors.py) and a bitwise check against a numpy
restatement of field_monitors.py:59-102 on the same snapshots."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _const_field(Nx=2, Ny=2, Nz=2, values=(1.0, 2.0, 3.0)):
    import wakis_amd as wa
    f = wa.Field(Nx, Ny, Nz, dtype=float)
    for c, v in zip("xyz", values):
        f.from_matrix(np.full((Nx, Ny, Nz), v), c)
    return f


def test_initial_state():
    from wakis_amd import FieldMonitor
    m = FieldMonitor([1.0e9, 2.0e9])
    assert isinstance(m.frequencies, np.ndarray) and np.allclose(m.frequencies, [1e9, 2e9])
    assert m.time_index == 0 and m.dt is None and m.shape is None and m.xp is None
    assert m.get_components() == {"Ex": None, "Ey": None, "Ez": None}


def test_single_update_zero_frequency():
    from wakis_amd import FieldMonitor
    f = _const_field()
    m = FieldMonitor([0.0])
    m.update(f, 1.0e-9)
    assert m.dt == 1.0e-9 and m.shape == (2, 2, 2) and m.time_index == 1 and m.xp is f.xp
    comps = m.get_components()
    for c in "xyz":
        a = comps["E" + c]
        assert tuple(a.shape) == (1, 2, 2, 2) and a.is_complex()
        assert np.allclose(a[0].cpu().numpy(), f.to_matrix(c).cpu().numpy())


def test_linear_accumulation_and_geometric_series():
    from wakis_amd import FieldMonitor
    f = _const_field(values=(1.0, 0.0, 0.0))
    m = FieldMonitor([0.0])
    for _ in range(5):
        m.update(f, 0.1)
    comps = m.get_components()
    assert np.allclose(comps["Ex"][0].cpu().numpy(), 5.0)
    assert np.allclose(comps["Ey"].cpu().numpy(), 0.0)
    g = _const_field(Nx=1, Ny=1, Nz=1, values=(1.0, 0.0, 0.0))
    m = FieldMonitor([0.5])  # phase (-1)^n
    for _ in range(4):
        m.update(g, 1.0)
    assert np.allclose(m.get_components()["Ex"][0].cpu().numpy(), 0.0, atol=1e-12)


@pytest.mark.parametrize("dtype,dt", [(np.float32, np.float32(1.1e-12)), (np.float32, 1.1e-12),
                                      (np.float64, np.float64(1.1e-12)),
                                      (np.float64, np.float32(1.1e-12))])
def test_update_bitwise_vs_numpy_promotions(dtype, dt):
    """FieldMonitor.update on random fields against the reference's numpy
    arithmetic (field_monitors.py:90-100) with its promotions: the phase is
    complex64 for an FP32 dt, and E * phase is complex64 for FP32 E with a
    complex64 phase, complex128 otherwise; the accumulators are complex128."""
    import wakis_amd as wa
    from wakis_amd import FieldMonitor
    print("monitor update", np.dtype(dtype).name, type(dt).__name__, flush=True)
    rng = np.random.default_rng(3)
    Nx, Ny, Nz = 6, 5, 7
    f = wa.Field(Nx, Ny, Nz, dtype=dtype)
    m = FieldMonitor([0.9e9, 2.5e9])
    ref = {c: np.zeros((2, Nx, Ny, Nz), np.complex128) for c in "xyz"}
    for n in range(6):
        for c in "xyz":
            f.from_matrix(rng.standard_normal((Nx, Ny, Nz)).astype(dtype), c)
        m.update(f, dt)
        t = n * m.dt
        for i, fr in enumerate(m.frequencies):
            ph = np.exp(-2j * np.pi * fr * t)
            for c in "xyz":
                ref[c][i] += f.to_matrix(c).cpu().numpy() * ph
    comps = m.get_components()
    for c in "xyz":
        assert np.array_equal(comps["E" + c].cpu().numpy(), ref[c]), c


def _monitor_run(dtype, per_step, path="auto", snaps=None):
    """A 16x16x32 cavity/pipe wakesolve with a 3-frequency monitor; per_step: a
    callback forces one step per call and FieldMonitor.update after each."""
    import wakis_amd as wa
    from oracle import cases
    from wakis_amd import FieldMonitor
    x, y, z, mask = cases.cavity_pipe_geometry(16, 16, 32)
    grid = wa.GridFIT3D(x=x, y=y, z=z, stl_solids={"cav": mask},
                        stl_materials={"cav": "vacuum"}, verbose=0)
    wake = wa.WakeSolver(q=1e-9, sigmaz=10e-3, beta=1.0, skip_cells=2, save=False, verbose=0)
    s = wa.SolverFIT3D(grid, wake, bc_low=["pec", "pec", "abc"], bc_high=["pec", "pec", "abc"],
                       use_stl=True, bg="pec", dtype=dtype, verbose=0, path=path)
    m = FieldMonitor([0.7e9, 2.1e9, 4.0e9])
    cb = None
    if per_step:
        def cb(solver, t):
            if snaps is not None:
                snaps.append(solver.E.to_numpy().copy())
    s.wakesolve(wakelength=0.03, use_field_monitor=True, field_monitor=m, callback=cb)
    return s, m


def test_monitor_in_wakesolve_bitwise_vs_numpy():
    """wakesolve with a monitor (per-step updates) against the reference's
    arithmetic (field_monitors.py:95-100) applied to the same per-step E
    snapshots."""
    snaps = []
    s, m = _monitor_run(np.float32, True, snaps=snaps)
    assert m.time_index == s.Nt == len(snaps) > 0 and m.shape == (16, 16, 32)
    N = 16 * 16 * 32
    ref = {c: np.zeros((3, 16, 16, 32), dtype=np.complex128) for c in "xyz"}
    for n, E in enumerate(snaps):
        t = n * m.dt
        for i, fr in enumerate(m.frequencies):
            ph = np.exp(-2j * np.pi * fr * t)
            for ci, c in enumerate("xyz"):
                Ec = np.reshape(E[ci * N:(ci + 1) * N], (16, 16, 32), order="F")
                ref[c][i] += Ec * ph
    comps = m.get_components()
    assert np.any(ref["z"] != 0)
    for c in "xyz":
        assert np.array_equal(comps["E" + c].cpu().numpy(), ref[c]), c


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("path", ["auto", "general"])
def test_monitor_accumulated_inside_the_step_call(dtype, path):
    """Without a callback wakesolve runs the whole loop in one step call that
    accumulates the monitor after every step from a device table of phases
    (no per-step host work): bitwise the per-step updates' accumulators."""
    import torch
    print("monitor in the step call", np.dtype(dtype).name, path, flush=True)
    s1, m1 = _monitor_run(dtype, True, path)
    s2, m2 = _monitor_run(dtype, False, path)
    assert m1.time_index == m2.time_index == s1.Nt == s2.Nt
    assert s2.last_path == ("fused" if path == "auto" else "general")
    assert torch.equal(s1.E.array, s2.E.array)
    for c in "xyz":
        a, b = m1.get_components()["E" + c], m2.get_components()["E" + c]
        assert torch.equal(a, b), c
    assert bool((m2.get_components()["Ez"] != 0).any())
