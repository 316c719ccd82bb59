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


def test_monitor_in_wakesolve_bitwise_vs_numpy():
    """wakesolve with a monitor (per-step updates) against the reference's
    arithmetic applied to the same per-step E snapshots."""
    import wakis_amd as wa
    from wakis_amd import FieldMonitor
    freqs = [0.5e9, 1.0e9, 3.0e9]
    grid = wa.GridFIT3D(0.0, 1.0, 0.0, 1.0, 0.0, 1.0, 4, 4, 4, verbose=0)
    wake = wa.WakeSolver(q=1e-9, sigmaz=1e-2, beta=1.0, save=False, verbose=0)
    wake.solve = lambda *a, **k: None
    s = wa.SolverFIT3D(grid, wake, bc_low=["pec"] * 3, bc_high=["pec"] * 3, use_stl=False,
                       bg="pec", dtype=np.float32, verbose=0)
    m = FieldMonitor(freqs)
    snaps = []

    def cb(solver, t):
        snaps.append(solver.E.to_numpy().copy())

    s.wakesolve(wakelength=0.1, add_space=0, use_field_monitor=True, field_monitor=m,
                callback=cb)
    assert m.time_index == s.Nt > 0 and m.shape == (4, 4, 4)
    ref = {c: np.zeros((len(freqs), 4, 4, 4), dtype=np.complex128) for c in "xyz"}
    for n, E in enumerate(snaps):
        t = n * m.dt
        for i, fr in enumerate(freqs):
            ph = np.exp(-2j * np.pi * fr * t)
            for ci, c in enumerate("xyz"):
                Ec = np.reshape(E[ci * 64:(ci + 1) * 64], (4, 4, 4), order="F")
                ref[c][i] += Ec * ph
    comps = m.get_components()
    for c in "xyz":
        assert np.array_equal(comps["E" + c].cpu().numpy(), ref[c]), c
