"""GPU: emsolve with host-side sources, the emsolve ABC face copies and the PML.

* Plane wave between two ABC z faces (periodic x/y): the device emsolve
  (wf_abc_low after each step) is bit-identical to the oracle's emsolve.
* The reference's PML test (tests/test_005_pml_boundaries.py: periodic x/y,
  PEC z-low, 30-layer PML z-high, FP32, plane wave of 7 periods): bit-identical
  to the oracle over the first steps, and over the full run the reflected
  H and E stay within the reference's 10% bound.
"""

import numpy as np
import pytest
from scipy.constants import c as c_light
from scipy.constants import mu_0

from oracle import fit_oracle as fo

pytestmark = pytest.mark.gpu


def _pair(x, y, z, bc, dtype, bg=(1.0, 1.0), n_pml=10):
    import wakis_amd as wa
    so = fo.OracleSolver(x, y, z, bc[0], bc[1], bg=list(bg), dtype=dtype, n_pml=n_pml)
    grid = wa.GridFIT3D(x=x, y=y, z=z, verbose=0)
    sa = wa.SolverFIT3D(grid, bc_low=bc[0], bc_high=bc[1], use_stl=False, bg=list(bg),
                        dtype=dtype, verbose=0, n_pml=n_pml)
    assert sa.dt == so.dt
    return so, sa


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_emsolve_abc_plane_wave_matches_oracle(dtype):
    import wakis_amd as wa
    print("emsolve abc", np.dtype(dtype).name, flush=True)
    x, y, z = fo.uniform_axes(-1.0, 1.0, -1.0, 1.0, 0.0, 1.0, 12, 10, 60)
    so, sa = _pair(x, y, z, (["periodic", "periodic", "abc"], ["periodic", "periodic", "abc"]),
                   dtype)
    f = 15 / (1.0 / c_light)
    po = fo.OraclePlaneWave(zs=5, f=f, nodes=2, amplitude=10.0, phase=np.pi / 2)
    pa = wa.PlaneWave(zs=5, f=f, nodes=2, amplitude=10.0, phase=np.pi / 2)
    fo.emsolve(so, 120, source=po)
    sa.emsolve(120, source=pa)
    assert so.activate_abc and sa.activate_abc
    for name in ("E", "H", "J"):
        assert np.array_equal(getattr(sa, name).to_numpy(), getattr(so, name)), name
    assert np.any(so.E != 0)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("bc", [
    (["abc", "abc", "abc"], ["abc", "abc", "abc"]),
    (["abc", "periodic", "pec"], ["pec", "periodic", "abc"]),
])
def test_emsolve_abc_faces_random_fields_match_oracle(dtype, bc):
    """ABC on x, y and z faces (edge and corner cells written by several faces in
    the reference's order), random initial E and H, no source."""
    print("emsolve abc faces", bc, np.dtype(dtype).name, flush=True)
    x, y, z = fo.uniform_axes(-1.0, 1.0, -1.0, 1.0, 0.0, 2.0, 11, 9, 14)
    so, sa = _pair(x, y, z, bc, dtype)
    rng = np.random.default_rng(11)
    so.E[:] = rng.standard_normal(so.E.shape).astype(dtype)
    so.H[:] = rng.standard_normal(so.H.shape).astype(dtype)
    sa.E.fromarray(so.E)
    sa.H.fromarray(so.H)
    fo.emsolve(so, 40)
    sa.emsolve(40)
    assert so.activate_abc and sa.activate_abc
    for name in ("E", "H", "J"):
        assert np.array_equal(getattr(sa, name).to_numpy(), getattr(so, name)), name


def _test005_setup(Nz=200):
    x, y, z = fo.uniform_axes(-1.0, 1.0, -1.0, 1.0, 0.0, 1.0, 20, 20, Nz)
    bc = (["periodic", "periodic", "pec"], ["periodic", "periodic", "pml"])
    return x, y, z, bc


def test_pml_user_loop_matches_oracle():
    """tests/test_005_pml_boundaries.py:24-77 set-up, first 200 steps."""
    import wakis_amd as wa
    print("pml loop", flush=True)
    x, y, z, bc = _test005_setup()
    so, sa = _pair(x, y, z, bc, np.float32, bg=(1.0, 1.0, 0.0), n_pml=30)
    f = 15 / ((z.max() - z.min()) / c_light)
    so.dt = 1 / f / 200
    sa.dt = 1 / f / 200
    kw = dict(xs=slice(0, 20), ys=slice(0, 20), zs=0, beta=1.0, amplitude=100.0, f=f, nodes=7,
              phase=np.pi / 2)
    po, pa = fo.OraclePlaneWave(**kw), wa.PlaneWave(**kw)
    for n in range(200):
        po.update(so, n * so.dt)
        pa.update(sa, n * sa.dt)
        so.one_step()
        sa.one_step()
    for name in ("E", "H", "J"):
        assert np.array_equal(getattr(sa, name).to_numpy(), getattr(so, name)), name


def test_pml_reflection_reference_bound():
    """The reference's test_005 run through the device engine: after the
    wave has entered the PML, |H| and |E|/(mu0 c) stay below 10 (10% of the
    amplitude 100)."""
    import wakis_amd as wa
    print("pml reflection (test_005)", flush=True)
    x, y, z, bc = _test005_setup()
    grid = wa.GridFIT3D(x=x, y=y, z=z, verbose=0)
    s = wa.SolverFIT3D(grid, bc_low=bc[0], bc_high=bc[1], use_stl=False, bg=[1.0, 1.0, 0.0],
                       dtype=np.float32, verbose=0, n_pml=30)
    f = 15 / ((s.z.max() - s.z.min()) / c_light)
    pw = wa.PlaneWave(xs=slice(0, 20), ys=slice(0, 20), zs=0, beta=1.0, amplitude=100.0, f=f,
                      nodes=7, phase=np.pi / 2)
    s.dt = 1 / f / 200
    z_pml = sum(s.dz[200 - 1 - k] for k in range(s.n_pml))
    Nt = int(1.0 * 2.0 * (1.0 - 0.0 - z_pml) / c_light / s.dt)
    for n in range(Nt):
        pw.update(s, n * s.dt)
        s.one_step()
    reflH = s.H.get_abs()[10, 10, :].max()
    reflE = s.E.get_abs()[10, 10, :].max() / (mu_0 * c_light)
    print(f"Nt={Nt} reflection H {float(reflH):.3f} E {float(reflE):.3f}", flush=True)
    assert reflH <= 10 and reflE <= 10


_SOURCES = {
    "wavepacket": (lambda m, dt: m.WavePacket(zs=3, sigmaz=0.05, sigmaxy=0.4, wavelength=0.2,
                                              amplitude=5.0),
                   lambda dt: fo.OracleWavePacket(zs=3, sigmaz=0.05, sigmaxy=0.4,
                                                  wavelength=0.2, amplitude=5.0)),
    "dipole_E": (lambda m, dt: m.Dipole(field="Ez", nodes=6, amplitude=2.0),
                 lambda dt: fo.OracleDipole(field="Ez", nodes=6, amplitude=2.0)),
    "dipole_H": (lambda m, dt: m.Dipole(field="H", component="x", xs=4, ys=5, zs=7, f=3e9),
                 lambda dt: fo.OracleDipole(field="H", component="x", xs=4, ys=5, zs=7, f=3e9)),
    "dipole_J": (lambda m, dt: m.Dipole(field="J", component="y", f=2e9, amplitude=1e-3),
                 lambda dt: fo.OracleDipole(field="J", component="y", f=2e9, amplitude=1e-3)),
    # L in metres of c t (the reference's default L = 50 dt mixes units and cuts the
    # pulse at once)
    "pulse_harris": (lambda m, dt: m.Pulse(field="Ez", shape="Harris", L=30 * dt * c_light),
                     lambda dt: fo.OraclePulse(field="Ez", shape="Harris",
                                               L=30 * dt * c_light)),
    "pulse_gauss": (lambda m, dt: m.Pulse(field="Ey", shape="Gaussian", L=40 * dt * c_light),
                    lambda dt: fo.OraclePulse(field="Ey", shape="Gaussian",
                                              L=40 * dt * c_light)),
    "pulse_rect": (lambda m, dt: m.Pulse(field="E", component="x", shape="Rectangular",
                                         L=20 * dt * c_light, delay=5 * dt * c_light),
                   lambda dt: fo.OraclePulse(field="E", component="x", shape="Rectangular",
                                             L=20 * dt * c_light, delay=5 * dt * c_light)),
}


@pytest.mark.parametrize("name", sorted(_SOURCES))
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_emsolve_sources_match_oracle(name, dtype):
    """WavePacket, Dipole (E, H, J) and Pulse (three profiles) through emsolve
    (sources.py:285-762): bit-identical to the oracle's emsolve with the same
    source."""
    import wakis_amd as wa
    print("emsolve source", name, np.dtype(dtype).name, flush=True)
    x, y, z = fo.uniform_axes(-1.0, 1.0, -1.0, 1.0, 0.0, 2.0, 12, 12, 24)
    so, sa = _pair(x, y, z, (["pec", "pec", "pec"], ["pec", "pec", "pec"]), dtype)
    make_a, make_o = _SOURCES[name]
    fo.emsolve(so, 60, source=make_o(float(so.dt)))
    sa.emsolve(60, source=make_a(wa, float(sa.dt)))
    assert np.any(so.E != 0) or np.any(so.H != 0)
    for fld in ("E", "H", "J"):
        assert np.array_equal(getattr(sa, fld).to_numpy(), getattr(so, fld)), fld
