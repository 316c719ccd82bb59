"""GPU: wakesolve at the lossy-pillbox settings and the remaining branches of
the wake potential, against the oracle.

* wakesolve with beta = 0.8 (the Beam and the WakeSolver use different
  injection times, sources.py:69-73 and wakeSolver.py:164-172), a source or
  test line 2 mm off axis, a conductive wall, 10 PML layers on z, with and
  without the surface-impedance substitution: Nt, s, WP_3d, WP_x, WP_y and
  the fields bit-identical, Z, Z_x, Z_y within 1e-12 of the largest bin.
* the counter-moving integral (wakeSolver.py:503-517), calc_long_WP on a 3x3
  window (the off-centre line of wakeSolver.py:385-388) and calc_long_WP on a
  given Ezt array, on random captures: bit-identical.
"""

import numpy as np
import pytest
import torch

from oracle import cases
from oracle import fit_oracle as fo
from oracle import wake_oracle as wo
from tests import gpu_cases as gc

pytestmark = pytest.mark.gpu

Q, SIGMAZ, BETA = 1e-9, 15e-3, 0.8
# scaled-down pillbox: cylinder r = 20 mm, |z| < 15 mm, pipe r = 5 mm, 3-cell
# wall, 50 x 50 x 90 mm on 25 x 25 x 48 cells (2 mm across, a cell centred on
# the axis: a 2 mm offset moves the line by one cell)
PILLBOX = dict(r_cav=20e-3, half_len=15e-3, r_pipe=5e-3, wall_cells=3, half_x=25e-3,
               half_z=45e-3)


def _say(*a):
    print(*a, flush=True)


def _pillbox_wake_pair(dtype, sibc, sigma_wall, xsource, xtest, wakelength=0.01, skip=3,
                       dims=(25, 25, 48)):
    import wakis_amd as wa
    x, y, z, vac, wall = cases.pillbox_geometry(*dims, **PILLBOX)
    lossy = [1.0, 1.0, sigma_wall]
    bc = (["pec", "pec", "pml"], ["pec", "pec", "pml"])
    # a solver built with a WakeSolver and the default fmax takes the wake's fmax
    # = v/(3 sigma_z) for the surface impedance (solverFIT3D.py:177-178)
    fmax = BETA * 299792458.0 / (3.0 * SIGMAZ)
    so = fo.OracleSolver(x, y, z, *bc, bg="pec", solids=[(wall, lossy), (vac, "vacuum")],
                         dtype=dtype, n_pml=10, use_sibc=sibc, fmax=fmax)
    ti = wo.wake_ti(SIGMAZ, BETA)
    cap = fo.wakesolve_fields(so, wakelength, Q, SIGMAZ, BETA, xsource, 0.0, xtest, 0.0, ti)
    ref = wo.wake_solve(cap, Q, SIGMAZ, BETA, wakelength, skip_cells=skip)
    grid = wa.GridFIT3D(x=x, y=y, z=z, stl_solids={"wall": wall, "vac": vac},
                        stl_materials={"wall": lossy, "vac": "vacuum"}, verbose=0)
    wake = wa.WakeSolver(q=Q, sigmaz=SIGMAZ, beta=BETA, xsource=xsource, xtest=xtest,
                         skip_cells=skip, save=False, verbose=0)
    sa = wa.SolverFIT3D(grid, wake=wake, bc_low=bc[0], bc_high=bc[1], use_stl=True, bg="pec",
                        dtype=dtype, verbose=0, n_pml=10, use_sibc=sibc)
    assert sa.fmax == fmax == wake.fmax
    for name in ("ieps", "imu", "sigma"):
        assert np.array_equal(getattr(sa, name).to_numpy(), getattr(so, name)), name
    sa.wakesolve(wakelength=wakelength)
    return so, sa, cap, ref, wake


@pytest.mark.parametrize("dtype,sibc,offset", [
    (np.float64, True, "test"), (np.float64, True, "source"), (np.float32, True, "source"),
    (np.float64, False, "test"), (np.float64, False, "source")])
def test_lossy_pillbox_wakesolve_beta08_offset_matches_oracle(dtype, sibc, offset):
    """Config 3's settings, scaled down. With use_sibc the wall keeps its
    surface cells (sigma = eps = 1/Z_s, sigma 1e4 S/m); without, the wall's
    sigma is 10 S/m so that its relaxation time equals the PML's and the
    reference's tau clamp leaves dt at the PML's eps0/10 (a 1e4 S/m volume wall
    clamps dt to 8.9e-16 s: millions of steps)."""
    xsource, xtest = (2e-3, 0.0) if offset == "source" else (0.0, 2e-3)
    sigma_wall = 1e4 if sibc else 10.0
    _say("pillbox wakesolve beta 0.8", np.dtype(dtype).name, "sibc" if sibc else "volume",
         offset)
    so, sa, cap, ref, wake = _pillbox_wake_pair(dtype, sibc, sigma_wall, xsource, xtest)
    # the two injection times (Finding: Beam without sqrt(beta), WakeSolver with it)
    v = BETA * 299792458.0
    assert sa._wake_beam.ti == 8.548921333333334 * SIGMAZ / v
    assert wake.ti == ref["ti"] == 8.548921333333334 * SIGMAZ / (np.sqrt(BETA) * v)
    assert sa.dt == so.dt and sa.Nt == cap["Nt"]
    assert (sa.ixt, sa.iyt) == (cap["ixt"], cap["iyt"])
    assert (int(sa._wake_beam.ixs), int(sa._wake_beam.iys)) == (cap["ixs"], cap["iys"])
    assert abs(int(cap["ixs"]) - int(cap["ixt"])) == 1  # the offset moved one of the lines
    for name in ("E", "H", "J"):
        got, r = getattr(sa, name).to_numpy(), getattr(so, name)
        assert np.array_equal(got, r), f"{name}: rel-L2 {gc.rel_l2(got, r):.3e}"
    assert np.array_equal(wake.s, ref["s"])
    assert np.array_equal(wake.WP_3d, ref["WP_3d"]), "WP_3d not bit-identical"
    assert np.array_equal(wake.WPx, ref["WPx"]) and np.array_equal(wake.WPy, ref["WPy"])
    assert np.any(wake.WPx != 0)
    assert np.array_equal(wake.f, ref["f"])
    for k in ("Z", "Zx", "Zy"):
        r = ref[k]
        assert np.max(np.abs(getattr(wake, k) - r)) <= 1e-12 * np.max(np.abs(r)), k


def _random_capture(Nz, nsamples, nt_axis, dtype, seed, t0=0.0):
    """A random 3x3 Ez capture on a non-uniform z axis (so dz[k] and dz[nz-1-k]
    differ) and its time axis of nt_axis samples starting at t0."""
    rng = np.random.default_rng(seed)
    zn = np.cumsum(np.r_[-0.04, rng.uniform(0.8e-3, 1.2e-3, Nz)])
    zf = zn[:-1] + np.diff(zn) / 2
    dt = 2e-12
    t = t0 + np.arange(nt_axis) * dt
    Ez = rng.standard_normal((nsamples, 3, 3, Nz)).astype(dtype)
    return dict(Ez=Ez, t=t, zf=zf, dz=np.diff(zn), xf=np.array([-1e-3, 0.0, 1.1e-3]),
                yf=np.array([-0.9e-3, 0.0, 1e-3]))


def _device_wake(cap, skip, counter_moving=False, sigmaz=5e-3, beta=1.0, wakelength=0.03):
    import wakis_amd as wa
    wake = wa.WakeSolver(q=Q, sigmaz=sigmaz, beta=beta, skip_cells=skip, save=False,
                         verbose=0, counter_moving=counter_moving)
    wake.wakelength = wakelength
    Ez = cap["Ez"]
    # device capture layout [line (i-major)][sample][z]
    dev = torch.as_tensor(np.ascontiguousarray(Ez.reshape(Ez.shape[0], 9, -1).transpose(1, 0, 2)),
                          device="cuda")
    wake._set_capture(dev, cap["t"], cap["xf"], cap["yf"], cap["zf"], None, None, cap["dz"],
                      Ez.dtype, Ez.shape[0])
    return wake


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("skip", [0, 3])
@pytest.mark.parametrize("beta", [1.0, 0.7])
def test_counter_moving_wp_matches_oracle(dtype, skip, beta):
    _say("counter-moving WP", np.dtype(dtype).name, skip, beta)
    cap = _random_capture(40, 300, 300, dtype, seed=3 + skip)
    wake = _device_wake(cap, skip, counter_moving=True, beta=beta)
    wake.calc_long_WP_3d()
    ti = wo.wake_ti(5e-3, beta)
    s, WP, WP3 = wo.long_WP_3d(cap, Q, beta * 299792458.0, ti, 0.03, skip_cells=skip,
                               counter_moving=True)
    assert np.array_equal(wake.s, s)
    assert np.any(WP3 != 0)
    assert np.array_equal(wake.WP_3d, WP3)
    # the forward integral of the same capture differs
    s2, WPf, _ = wo.long_WP_3d(cap, Q, beta * 299792458.0, ti, 0.03, skip_cells=skip)
    assert not np.array_equal(WPf, WP)


@pytest.mark.parametrize("nsamples,t0", [(300, 0.0), (280, 0.0), (300, 3e-11)])
def test_long_wp_3d_time_index_wrap(nsamples, t0):
    """Fewer captured samples than time-axis points, and a time axis that does
    not start at 0 (negative time indices read from the end of the reference's
    Ezt of len(t) columns): bit-identical."""
    _say("WP_3d wrap", nsamples, t0)
    cap = _random_capture(32, nsamples, 300, np.float64, seed=9, t0=t0)
    wake = _device_wake(cap, 2)
    wake.calc_long_WP_3d()
    s, WP, WP3 = wo.long_WP_3d(cap, Q, 299792458.0, wo.wake_ti(5e-3, 1.0), 0.03, skip_cells=2)
    assert np.array_equal(wake.WP_3d, WP3)


@pytest.mark.parametrize("skip", [0, 2])
def test_long_wp_on_window_reads_the_reference_line(skip):
    """calc_long_WP on a 3x3 capture integrates line [2, 2] (shape//2 + 1 on
    both axes, wakeSolver.py:385-388) with skip_cells slicing z."""
    _say("calc_long_WP on a window", skip)
    cap = _random_capture(36, 250, 250, np.float64, seed=5)
    wake = _device_wake(cap, skip)
    wake.calc_long_WP()
    s, WP = wo.long_WP_1d(cap, Q, 299792458.0, wo.wake_ti(5e-3, 1.0), 0.03, skip_cells=skip)
    assert np.array_equal(wake.s, s)
    assert np.array_equal(wake.WP, WP)
    centre = dict(cap, Ez=cap["Ez"][:, 1, 1, :])
    _, WPc = wo.long_WP_1d(centre, Q, 299792458.0, wo.wake_ti(5e-3, 1.0), 0.03)
    assert skip != 0 or not np.array_equal(WP, WPc)  # not the centre line


@pytest.mark.parametrize("ncols,t0", [(250, 0.0), (200, 0.0), (250, 4e-11)])
@pytest.mark.parametrize("skip", [0, 2])
def test_long_wp_from_given_ezt(ncols, t0, skip):
    """calc_long_WP(Ezt=...) integrates the given (nz, nt) samples
    (wakeSolver.py:395-408); bit-identical to the oracle, including a time
    axis longer than Ezt and negative time indices."""
    import wakis_amd as wa
    _say("calc_long_WP(Ezt)", ncols, t0, skip)
    cap = _random_capture(30, 1, 250, np.float64, seed=7, t0=t0)
    nz = 30 - 2 * skip
    Ezt = np.random.default_rng(8).standard_normal((nz, ncols))
    wake = wa.WakeSolver(q=Q, sigmaz=5e-3, beta=1.0, skip_cells=skip, save=False, verbose=0)
    wake.calc_long_WP(Ezt=Ezt, t=cap["t"], zf=cap["zf"], dz=cap["dz"], wakelength=0.03)
    s, WP = wo.long_WP_from_Ezt(Ezt, cap["t"], cap["zf"], cap["dz"], Q, 299792458.0,
                                wo.wake_ti(5e-3, 1.0), 0.03, skip_cells=skip,
                                nt=min(250, ncols))
    assert np.array_equal(wake.s, s)
    assert np.any(WP != 0) and np.array_equal(wake.WP, WP)
    assert wake.Ezt is not None
    with pytest.raises(ValueError):
        wake.calc_long_WP(Ezt=Ezt[:-1])


def test_counter_moving_wakesolve_matches_oracle():
    """A wakefield run whose WakeSolver integrates counter-moving: the whole
    post-processing (WP_3d, WP_x/y, Z) against the oracle."""
    import wakis_amd as wa
    _say("wakesolve counter-moving 16x16x32")
    x, y, z, mask = cases.cavity_pipe_geometry(16, 16, 32)
    so = fo.OracleSolver(x, y, z, ["pec", "pec", "abc"], ["pec", "pec", "abc"], bg="pec",
                         solids=[(mask, "vacuum")], dtype=np.float64)
    sigmaz, wl = 10e-3, 0.04
    ti = wo.wake_ti(sigmaz, 1.0)
    cap = fo.wakesolve_fields(so, wl, Q, sigmaz, 1.0, 0.0, 0.0, 0.0, 0.0, ti)
    ref = wo.wake_solve(cap, Q, sigmaz, 1.0, wl, skip_cells=2, counter_moving=True)
    grid = wa.GridFIT3D(x=x, y=y, z=z, stl_solids={"cav": mask},
                        stl_materials={"cav": "vacuum"}, verbose=0)
    wake = wa.WakeSolver(q=Q, sigmaz=sigmaz, beta=1.0, skip_cells=2, save=False, verbose=0,
                         counter_moving=True)
    sa = wa.SolverFIT3D(grid, wake=wake, bc_low=["pec", "pec", "abc"],
                        bc_high=["pec", "pec", "abc"], use_stl=True, bg="pec",
                        dtype=np.float64, verbose=0)
    sa.wakesolve(wakelength=wl)
    assert np.array_equal(wake.WP_3d, ref["WP_3d"])
    assert np.array_equal(wake.WPx, ref["WPx"]) and np.array_equal(wake.WPy, ref["WPy"])
    assert np.max(np.abs(wake.Z - ref["Z"])) <= 1e-12 * np.max(np.abs(ref["Z"]))
