"""GPU: the named workloads against the FP64 oracle.

* 48x48x128 cavity/pipe wakefield run (2062 steps, 0.5 m wake, Z up to
  5 GHz): the FP64 device run is bit-identical to the oracle's committed
  outputs (tests/golden/cavity_48x48x128_fp64_oracle.npz, made by
  tests/golden/make_cavity_48x48x128_golden.py); the FP32 device run is within 1e-4
  relative L2 of it on WP_z.
* Lossy pillbox with PML (conductive wall shell, beta = 0.8, tau-clamped dt),
  scaled down to 40x40x80: FP32 and FP64 device runs bit-identical to the
  oracle run in the same process.
"""

import math
from pathlib import Path

import numpy as np
import pytest

from oracle import cases
from oracle import fit_oracle as fo
from tests import gpu_cases as gc

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden" / "cavity_48x48x128_fp64_oracle.npz"


def _cavity_48x48x128(dtype):
    import wakis_amd as wa
    x, y, z, mask = cases.cavity_pipe_geometry(48, 48, 128)
    grid = wa.GridFIT3D(x=x, y=y, z=z, stl_solids={"cav": mask},
                        stl_materials={"cav": "vacuum"}, verbose=0)
    wake = wa.WakeSolver(q=1e-9, sigmaz=10e-3, beta=1.0, skip_cells=10, fmax=5e9,
                         save=False, verbose=0)
    s = wa.SolverFIT3D(grid, wake=wake, bc_low=["pec", "pec", "abc"],
                       bc_high=["pec", "pec", "abc"], use_stl=True, bg="pec", dtype=dtype,
                       verbose=0)
    s.wakesolve(wakelength=0.5)
    return s, wake


def _digest(F, seed, stride=97):
    w = np.random.default_rng(seed).standard_normal(F.size)
    return F[::stride], math.fsum((F * w).tolist())


def test_cavity_48x48x128_fp64_bit_identical_to_oracle():
    print("cavity 48x48x128 fp64, 2062 steps", flush=True)
    g = np.load(GOLD, allow_pickle=False)
    s, wake = _cavity_48x48x128(np.float64)
    assert s.Nt == int(g["Nt"]) == 2062
    assert s.dt == float(g["dt"])
    assert s.last_path == "fused"
    assert np.array_equal(wake.s, g["s"])
    assert np.array_equal(wake.WP_3d, g["WP_3d"])
    assert np.array_equal(wake.WPx, g["WPx"]) and np.array_equal(wake.WPy, g["WPy"])
    assert np.array_equal(wake.f, g["f"])
    for k in ("Z", "Zx", "Zy"):
        ref = g[k]
        assert np.max(np.abs(getattr(wake, k) - ref)) <= 1e-12 * np.max(np.abs(ref)), k
    for name, seed in (("E", 1), ("H", 2), ("J", 3)):
        sample, dot = _digest(getattr(s, name).to_numpy(), seed)
        assert np.array_equal(sample, g[name + "_sample"]), name
        assert dot == float(g[name + "_dot"]), name


def test_cavity_48x48x128_fp32_wake_within_1e4_of_fp64_oracle():
    print("cavity 48x48x128 fp32", flush=True)
    g = np.load(GOLD, allow_pickle=False)
    s, wake = _cavity_48x48x128(np.float32)
    # the FP32 run's s and f axes come from its FP32 dt: compare on the common length
    n, m = min(len(wake.s), len(g["s"])), min(len(wake.f), len(g["f"]))
    assert np.max(np.abs(wake.s[:n] - g["s"][:n])) <= 1e-6 * np.max(np.abs(g["s"]))
    err_wp = gc.rel_l2(wake.WP[:n], g["WP_3d"][1, 1][:n])
    err_z = gc.rel_l2(np.abs(wake.Z[:m]), np.abs(g["Z"][:m]))
    print(f"FP32 vs FP64 oracle: WP_z rel-L2 {err_wp:.3e}, |Z_z| rel-L2 {err_z:.3e}", flush=True)
    assert err_wp <= 1e-4
    assert err_z <= 1e-4


def _pillbox_pair(dtype, Nx=40, Ny=40, Nz=80, steps=150, sibc=False):
    import wakis_amd as wa
    x, y, z, vac, wall = cases.pillbox_geometry(Nx, Ny, Nz)
    lossy = [1.0, 1.0, 1e4]
    bc = (["pec", "pec", "pml"], ["pec", "pec", "pml"])
    so = fo.OracleSolver(x, y, z, *bc, bg="pec", solids=[(wall, lossy), (vac, "vacuum")],
                         dtype=dtype, n_pml=10, use_sibc=sibc)
    grid = wa.GridFIT3D(x=x, y=y, z=z, stl_solids={"wall": wall, "vac": vac},
                        stl_materials={"wall": lossy, "vac": "vacuum"}, verbose=0)
    # use_sibc=False keeps the volume conductivity of the wall; True (the
    # reference's default) replaces it by a surface-impedance layer
    sa = wa.SolverFIT3D(grid, bc_low=bc[0], bc_high=bc[1], use_stl=True, bg="pec",
                        dtype=dtype, verbose=0, n_pml=10, use_sibc=sibc)
    assert sa.dt == so.dt and sa.use_conductivity and so.use_conductivity
    if sibc:
        assert np.array_equal(grid.grid["wall"], fo.sibc_surface(wall))
        for name in ("ieps", "sigma"):
            assert np.array_equal(getattr(sa, name).to_numpy(), getattr(so, name)), name
    bo = fo.OracleBeam(q=1e-9, sigmaz=15e-3, beta=0.8)
    ba = wa.Beam(q=1e-9, sigmaz=15e-3, beta=0.8)
    # start with the bunch centre just below the domain so the short run sees it
    t0 = 0.9 * bo.ti
    for n in range(steps):
        t = t0 + n * so.dt
        bo.update(so, t)
        ba.update(sa, t)
        so.one_step()
        sa.one_step()
    return so, sa


@pytest.mark.parametrize("sibc", [False, True])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_lossy_pillbox_pml_bit_identical(dtype, sibc):
    print("pillbox", np.dtype(dtype).name, "sibc" if sibc else "volume", flush=True)
    so, sa = _pillbox_pair(dtype, sibc=sibc)
    assert sa.last_path == "fused"
    assert np.any(so.E != 0)
    for name in ("E", "H", "J"):
        got, ref = getattr(sa, name).to_numpy(), getattr(so, name)
        assert np.array_equal(got, ref), f"{name}: rel-L2 {gc.rel_l2(got, ref):.3e}"
