"""GPU: the energy diagnostic and the loss factor.

The FIT energy W^n = 1/2 sum eps*tA*L*(E^n)^2 + 1/2 sum mu*A*tL*H^(n-1/2)*H^(n+1/2)
is conserved by the lossless leapfrog up to rounding and decreases with
conductivity; the device reduction must agree with a numpy evaluation of the
same sum on the oracle's (bit-identical) fields. The loss factor
k = -sum WP*lambda*ds must agree with numpy and, through Parseval, with the
impedance: k = (df / (v^2 1e12)) * [Re Z_0 |lambda_0|^2 + 2 sum_k>0 Re Z_k |lambda_k|^2].
"""

import numpy as np
import pytest

from oracle import cases
from tests import gpu_cases as gc
from tests.test_oracle_core import _energy

pytestmark = pytest.mark.gpu

PEC_ALL = (["pec", "pec", "pec"], ["pec", "pec", "pec"])


def _box_case(conductive, seed=7):
    case = gc.random_case(PEC_ALL, np.float64, conductive, seed=seed, dims=(16, 12, 20))
    case["fields"][2] = np.zeros_like(case["fields"][2])
    return case


@pytest.mark.parametrize("path", ["fused", "general"])
def test_energy_conserved_and_matches_numpy(path):
    print("energy lossless", path, flush=True)
    case = _box_case(False)
    so = gc.oracle_solver(case)
    sa = gc.amd_solver(case, path=path)
    sa.keep_previous_fields = True
    # frozen E on PEC-material edges would act as a source: start those at zero
    live = (so.ieps > 0) & (so.itA > 0)
    so.E[:] = np.where(live, so.E, 0.0)
    sa.E.fromarray(so.E)
    W = []
    for n in range(120):
        E0, H0 = so.E.copy(), so.H.copy()
        so.one_step()
        sa.one_step()
        w = sa.energy()
        W.append(w)
        if n in (0, 1, 119):
            ref = _energy(so, E0, H0, so.H) / 2  # numpy sum, same definition
            assert w == pytest.approx(ref, rel=1e-12), n
    W = np.array(W)
    assert sa.last_path == path
    assert np.max(np.abs(W - W[0])) / W[0] < 1e-11


def test_energy_decreases_with_conductivity():
    print("energy lossy", flush=True)
    case = _box_case(True, seed=9)
    sa = gc.amd_solver(case, path="general")
    sa.one_step()          # J = sigma*E from here on
    sa.path = "fused"
    W = []
    for _ in range(60):
        sa.one_step()
        W.append(sa.energy())
    assert sa.last_path == "fused"
    W = np.array(W)
    assert np.all(np.diff(W) <= 1e-13 * abs(W[0]))
    assert W[-1] < W[0]


def test_energy_needs_the_previous_fields():
    case = _box_case(False)
    sa = gc.amd_solver(case, path="general")
    sa.one_step()
    with pytest.raises(RuntimeError, match="before the last step"):
        sa.energy()
    sa.keep_previous_fields = True
    sa.one_step()
    assert sa.energy() > 0
    sa.E[0, 0, 0, "x"] = 1.0       # a host write invalidates the kept state
    with pytest.raises(RuntimeError):
        sa.energy()


def test_loss_factor_matches_sum_and_impedance():
    import wakis_amd as wa
    print("loss factor", flush=True)
    x, y, z, mask = cases.cavity_pipe_geometry(20, 20, 40)
    sigmaz = 10e-3
    v = 299792458.0
    # a wide band (fmax = 3v/sigma) so the Parseval sum over the kept bins is complete
    wake = wa.WakeSolver(q=1e-9, sigmaz=sigmaz, beta=1.0, skip_cells=2, save=False,
                         verbose=0, fmax=3 * v / sigmaz)
    grid = wa.GridFIT3D(x=x, y=y, z=z, stl_solids={"cav": mask},
                        stl_materials={"cav": "vacuum"}, verbose=0)
    s = wa.SolverFIT3D(grid, wake=wake, bc_low=["pec", "pec", "abc"],
                       bc_high=["pec", "pec", "abc"], use_stl=True, bg="pec",
                       dtype=np.float64, verbose=0)
    s.wakesolve(wakelength=0.05)
    ds = wake.s[2] - wake.s[1]
    k_np = -np.sum(wake.WP * wake.lambdas) * ds
    assert wake.loss_factor == pytest.approx(k_np, rel=1e-12, abs=1e-15)
    # Parseval through the device DFT bins
    dsm = np.mean(wake.s[1:] - wake.s[:-1])
    N = int((v / dsm) // wake.fmax * 1001)
    df = v / (N * dsm)
    w = np.full(len(wake.f), 2.0)
    w[0] = 1.0
    k_z = df / (v ** 2 * 1e12) * np.sum(w * np.real(wake.Z) * np.abs(wake.lambdaf) ** 2)
    assert wake.loss_factor == pytest.approx(k_z, rel=1e-6)
