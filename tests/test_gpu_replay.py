"""GPU: deterministic replay. Running the same problem twice in one process
gives bitwise the same fields, wake potentials, impedance, energy and
gradients: no kernel of the path uses atomics or an order that depends on
scheduling (the reductions combine per-block partials in a fixed order)."""

import numpy as np
import pytest

from oracle import cases

pytestmark = pytest.mark.gpu


def _wake_run(dtype):
    import wakis_amd as wa
    x, y, z, mask = cases.cavity_pipe_geometry(24, 24, 48)
    grid = wa.GridFIT3D(x=x, y=y, z=z, stl_solids={"cav": mask},
                        stl_materials={"cav": "vacuum"}, verbose=0)
    wake = wa.WakeSolver(q=1e-9, sigmaz=10e-3, beta=1.0, skip_cells=2, save=False, verbose=0)
    s = wa.SolverFIT3D(grid, wake=wake, bc_low=["pec", "pec", "abc"],
                       bc_high=["pec", "pec", "abc"], use_stl=True, bg="pec", dtype=dtype,
                       verbose=0)
    s.wakesolve(wakelength=0.05)
    s.one_step()
    return dict(E=s.E.to_numpy(), H=s.H.to_numpy(), WP_3d=wake.WP_3d, WPx=wake.WPx,
                Z=wake.Z, Zx=wake.Zx, k=np.array(wake.loss_factor), U=np.array(s.energy()))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_wakefield_run_replays_bitwise(dtype):
    print("replay wakesolve", np.dtype(dtype).name, flush=True)
    a, b = _wake_run(dtype), _wake_run(dtype)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_adjoint_gradient_replays_bitwise():
    import wakis_amd as wa
    from wakis_amd.adjoint import WakeAdjoint
    print("replay adjoint", flush=True)
    dims, region = (16, 16, 40), ((6, 10), (6, 10), (17, 23))
    x, y, z, mask = cases.cavity_pipe_geometry(*dims)
    ins = np.zeros(dims, dtype=bool)
    ins[tuple(slice(a, b) for a, b in region)] = True
    rng = np.random.default_rng(4)
    er, sg = rng.uniform(1.0, 4.0, (4, 4, 6)), rng.uniform(0.0, 1.0, (4, 4, 6))
    out = []
    for _ in range(2):
        grid = wa.GridFIT3D(x=x, y=y, z=z, stl_solids={"cav": mask, "ins": ins},
                            stl_materials={"cav": "vacuum", "ins": [2.0, 1.0, 0.5]}, verbose=0)
        wake = wa.WakeSolver(q=1e-9, sigmaz=10e-3, beta=1.0, skip_cells=2, save=False,
                             verbose=0)
        s = wa.SolverFIT3D(grid, wake=wake, bc_low=["pec", "pec", "abc"],
                           bc_high=["pec", "pec", "abc"], use_stl=True, bg="pec",
                           dtype=np.float32, verbose=0, use_sibc=False, tau_clamp=False)
        adj = WakeAdjoint(s, region, 0.03, w_loss=1.0, w_band=1e-13)
        L = adj.forward(er, sg)
        out.append((L,) + tuple(g.cpu().numpy() for g in adj.backward()))
    (La, ea, sa), (Lb, eb, sb) = out
    assert La == Lb and np.array_equal(ea, eb) and np.array_equal(sa, sb)
