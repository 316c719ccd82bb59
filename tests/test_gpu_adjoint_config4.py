"""GPU: the reverse mode at the differentiable workload's size.

128x128x256 cavity/pipe (PEC background, PEC x/y, ABC z) with a 32x32x24 block
of cells on the axis carrying their own eps_r ~ U[1, 10] and sigma ~ U[0, 2]
S/m (seeded), 4,000 steps at the CFL dt, L = k_loss + 1e-15 * band power (the
band weight keeps |L| near 10^2: the central differences of the eps_r
gradients, ~250 times smaller than the sigma ones, then stay well above the
rounding noise of L):

* the FP64 adjoint gradient against central finite differences (h = 1e-4) of
  the FP64 forward on 16 seeded block cells, for eps_r and for sigma: 1e-3;
* the FP32 run's gradient (FP32 fields, FP32 adjoint arithmetic) against the
  FP64 gradient: relative L2 stated and bounded by 1e-2;
* autograd through the plain FP64 torch-sparse leapfrog at the 32x32x64
  downscale (an 8x8x6 block, 1,000 steps: the same physical time).
"""

import numpy as np
import pytest
import torch

from oracle import cases
from tests.test_gpu_adjoint import _torch_reference

pytestmark = pytest.mark.gpu

DIMS4 = (128, 128, 256)
INSERT4 = (32, 32, 24)
W_BAND = 1e-15


def _region(dims, insert):
    lo = [(n - m) // 2 for n, m in zip(dims, insert)]
    return tuple((a, a + m) for a, m in zip(lo, insert))


def _problem(dims, insert, dtype, steps, seed=0):
    import wakis_amd as wa
    from wakis_amd.adjoint import WakeAdjoint
    x, y, z, cav = cases.cavity_pipe_geometry(*dims)
    region = _region(dims, insert)
    ins = np.zeros(dims, dtype=bool)
    ins[tuple(slice(a, b) for a, b in region)] = True
    grid = wa.GridFIT3D(x=x, y=y, z=z, stl_solids={"cav": cav, "ins": ins},
                        stl_materials={"cav": "vacuum", "ins": [2.0, 1.0, 1.0]}, verbose=0)
    wake = wa.WakeSolver(q=1e-9, sigmaz=10e-3, beta=1.0, skip_cells=2, save=False, verbose=0)
    s = wa.SolverFIT3D(grid, wake=wake, bc_low=["pec", "pec", "abc"],
                       bc_high=["pec", "pec", "abc"], use_stl=True, bg="pec", dtype=dtype,
                       verbose=0, use_sibc=False, tau_clamp=False)
    v = wake.v
    wl = steps * float(s.dt) * v - wake.ti * v - (s.z.max() - s.z.min())
    adj = WakeAdjoint(s, region, wl, w_loss=1.0, w_band=W_BAND)
    rng = np.random.default_rng(seed)
    er = rng.uniform(1.0, 10.0, insert)
    sg = rng.uniform(0.0, 2.0, insert)
    return s, adj, er, sg, region, wl


def test_config4_adjoint_matches_finite_differences_and_fp32():
    print("config 4: 128x128x256, 32x32x24 block, 4000 steps, FD on 16 cells", flush=True)
    s, adj, er, sg, region, wl = _problem(DIMS4, INSERT4, np.float64, 4000)
    L0 = adj.forward(er, sg)
    assert abs(s.Nt - 4000) <= 1
    g_er, g_sg = (g.cpu().numpy() for g in adj.backward())
    rng = np.random.default_rng(7)
    cells = [tuple(int(rng.integers(0, n)) for n in INSERT4) for _ in range(16)]
    h = 1e-4
    fd_er, fd_sg = [], []
    for c in cells:
        for P, out in ((er, fd_er), (sg, fd_sg)):
            Pp, Pm = P.copy(), P.copy()
            Pp[c] += h
            Pm[c] -= h
            if P is er:
                Lp, Lm = adj.forward(Pp, sg), adj.forward(Pm, sg)
            else:
                Lp, Lm = adj.forward(er, Pp), adj.forward(er, Pm)
            out.append((Lp - Lm) / (2 * h))
    ga = np.array([g_er[c] for c in cells])
    gs = np.array([g_sg[c] for c in cells])
    fd_er, fd_sg = np.array(fd_er), np.array(fd_sg)
    err_er = np.abs(ga - fd_er).max() / np.abs(fd_er).max()
    err_sg = np.abs(gs - fd_sg).max() / np.abs(fd_sg).max()
    print("L", L0, "max|FD eps_r|", np.abs(fd_er).max(), "max|FD sigma|", np.abs(fd_sg).max(),
          "max err / max|FD|: eps_r", err_er, "sigma", err_sg, flush=True)
    assert np.abs(fd_er).max() > 0 and np.abs(fd_sg).max() > 0
    np.testing.assert_allclose(ga, fd_er, rtol=1e-3, atol=1e-3 * np.abs(fd_er).max())
    np.testing.assert_allclose(gs, fd_sg, rtol=1e-3, atol=1e-3 * np.abs(fd_sg).max())
    del s, adj
    torch.cuda.empty_cache()
    # the FP32 form: FP32 fields, FP32 adjoint fields and arithmetic
    s32, adj32, *_ = _problem(DIMS4, INSERT4, np.float32, 4000)
    adj32.forward(er, sg)
    e32, s32g = (g.cpu().numpy() for g in adj32.backward())
    assert adj32.used_adjoint_compute == "float32"
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    r_er, r_sg = rel(e32, g_er), rel(s32g, g_sg)
    print("FP32 run gradient vs FP64 gradient, relative L2: eps_r", r_er, "sigma", r_sg,
          flush=True)
    assert r_er < 1e-2 and r_sg < 1e-2


def test_config4_downscale_matches_torch_autograd():
    """32x32x64 (the differentiable workload at a quarter of the resolution):
    the device FP64 adjoint against autograd through the plain FP64 torch-sparse
    leapfrog of the same run."""
    print("config 4 downscale 32x32x64 vs torch autograd", flush=True)
    dims, insert = (32, 32, 64), (8, 8, 6)
    s, adj, er, sg, region, wl = _problem(dims, insert, np.float64, 1000, seed=1)
    L = adj.forward(er, sg)
    g_er, g_sg = (g.cpu().numpy() for g in adj.backward())
    Lt, t_er, t_sg = _torch_reference(er, sg, 1.0, W_BAND, s, dims=dims, region=region,
                                      wakelength=wl, ins_material=[2.0, 1.0, 1.0])
    print("L device", L, "torch", Lt, flush=True)
    assert L == pytest.approx(Lt, rel=1e-9)
    np.testing.assert_allclose(g_er, t_er, rtol=1e-6, atol=1e-6 * np.abs(t_er).max())
    np.testing.assert_allclose(g_sg, t_sg, rtol=1e-6, atol=1e-6 * np.abs(t_sg).max())
