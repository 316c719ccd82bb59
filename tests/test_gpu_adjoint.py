"""GPU: reverse-mode gradients of the wake figure of merit (wakis_amd.adjoint).

L = w_loss * k_loss + w_band * sum_{f < 3 GHz} |Z_z|^2 df over a lossy
dielectric insert (per-cell eps_r and sigma) in the cavity/pipe geometry,
FP64, checked three ways:

* central finite differences (h = 1e-4) on 16 seeded cells, agreement 1e-3;
* autograd through a plain FP64 torch-sparse leapfrog of the same run
  (operators assembled by the oracle, the same loss written with torch.fft);
* the loss value of that torch run equals the device forward's.
"""

import numpy as np
import pytest
import scipy.sparse as sp
import torch
from scipy.constants import c as c_light
from scipy.constants import epsilon_0 as eps_0

from oracle import cases
from oracle import fit_oracle as fo

pytestmark = pytest.mark.gpu

DIMS = (16, 16, 40)
REGION = ((6, 10), (6, 10), (17, 23))
WAKELENGTH = 0.03


def _setup(w_loss, w_band, mode="auto", kernel="auto", dtype=np.float64, adjoint_dtype="auto"):
    import wakis_amd as wa
    from wakis_amd.adjoint import WakeAdjoint
    x, y, z, mask = cases.cavity_pipe_geometry(*DIMS)
    ins = np.zeros(DIMS, dtype=bool)
    (i0, i1), (j0, j1), (k0, k1) = REGION
    ins[i0:i1, j0:j1, k0:k1] = True
    grid = wa.GridFIT3D(x=x, y=y, z=z, stl_solids={"cav": mask, "ins": ins},
                        stl_materials={"cav": "vacuum", "ins": [2.0, 1.0, 0.5]}, verbose=0)
    wake = wa.WakeSolver(q=1e-9, sigmaz=10e-3, beta=1.0, skip_cells=2, save=False, verbose=0)
    s = wa.SolverFIT3D(grid, wake=wake, bc_low=["pec", "pec", "abc"],
                       bc_high=["pec", "pec", "abc"], use_stl=True, bg="pec",
                       dtype=dtype, verbose=0, use_sibc=False, tau_clamp=False)
    return s, WakeAdjoint(s, REGION, WAKELENGTH, w_loss=w_loss, w_band=w_band, mode=mode,
                          kernel=kernel, adjoint_dtype=adjoint_dtype)


def _params(seed=0):
    rng = np.random.default_rng(seed)
    shape = tuple(b - a for a, b in REGION)
    return rng.uniform(1.0, 4.0, shape), rng.uniform(0.0, 1.0, shape)


@pytest.mark.parametrize("weights", [(1.0, 0.0), (0.0, 1e-13), (1.0, 1e-13)])
def test_adjoint_matches_finite_differences(weights):
    print("adjoint vs FD", weights, flush=True)
    s, adj = _setup(*weights)
    er, sg = _params()
    L0 = adj.forward(er, sg)
    g_er, g_sg = (g.cpu().numpy() for g in adj.backward())
    rng = np.random.default_rng(7)
    shape = er.shape
    cells = [tuple(int(rng.integers(0, n)) for n in shape) for _ in range(16)]
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
    print("L", L0, "max |dL/deps_r|", np.abs(fd_er).max(), "max |dL/dsigma|",
          np.abs(fd_sg).max(), flush=True)
    assert np.abs(fd_er).max() > 0 and np.abs(fd_sg).max() > 0
    np.testing.assert_allclose(ga, fd_er, rtol=1e-3, atol=1e-3 * np.abs(fd_er).max())
    np.testing.assert_allclose(gs, fd_sg, rtol=1e-3, atol=1e-3 * np.abs(fd_sg).max())


def _torch_reference(er, sg, w_loss, w_band, s_dev, dims=DIMS, region=REGION,
                     wakelength=WAKELENGTH, ins_material=(2.0, 1.0, 0.5)):
    """The same run and loss as a plain FP64 torch-sparse leapfrog on the CPU."""
    x, y, z, mask = cases.cavity_pipe_geometry(*dims)
    ins = np.zeros(dims, dtype=bool)
    (i0, i1), (j0, j1), (k0, k1) = region
    ins[i0:i1, j0:j1, k0:k1] = True
    so = fo.OracleSolver(x, y, z, ["pec", "pec", "abc"], ["pec", "pec", "abc"], bg="pec",
                         solids=[(mask, "vacuum"), (ins, list(ins_material))],
                         dtype=np.float64)
    so.dt = np.float64(s_dev.dt)
    n3 = 3 * so.N

    def tsp(M):
        M = M.tocoo()
        return torch.sparse_coo_tensor(np.vstack([M.row, M.col]), M.data, M.shape,
                                       dtype=torch.float64).coalesce()

    M1 = tsp(so.M1)
    K2 = tsp(sp.diags(so.itA) @ so.C.T.tocsr() @ sp.diags(so.tL))
    # parameterised ieps, sigma
    idx = []
    for c in range(3):
        for kk in range(k0, k1):
            for jj in range(j0, j1):
                for ii in range(i0, i1):
                    idx.append(c * so.N + ii + so.Nx * jj + so.Nx * so.Ny * kk)
    idx = torch.as_tensor(idx)
    er_t = torch.tensor(er, dtype=torch.float64, requires_grad=True)
    sg_t = torch.tensor(sg, dtype=torch.float64, requires_grad=True)
    flat = lambda P: P.permute(2, 1, 0).reshape(-1).repeat(3)  # x fastest, 3 components
    ieps = torch.as_tensor(so.ieps.copy()).index_put((idx,), 1.0 / (flat(er_t) * eps_0))
    sig = torch.as_tensor(so.sigma.copy()).index_put((idx,), flat(sg_t))
    ieps = ieps * torch.as_tensor(np.where(so.pec_mask == 0, 0.0, 1.0) if so.pec_mask is not None
                                  else np.ones(n3))
    # wakesolve: beam, window capture, Nt
    w = s_dev.wake
    v = w.v
    beam = fo.OracleBeam(q=w.q, sigmaz=w.sigmaz, beta=w.beta)
    tmax = (wakelength + w.ti * v + (so.z.max() - so.z.min())) / v
    Nt = int(tmax / so.dt)
    E = torch.zeros(n3, dtype=torch.float64)
    H = torch.zeros_like(E)
    J = torch.zeros_like(E)
    ixt, iyt = np.abs(so.x).argmin(), np.abs(so.y).argmin()
    win = [2 * so.N + (ixt + di) + so.Nx * (iyt + dj) + so.Nx * so.Ny * np.arange(so.Nz)
           for di in (-1, 0, 1) for dj in (-1, 0, 1)]
    caps = []
    jl = 2 * so.N + beam_line(so, beam) + so.Nx * so.Ny * np.arange(so.Nz)
    for n in range(Nt):
        jrow = torch.as_tensor(beam.profile(so, n * so.dt))
        J = J.index_put((torch.as_tensor(jl),), jrow)
        caps.append(torch.stack([E[torch.as_tensor(l)] for l in win]))
        H = H - so.dt * torch.sparse.mm(M1, E[:, None])[:, 0]
        E = E + so.dt * (ieps * torch.sparse.mm(K2, H[:, None])[:, 0] - ieps * J)
        J = sig * E
    cap = torch.stack(caps, dim=1)  # [9][Nt][Nz]
    # WP line integrals with the carry, loss
    t = np.arange(0, Nt * so.dt, so.dt)
    dt_w = t[2] - t[1]
    s_ax = np.arange(-w.ti * v, wakelength, dt_w * v)
    zf = so.z[w.skip_cells:-w.skip_cells]
    dz = so.dz[:len(zf)]
    zmin = so.z.min()
    Q = w.q * 1e12
    WP = torch.zeros(len(s_ax), dtype=torch.float64)
    for l in range(9):
        acc = torch.zeros(len(s_ax), dtype=torch.float64)
        for kk, zk in enumerate(zf):
            it = (((zk + s_ax) / v - zmin / v - t[0] + w.ti) / dt_w).astype(np.int64)
            ok = np.flatnonzero(it < Nt)
            acc = acc.index_add(0, torch.as_tensor(ok),
                                cap[l, torch.as_tensor(it[ok]), kk + w.skip_cells] * dz[kk])
        WP = (WP + acc) / Q
        if l == 4:
            break
    lam = torch.as_tensor(np.exp(-(s_ax ** 2) / (2 * w.sigmaz ** 2))
                          / (w.sigmaz * np.sqrt(2 * np.pi)))
    ds = s_ax[2] - s_ax[1]
    k_loss = -torch.sum(WP * lam) * ds
    dsm = np.mean(s_ax[1:] - s_ax[:-1])
    N = int((v / dsm) // w.fmax * 1001)
    f = np.fft.fftfreq(N, dsm / v)
    keep = np.flatnonzero((f >= 0) & (f < w.fmax))
    X = torch.fft.fft(WP * 1e12, n=N)[keep] * dsm
    Lf = torch.fft.fft(lam * v, n=N)[keep] * dsm
    Z = -X / Lf
    band = np.flatnonzero(f[keep] < 3e9)
    df = f[keep][1] - f[keep][0]
    L = w_loss * k_loss + w_band * torch.sum(torch.abs(Z[band]) ** 2) * df
    L.backward()
    return float(L), er_t.grad.numpy(), sg_t.grad.numpy()


def beam_line(so, beam):
    ixs = np.abs(so.x - beam.xsource).argmin()
    iys = np.abs(so.y - beam.ysource).argmin()
    beam.ixs, beam.iys, beam.zmin, beam.is_first_update = ixs, iys, so.z.min(), False
    return ixs + so.Nx * iys


@pytest.mark.parametrize("weights", [(1.0, 1e-13)])
def test_adjoint_matches_torch_autograd(weights):
    print("adjoint vs torch autograd", weights, flush=True)
    s, adj = _setup(*weights)
    er, sg = _params(3)
    L = adj.forward(er, sg)
    g_er, g_sg = (g.cpu().numpy() for g in adj.backward())
    Lt, t_er, t_sg = _torch_reference(er, sg, *weights, s)
    print("L device", L, "torch", Lt, flush=True)
    assert L == pytest.approx(Lt, rel=1e-9)
    np.testing.assert_allclose(g_er, t_er, rtol=1e-6, atol=1e-6 * np.abs(t_er).max())
    np.testing.assert_allclose(g_sg, t_sg, rtol=1e-6, atol=1e-6 * np.abs(t_sg).max())


def test_wake_loss_autograd_function():
    from wakis_amd.adjoint import wake_loss
    s, adj = _setup(1.0, 0.0)
    er, sg = _params(5)
    e = torch.tensor(er, requires_grad=True)
    g = torch.tensor(sg, requires_grad=True)
    L = wake_loss(e, g, adj)
    L.backward()
    ge, gs = adj.backward()
    assert np.allclose(e.grad.numpy(), ge.cpu().numpy())
    assert np.allclose(g.grad.numpy(), gs.cpu().numpy())


def test_trajectory_and_checkpoint_modes_agree_bitwise():
    """The region trajectory and the checkpoint/recompute schedule feed the
    reverse step the same forward values: identical gradients."""
    print("adjoint trajectory vs checkpoint", flush=True)
    out = {}
    for mode in ("trajectory", "checkpoint"):
        s, adj = _setup(1.0, 1e-13, mode=mode)
        er, sg = _params(11)
        L = adj.forward(er, sg)
        assert adj.used_mode == mode
        out[mode] = (L,) + tuple(g.cpu().numpy() for g in adj.backward())
    (La, ea, sa), (Lb, eb, sb) = out["trajectory"], out["checkpoint"]
    assert La == Lb
    assert np.array_equal(ea, eb) and np.array_equal(sa, sb)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("segment", [3, 7])
@pytest.mark.parametrize("full_rows", [0, 2])
def test_checkpoint_recompute_on_the_dependence_box(dtype, segment, full_rows):
    """Checkpoint mode recomputes each segment only on the tiles within
    segment + 3 cells of the block (a strict part of the grid here); the
    gradients equal the trajectory mode's bitwise."""
    import wakis_amd as wa
    from wakis_amd.adjoint import WakeAdjoint
    print("adjoint recompute box", np.dtype(dtype).name, segment, full_rows, flush=True)
    dims, region = (24, 40, 64), ((10, 14), (18, 22), (30, 34))
    x, y, z, mask = cases.cavity_pipe_geometry(*dims)
    ins = np.zeros(dims, dtype=bool)
    ins[tuple(slice(a, b) for a, b in region)] = True
    out = {}
    for mode in ("trajectory", "checkpoint"):
        grid = wa.GridFIT3D(x=x, y=y, z=z, stl_solids={"cav": mask, "ins": ins},
                            stl_materials={"cav": "vacuum", "ins": [2.0, 1.0, 0.5]},
                            verbose=0)
        wake = wa.WakeSolver(q=1e-9, sigmaz=10e-3, beta=1.0, skip_cells=2, save=False,
                             verbose=0)
        s = wa.SolverFIT3D(grid, wake=wake, bc_low=["pec", "pec", "abc"],
                           bc_high=["pec", "pec", "abc"], use_stl=True, bg="pec",
                           dtype=dtype, verbose=0, use_sibc=False, tau_clamp=False)
        s.fused_full_rows = full_rows
        adj = WakeAdjoint(s, region, WAKELENGTH, w_loss=1.0, w_band=1e-13, mode=mode,
                          segment=segment)
        er, sg = _params_of(region)
        L = adj.forward(er, sg)
        assert adj.used_mode == mode
        out[mode] = (L,) + tuple(g.cpu().numpy() for g in adj.backward())
    (La, ea, sa), (Lb, eb, sb) = out["trajectory"], out["checkpoint"]
    assert La == Lb and np.any(ea != 0)
    assert np.array_equal(ea, eb) and np.array_equal(sa, sb)


def _block_problem(dtype, dims, region):
    import wakis_amd as wa
    from wakis_amd.adjoint import WakeAdjoint
    x, y, z, mask = cases.cavity_pipe_geometry(*dims)
    ins = np.zeros(dims, dtype=bool)
    ins[tuple(slice(a, b) for a, b in region)] = True
    grid = wa.GridFIT3D(x=x, y=y, z=z, stl_solids={"cav": mask, "ins": ins},
                        stl_materials={"cav": "vacuum", "ins": [2.0, 1.0, 0.5]}, verbose=0)
    wake = wa.WakeSolver(q=1e-9, sigmaz=10e-3, beta=1.0, skip_cells=2, save=False, verbose=0)
    s = wa.SolverFIT3D(grid, wake=wake, bc_low=["pec", "pec", "abc"],
                       bc_high=["pec", "pec", "abc"], use_stl=True, bg="pec",
                       dtype=dtype, verbose=0, use_sibc=False, tau_clamp=False)
    return s, WakeAdjoint(s, region, WAKELENGTH, w_loss=1.0, w_band=1e-13)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("use_kappa", [True, False])
@pytest.mark.parametrize("escape", [False, True])
def test_reverse_kernel_codes_equal_arrays(dtype, use_kappa, escape):
    """The reverse kernel recomputing a, b, kappa (or sigma) and ieps from the
    forward's material codes gives bitwise the gradients of the run that reads
    the arrays; with more than 256 materials (an 8x8x8 block of its own
    values) the block's cells carry the escape code and read the arrays."""
    print("adjoint codes vs arrays", np.dtype(dtype).name, use_kappa, escape, flush=True)
    dims, region = ((24, 40, 64), ((8, 16), (16, 24), (28, 36))) if escape else (DIMS, REGION)
    out = {}
    for codes in (True, False):
        s, adj = _block_problem(dtype, dims, region)
        adj.use_codes = codes
        adj.use_kappa = use_kappa
        er, sg = _params_of(region, seed=17)
        L = adj.forward(er, sg)
        out[codes] = (L,) + tuple(g.cpu().numpy() for g in adj.backward())
        assert adj.used_kernel == "fused"
        assert s._code is not None and s._code_escape == escape
    (La, ea, sa), (Lb, eb, sb) = out[True], out[False]
    assert La == Lb and np.any(ea != 0)
    assert np.array_equal(ea, eb) and np.array_equal(sa, sb)


def test_trajectory_mode_needs_interior_region():
    import wakis_amd as wa
    from wakis_amd.adjoint import WakeAdjoint
    s, _ = _setup(1.0, 0.0)
    region = ((0, 4), (6, 10), (17, 23))
    with pytest.raises(ValueError):
        WakeAdjoint(s, region, WAKELENGTH, mode="trajectory").forward(*_params_of(region))
    adj = WakeAdjoint(s, region, WAKELENGTH)  # auto falls back to checkpoints
    adj.forward(*_params_of(region))
    assert adj.used_mode == "checkpoint"


def _params_of(region, seed=1):
    rng = np.random.default_rng(seed)
    shape = tuple(b - a for a, b in region)
    return rng.uniform(1.0, 4.0, shape), rng.uniform(0.0, 1.0, shape)


@pytest.mark.parametrize("mode", ["trajectory", "checkpoint"])
def test_fused_reverse_kernel_equals_two_pass(mode):
    """The single-pass reverse kernel sums the same terms in the same order as
    the two-kernel path: identical gradients, or within 1e-12 when it reads the
    precomputed conduction factor (and the default picks it here)."""
    print("adjoint fused vs two-pass", mode, flush=True)
    out = {}
    for kernel, kappa in (("fused", False), ("fused", True), ("two-pass", False)):
        s, adj = _setup(1.0, 1e-13, mode=mode, kernel=kernel)
        adj.use_kappa = kappa
        er, sg = _params(13)
        adj.forward(er, sg)
        out[kernel, kappa] = tuple(g.cpu().numpy() for g in adj.backward())
        assert adj.used_kernel == kernel
    ref = out["two-pass", False]
    assert np.abs(ref[0]).max() > 0
    for a, b in zip(out["fused", False], ref):
        assert np.array_equal(a, b)
    for a, b in zip(out["fused", True], ref):  # kappa = -dt ieps sigma: one rounding apart
        assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(b)
    s, adj = _setup(1.0, 0.0)
    adj.forward(*_params(2))
    adj.backward()
    assert adj.used_kernel == "fused"


def test_fused_reverse_kernel_rejects_low_face_region():
    from wakis_amd.adjoint import WakeAdjoint
    s, _ = _setup(1.0, 0.0)
    region = ((0, 4), (6, 10), (17, 23))
    adj = WakeAdjoint(s, region, WAKELENGTH, kernel="fused")
    adj.forward(*_params_of(region))
    with pytest.raises(ValueError):
        adj.backward()
    adj = WakeAdjoint(s, region, WAKELENGTH)
    adj.forward(*_params_of(region))
    adj.backward()
    assert adj.used_kernel == "two-pass"


def test_fp32_adjoint_fields_close_to_fp64():
    """FP32 forward run: adjoint fields stored in FP32, with FP64 or (the
    default) FP32 arithmetic, give the gradient of FP64-stored adjoint fields
    to 1e-4 relative L2, and all are within 1e-2 of the FP64 run's gradient
    (the FP32 forward's own error)."""
    print("adjoint FP32 storage", flush=True)
    er, sg = _params(21)
    got = {}
    for adt, comp in (("float32", "float32"), ("float32", "float64"), ("float64", "float64")):
        s, adj = _setup(1.0, 1e-13, dtype=np.float32, adjoint_dtype=adt)
        adj.adjoint_compute = comp
        adj.forward(er, sg)
        got[adt, comp] = tuple(g.cpu().numpy() for g in adj.backward())
        assert adj.used_kernel == "fused" and adj.used_adjoint_dtype == adt
        assert adj.used_adjoint_compute == comp
    s, adj = _setup(1.0, 1e-13, dtype=np.float32)
    adj.forward(er, sg)
    default = tuple(g.cpu().numpy() for g in adj.backward())
    assert adj.used_adjoint_compute == "float32"
    s, adj = _setup(1.0, 1e-13)
    adj.forward(er, sg)
    ref = tuple(g.cpu().numpy() for g in adj.backward())
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    for k in range(2):
        assert np.array_equal(default[k], got["float32", "float32"][k])
        base = got["float64", "float64"][k]
        for key in (("float32", "float32"), ("float32", "float64")):
            r32 = rel(got[key][k], base)
            r64 = rel(got[key][k], ref[k])
            print("grad", k, key, "vs fp64 storage", r32, "vs fp64 run", r64, flush=True)
            assert r32 < 1e-4
            assert r64 < 1e-2
    with pytest.raises(ValueError):
        s, adj = _setup(1.0, 1e-13, dtype=np.float32, adjoint_dtype="float64")
        adj.adjoint_compute = "float32"
        adj.forward(er, sg)
        adj.backward()


def test_lossless_adjoint_matches_finite_differences():
    """A block without conductivity (and a solver without any): gradient with
    respect to eps_r only, against central differences (h = 1e-4)."""
    import wakis_amd as wa
    from wakis_amd.adjoint import WakeAdjoint
    print("adjoint lossless vs FD", flush=True)
    x, y, z, mask = cases.cavity_pipe_geometry(*DIMS)
    ins = np.zeros(DIMS, dtype=bool)
    (i0, i1), (j0, j1), (k0, k1) = REGION
    ins[i0:i1, j0:j1, k0:k1] = True
    grid = wa.GridFIT3D(x=x, y=y, z=z, stl_solids={"cav": mask, "ins": ins},
                        stl_materials={"cav": "vacuum", "ins": [2.0, 1.0]}, verbose=0)
    wake = wa.WakeSolver(q=1e-9, sigmaz=10e-3, beta=1.0, skip_cells=2, save=False, verbose=0)
    s = wa.SolverFIT3D(grid, wake=wake, bc_low=["pec", "pec", "abc"],
                       bc_high=["pec", "pec", "abc"], use_stl=True, bg="pec",
                       dtype=np.float64, verbose=0, use_sibc=False)
    assert not s.use_conductivity
    adj = WakeAdjoint(s, REGION, WAKELENGTH, w_loss=1.0, w_band=1e-13)
    er, _ = _params(4)
    adj.forward(er)
    ge, gs = adj.backward()
    assert gs is None and adj.used_kernel == "fused"
    ge = ge.cpu().numpy()
    rng = np.random.default_rng(9)
    cells = [tuple(int(rng.integers(0, n)) for n in er.shape) for _ in range(8)]
    h = 1e-4
    fd = []
    for c in cells:
        Pp, Pm = er.copy(), er.copy()
        Pp[c] += h
        Pm[c] -= h
        fd.append((adj.forward(Pp) - adj.forward(Pm)) / (2 * h))
    fd = np.array(fd)
    assert np.abs(fd).max() > 0
    np.testing.assert_allclose([ge[c] for c in cells], fd, rtol=1e-3,
                               atol=1e-3 * np.abs(fd).max())


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_checkpoint_mode_with_live_wrap_terms(dtype):
    """x faces periodic with material on them: the flat-index wrap terms are
    live, so the run takes the exact general path, which steps every cell and
    couples i = Nx-1 to i = 0 of the next row. A block two cells from the x-low
    face must then get full-field checkpoints (a dependence-box restore would
    leave stale values at the far x face that the wrap carries into the block's
    neighbourhood); its gradients equal the trajectory mode's bitwise."""
    import wakis_amd as wa
    from wakis_amd.adjoint import WakeAdjoint
    print("adjoint checkpoint with live wraps", np.dtype(dtype).name, flush=True)
    dims, region = (16, 16, 40), ((2, 6), (6, 10), (17, 23))
    x, y, z, _ = cases.cavity_pipe_geometry(*dims)
    ins = np.zeros(dims, dtype=bool)
    ins[tuple(slice(a, b) for a, b in region)] = True
    out = {}
    for mode in ("trajectory", "checkpoint"):
        grid = wa.GridFIT3D(x=x, y=y, z=z, stl_solids={"ins": ins},
                            stl_materials={"ins": [2.0, 1.0, 0.5]}, verbose=0)
        wake = wa.WakeSolver(q=1e-9, sigmaz=10e-3, beta=1.0, skip_cells=2, save=False,
                             verbose=0)
        s = wa.SolverFIT3D(grid, wake=wake, bc_low=["periodic", "pec", "abc"],
                           bc_high=["periodic", "pec", "abc"], use_stl=True, bg="vacuum",
                           dtype=dtype, verbose=0, use_sibc=False, tau_clamp=False)
        adj = WakeAdjoint(s, region, WAKELENGTH, w_loss=1.0, w_band=1e-13, mode=mode,
                          segment=3)
        er, sg = _params_of(region, seed=23)
        L = adj.forward(er, sg)
        assert adj.used_mode == mode
        assert not s._wraps_dead() and s.last_path == "general"
        if mode == "checkpoint":
            assert adj._fwd["box"] is None
        out[mode] = (L,) + tuple(g.cpu().numpy() for g in adj.backward())
        assert adj.used_kernel == "two-pass"
    (La, ea, sa), (Lb, eb, sb) = out["trajectory"], out["checkpoint"]
    assert La == Lb and np.any(ea != 0) and np.any(sa != 0)
    assert np.array_equal(ea, eb) and np.array_equal(sa, sb)


@pytest.mark.parametrize("counter_moving", [False, True])
def test_wake_integral_adjoint_is_the_transpose(counter_moving):
    """<gI, I(cap)> = <gcap(gI), cap> for the line integral of wf_wake_wp (one
    line, no scaling) and wf_wake_wp_adjoint, on a random capture with a
    non-uniform dz (so that dz[k] and dz[nz-1-k] differ) and skipped rows."""
    from wakis_amd import _lib
    print("wake integral transpose", counter_moving, flush=True)
    rng = np.random.default_rng(21 + int(counter_moving))
    Nz, k0, nz, samples = 40, 3, 34, 300
    zn = np.cumsum(np.r_[-0.02, rng.uniform(0.8e-3, 1.2e-3, Nz)])
    zc = (zn[:-1] + np.diff(zn) / 2)[k0:k0 + nz]
    dz = np.diff(zn)[k0:k0 + nz].copy()
    v, dt, ti = c_light, 2e-12, 1.5e-10
    s_ax = np.arange(-ti * v, 0.05, dt * v)
    zref_v = (zc.max() / (-v)) if counter_moving else (zc.min() / v)
    dev = torch.device("cuda")
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev)
    cap = t(rng.standard_normal((1, samples, Nz)))
    gI = t(rng.standard_normal((1, len(s_ax))))
    I = torch.zeros((1, len(s_ax)), dtype=torch.float64, device=dev)
    zt, dzt, st_ = t(zc), t(dz), t(s_ax)
    L = _lib.lib()
    _lib.check(L.wf_wake_wp(_lib.ptr(cap), _lib.WF_F64, 1, samples, Nz, k0, nz, _lib.ptr(zt),
                            _lib.ptr(dzt), _lib.ptr(st_), len(s_ax), v, zref_v, 0.0, ti, dt,
                            samples, samples, int(counter_moving), 1.0, _lib.ptr(I),
                            _lib.stream_ptr()), "wf_wake_wp")
    gcap = torch.zeros_like(cap)
    _lib.check(L.wf_wake_wp_adjoint(_lib.ptr(gI), 1, samples, Nz, k0, nz, _lib.ptr(zt),
                                    _lib.ptr(dzt), _lib.ptr(st_), len(s_ax), v, zref_v, 0.0, ti,
                                    dt, samples, int(counter_moving), _lib.ptr(gcap),
                                    _lib.stream_ptr()), "wf_wake_wp_adjoint")
    lhs = float((gI * I).sum())
    rhs = float((gcap * cap).sum())
    scale = float((gI.abs() * I.abs()).sum())
    assert torch.count_nonzero(I) > len(s_ax) // 2
    assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs)
    # rows outside the kept range get no gradient
    assert not torch.any(gcap[..., :k0]) and not torch.any(gcap[..., k0 + nz:])


def test_counter_moving_adjoint_matches_finite_differences():
    """WakeAdjoint with a counter-moving WakeSolver (wakeSolver.py:503-517):
    dL/d eps_r and dL/d sigma against central differences (h = 1e-4)."""
    print("adjoint counter-moving vs FD", flush=True)
    s, adj = _setup(1.0, 1e-13)
    s.wake.counter_moving = True
    er, sg = _params(3)
    adj.forward(er, sg)
    g_er, g_sg = (g.cpu().numpy() for g in adj.backward())
    rng = np.random.default_rng(9)
    cells = [tuple(int(rng.integers(0, n)) for n in er.shape) for _ in range(8)]
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
    fd_er, fd_sg = np.array(fd_er), np.array(fd_sg)
    assert np.abs(fd_er).max() > 0 and np.abs(fd_sg).max() > 0
    np.testing.assert_allclose([g_er[c] for c in cells], fd_er, rtol=1e-3,
                               atol=1e-3 * np.abs(fd_er).max())
    np.testing.assert_allclose([g_sg[c] for c in cells], fd_sg, rtol=1e-3,
                               atol=1e-3 * np.abs(fd_sg).max())
