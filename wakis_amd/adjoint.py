"""Gradients of wake figures of merit with respect to per-cell eps_r and sigma.

No reference counterpart. The figure of merit of a wakefield run is

    L = w_loss * k_loss + w_band * sum_{f < f_band} |Z_z(f)|^2 * df

with k_loss = -sum_n WP(s_n) lambda(s_n) ds (WakeSolver.calc_loss_factor) and
Z_z the longitudinal impedance on the kept DFT bins (calc_long_Z). The
parameters are eps_r and sigma of every cell of a block of cells (all three
components of a cell share them, as a material fill does).

The gradient is the discrete adjoint of exactly what the forward computes:
the wake post-processing is reversed with wf_idft_real (impedance bins) and
wf_wake_wp_adjoint (line integrals, with the reference's line-to-line carry),
and the time stepping with wf_adjoint_step (transposed stencils).

The reverse step is linear in the adjoint fields and its coefficients do not
depend on the forward fields; only the parameter gradients read them, and only
on the block (E) and the block plus one cell below it on each axis (H). So the
forward pass keeps that trajectory of every step (wf_step's traj_* capture,
a few hundred KB per step for a 32x32x24 block) and the backward pass walks
the steps back without recomputing anything ("trajectory" mode). When the
trajectory would not fit in half of the free device memory, or the block
touches a low face of the grid, checkpoints of (E, H, J) are kept instead and
each segment is recomputed from its checkpoint before it is reversed
("checkpoint" mode): for a block off the low faces whose run takes the fused
step (flat-index wrap terms shown dead), segments of max(8, sqrt(Nt)/4) steps
with checkpoints and recompute confined to the block's domain of dependence
over a segment (a value moves at most one cell per step); otherwise
full-field checkpoints every sqrt(Nt) steps.
``wake_loss`` wraps it as a torch.autograd.Function.

The time step must not depend on sigma: build the solver with
``tau_clamp=False`` (and a dt for which dt*sigma/(eps_r*eps0) stays well below
2 over the parameter range).
"""

import ctypes
import math

import numpy as np
import torch
from scipy.constants import epsilon_0 as eps_0

from . import _lib

# The fused reverse kernel addresses 3N-double arrays with 32-bit byte offsets.
FUSED_MAX_BYTES = 2 ** 32


class WakeAdjoint:
    def __init__(self, solver, region, wakelength, f_band=3e9, w_loss=1.0, w_band=1.0,
                 segment=None, mode="auto", kernel="auto", adjoint_dtype="auto",
                 adjoint_compute="auto"):
        """``region`` = ((i0, i1), (j0, j1), (k0, k1)) cells, half-open.
        ``mode``: "trajectory", "checkpoint" or "auto" (trajectory when it fits).
        ``kernel``: "fused" (one single-pass reverse kernel per step), "two-pass"
        or "auto" (fused when the flat-index wrap terms are dead, as for the
        fused forward step, the region starts at cell 1 or later in x and y, and
        3N doubles fit in 4 GiB).
        ``adjoint_dtype``: storage of the adjoint fields, "float64", "float32"
        (fused kernel only) or "auto" (float32 with the fused kernel when the
        solver's fields are FP32, else float64).
        ``adjoint_compute``: arithmetic of the reverse step, "float64", or
        "float32" (FP32 adjoint fields only) or "auto" (the storage type of the
        adjoint fields); the adjoint of J on the block and the parameter
        gradients are accumulated in FP64 either way."""
        if adjoint_compute not in ("auto", "float32", "float64"):
            raise ValueError("WakeAdjoint: adjoint_compute must be 'auto', 'float32' or "
                             "'float64'")
        self.adjoint_compute = adjoint_compute
        self.used_adjoint_compute = None
        if adjoint_dtype not in ("auto", "float32", "float64"):
            raise ValueError("WakeAdjoint: adjoint_dtype must be 'auto', 'float32' or 'float64'")
        self.adjoint_dtype = adjoint_dtype
        self.used_adjoint_dtype = None
        if mode not in ("auto", "trajectory", "checkpoint"):
            raise ValueError("WakeAdjoint: mode must be 'auto', 'trajectory' or 'checkpoint'")
        if kernel not in ("auto", "fused", "two-pass"):
            raise ValueError("WakeAdjoint: kernel must be 'auto', 'fused' or 'two-pass'")
        self.kernel = kernel
        self.fused_rows = 0  # rows per workgroup of the fused reverse kernel (0: default)
        # fused kernel: read kappa = -dt ieps sigma instead of ieps and sigma (one
        # rounding of the conduction product differs from the two-kernel path)
        self.use_kappa = True
        # fused kernel: recompute the coefficients from the forward's material codes
        self.use_codes = True
        self.used_kernel = None
        s = solver
        if s.wake is None:
            raise ValueError("WakeAdjoint: the solver needs a WakeSolver")
        # sigma is a parameter only on a conductive solver (a conductive material,
        # background or PML); otherwise the gradient is with respect to eps_r alone
        self.conductive = bool(s.use_conductivity)
        (self.i0, self.i1), (self.j0, self.j1), (self.k0, self.k1) = region
        if not (0 <= self.i0 < self.i1 <= s.Nx and 0 <= self.j0 < self.j1 <= s.Ny
                and 0 <= self.k0 < self.k1 <= s.Nz):
            raise IndexError("WakeAdjoint: region outside the grid")
        self.s = s
        self.shape = (self.i1 - self.i0, self.j1 - self.j0, self.k1 - self.k0)
        self.wakelength = wakelength
        self.f_band, self.w_loss, self.w_band = f_band, w_loss, w_band
        self.segment = segment
        self.mode = mode
        self._fwd = None
        e = self.shape
        self._nE = e[0] * e[1] * e[2]
        self._nH = (e[0] + 1) * (e[1] + 1) * (e[2] + 1)

    # ------------------------------------------------------------- forward
    def _set_params(self, eps_r, sigma=None):
        s = self.s
        er = np.asarray(eps_r, dtype=np.float64).reshape(self.shape)
        if sigma is not None:
            sg = np.asarray(sigma, dtype=np.float64).reshape(self.shape)
            if not self.conductive and np.any(sg != 0):
                raise ValueError("WakeAdjoint: nonzero sigma needs a conductive solver")
        ie = (1.0 / (er * eps_0)).astype(s.dtype)
        sl = (slice(self.i0, self.i1), slice(self.j0, self.j1), slice(self.k0, self.k1))
        for c in "xyz":
            s.ieps[sl + (c,)] = ie
            if self.conductive and sigma is not None:
                s.sigma[sl + (c,)] = sg.astype(s.dtype)
        s.update_tensors()
        s.reset_fields()
        s.step_0 = True

    def _use_trajectory(self, Nt):
        if self.mode == "checkpoint":
            return False
        inside = min(self.i0, self.j0, self.k0) >= 1
        if self.mode == "trajectory":
            if not inside:
                raise ValueError("WakeAdjoint: trajectory mode needs the region to start at "
                                 "cell 1 or later on every axis")
            return True
        need = (Nt + 1) * 3 * (self._nE + self._nH) * self.s.E.array.element_size()
        need += (Nt + 1) * 3 * self._nE * 16  # region W and lJ of every reverse step
        free, _ = torch.cuda.mem_get_info(self.s.device)
        return inside and need <= free // 2

    def _dep_box(self, K):
        """Slices (z, y, x) of the cells within K + 3 cells of the block: what
        the block's trajectory over K steps depends on (a value moves at most
        one cell per step; the H box starts one cell lower)."""
        m = K + 3
        s = self.s
        return (slice(max(self.k0 - m, 0), min(self.k1 + m, s.Nz)),
                slice(max(self.j0 - m, 0), min(self.j1 + m, s.Ny)),
                slice(max(self.i0 - m, 0), min(self.i1 + m, s.Nx)))

    def _lohi(self):
        lo = (ctypes.c_int32 * 3)(self.i0, self.j0, self.k0)
        hi = (ctypes.c_int32 * 3)(self.i1, self.j1, self.k1)
        return lo, hi

    def forward(self, eps_r, sigma=None):
        """Run the wakefield simulation with these parameters (sigma: None keeps
        the solver's, as for a solver without conductivity); return L."""
        s = self.s
        self._fwd = None  # release the previous run's trajectory or checkpoints
        self._set_params(eps_r, sigma)
        ctx = s._wake_setup(self.wakelength, "both")
        Nt = ctx["Nt"]
        if self._use_trajectory(Nt):
            dev, dt_ = s.device, s.E.array.dtype
            tE = torch.empty((Nt + 1, 3 * self._nE), dtype=dt_, device=dev)
            tH = torch.empty((Nt + 1, 3 * self._nH), dtype=dt_, device=dev)
            lo, hi = self._lohi()
            _lib.check(_lib.lib().wf_trajectory_gather(
                s._handle, _lib.ptr(s.E.array), _lib.ptr(s.H.array), _lib.ptr(tE), _lib.ptr(tH),
                lo, hi, Nt + 1, 0, _lib.stream_ptr()), "wf_trajectory_gather")
            if Nt > 0:
                args = s._wake_args(ctx, 0, Nt)
                args.traj_E, args.traj_H = _lib.ptr(tE), _lib.ptr(tH)
                args.traj_lo[:] = tuple(lo)
                args.traj_hi[:] = tuple(hi)
                args.traj_slots, args.traj_slot0 = Nt + 1, 0
                s._run(args)
            fwd = dict(ctx=ctx, Nt=Nt, traj=(tE, tH))
        else:
            root = int(math.ceil(math.sqrt(max(Nt, 1))))
            interior = min(self.i0, self.j0, self.k0) >= 1
            # interior block whose recompute runs on the fused path: each
            # checkpoint keeps only the block's domain of dependence over one
            # segment, so shorter segments (sqrt(Nt)/4) shrink both the
            # recompute and the checkpoints. The fused path has shown the
            # flat-index wrap terms dead and launches only the tiles that meet
            # the box; the exact general path steps every cell, and its live
            # wrap terms (x/y periodic, PMC or PML faces with material on them)
            # carry values from the far x/y faces into the block's
            # neighbourhood in one step, so it needs full-field checkpoints.
            local = interior and s._choose_fused(s._wake_args(ctx, 0, 1))
            K = self.segment or (max(8, root // 4) if local else root)
            box = self._dep_box(K) if local else None
            ckpt = []
            for n0 in range(0, Nt, K):
                ckpt.append((n0,) + tuple(
                    F.array.clone() if box is None
                    else F.array.view(3, s.Nz, s.Ny, s.Nx)[(slice(None),) + box].clone()
                    for F in (s.E, s.H, s.J)))
                k = min(K, Nt - n0)
                s._run(s._wake_args(ctx, n0, k))
            fwd = dict(ctx=ctx, ckpt=ckpt, K=K, Nt=Nt, box=box)
        torch.cuda.synchronize(s.device)
        s._wake_finish(ctx, "both")
        self._fwd = fwd
        return self._loss_value()

    def _use_fused_kernel(self):
        if self.kernel == "two-pass":
            return False
        # the fused reverse kernel addresses its FP64-sized 3N arrays with 32-bit
        # byte offsets; larger grids take the two-pass kernels (64-bit indices)
        small = 3 * self.s.N * 8 < FUSED_MAX_BYTES
        ok = small and self.i0 >= 1 and self.j0 >= 1 and self.s._wraps_dead()
        if self.kernel == "fused" and not ok:
            raise ValueError("WakeAdjoint: the fused reverse kernel needs dead flat-index wrap "
                             "terms, the region at cell 1 or later in x and y, and 3N "
                             "doubles within 4 GiB")
        return ok

    def _adjoint_f32(self, adj_fused):
        """Whether the adjoint fields are stored in FP32: asked for, or "auto"
        with the fused reverse kernel on a solver with FP32 fields."""
        f32 = self.adjoint_dtype == "float32" or (self.adjoint_dtype == "auto" and adj_fused
                                                   and self.s.E.array.dtype == torch.float32)
        if f32 and not adj_fused:
            raise ValueError("WakeAdjoint: FP32 adjoint fields need the fused reverse kernel")
        return f32

    @property
    def used_mode(self):
        """"trajectory" or "checkpoint": how the last forward kept its states."""
        if self._fwd is None:
            return None
        return "trajectory" if "traj" in self._fwd else "checkpoint"

    def _band(self):
        w = self.s.wake
        return np.flatnonzero(w.f < self.f_band)

    def _loss_value(self):
        w = self.s.wake
        k = w.loss_factor
        band = self._band()
        df = w.f[1] - w.f[0] if len(w.f) > 1 else 0.0
        zb = float(np.sum(np.abs(w.Z[band]) ** 2) * df)
        self.k_loss, self.band_power = k, zb
        return self.w_loss * k + self.w_band * zb

    # ------------------------------------------------------------ backward
    def _grad_wp(self):
        """dL/dWP on the centre line (host arrays of len(s))."""
        s, w = self.s, self.s.wake
        dev = s.device
        ds = w.s[2] - w.s[1]
        g = -self.w_loss * w.lambdas * ds
        band = self._band()
        if self.w_band != 0 and len(band):
            dsm = np.mean(w.s[1:] - w.s[:-1])
            N = int((w.v / dsm) // w.fmax * 1001)
            df = w.f[1] - w.f[0]
            # L_band = sum |Z_f|^2 df, Z_f = -1e12 dsm DFT(WP)_f / lambdaf_f
            coef = -2.0 * df * 1e12 * dsm * np.conj(w.Z[band]) / w.lambdaf[band]
            nb = int(band.max()) + 1
            wv = np.zeros((nb, 2))
            wv[band, 0] = coef.real
            wv[band, 1] = coef.imag
            wt = torch.as_tensor(wv, device=dev)
            out = torch.zeros(len(w.s), dtype=torch.float64, device=dev)
            _lib.check(_lib.lib().wf_idft_real(_lib.ptr(wt), nb, N, len(w.s), _lib.ptr(out),
                                               _lib.stream_ptr()), "wf_idft_real")
            g = g + self.w_band * out.cpu().numpy()
        return g

    def _grad_capture(self, gWP):
        """dL/d(capture) through the line integrals with the line-to-line carry."""
        s, w = self.s, self.s.wake
        cap = w._cap
        nl = cap.shape[0]
        Q = w.q * 1e12
        centre = nl // 2
        gI = np.zeros((nl, len(w.s)))
        for l in range(centre + 1):  # WP_3d[l] = (WP_3d[l-1] + I_l) / Q, loss on the centre
            gI[l] = gWP / Q ** (centre - l + 1)
        s_ax, dt, zmin = w._s_axis()
        zz = slice(w.skip_cells, -w.skip_cells) if w.skip_cells != 0 else np.s_[:]
        z = np.ascontiguousarray(w.zf[zz], dtype=np.float64)
        nz = len(z)
        k0 = w.skip_cells if w.skip_cells != 0 else 0
        nt = min(len(w.t), w._cap_samples)
        dev = cap.device
        gcap = torch.zeros(cap.shape, dtype=torch.float64, device=dev)
        t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev)
        gIt, zt, dzt, st_ = t(gI), t(z), t(w.dz[:max(nz, 1)]), t(s_ax)
        if w.counter_moving:  # reference time zmax/(-v) (wakeSolver.py:505-511)
            zref_v = float(np.max(w.zf) / (-1 * w.v))
        else:  # zmin/v (wakeSolver.py:524-528)
            zref_v = float(zmin / w.v)
        _lib.check(_lib.lib().wf_wake_wp_adjoint(
            _lib.ptr(gIt), nl, cap.shape[1], cap.shape[2], k0, nz, _lib.ptr(zt), _lib.ptr(dzt),
            _lib.ptr(st_), len(s_ax), float(w.v), zref_v, float(w.t[0]), float(w.ti), float(dt),
            nt, int(bool(w.counter_moving)), _lib.ptr(gcap), _lib.stream_ptr()),
            "wf_wake_wp_adjoint")
        return gcap

    def _reverse_setup(self):
        """Adjoint fields, wf_adjoint_args and wf_adjoint_grad_args of the last
        forward run (the reverse pass's state), as a dict."""
        if self._fwd is None:
            raise RuntimeError("WakeAdjoint.backward: run forward first")
        s = self.s
        f = self._fwd
        ctx, Nt = f["ctx"], f["Nt"]
        dev = s.device
        gcap = self._grad_capture(self._grad_wp())
        n3 = 3 * s.N
        adj_fused = self._use_fused_kernel()
        self.used_kernel = "fused" if adj_fused else "two-pass"
        f32 = self._adjoint_f32(adj_fused)
        self.used_adjoint_dtype = "float32" if f32 else "float64"
        c32 = f32 and self.adjoint_compute in ("auto", "float32")
        if self.adjoint_compute == "float32" and not f32:
            raise ValueError("WakeAdjoint: FP32 arithmetic needs FP32 adjoint fields")
        self.used_adjoint_compute = "float32" if c32 else "float64"
        adt = torch.float32 if f32 else torch.float64
        W = torch.zeros(n3, dtype=adt, device=dev)
        lH = torch.zeros_like(W)
        nreg = int(np.prod(self.shape))
        lJr = torch.zeros(3 * nreg, dtype=torch.float64, device=dev)
        g_ieps = torch.zeros(3 * nreg, dtype=torch.float64, device=dev)
        g_sig = torch.zeros_like(g_ieps)
        ix0, iy0, nx, ny = ctx["win"]
        beam = ctx["beam"]
        table = ctx["table"]
        a = _lib.WfAdjointArgs()
        a.W, a.lH, a.lJ_region = _lib.ptr(W), _lib.ptr(lH), _lib.ptr(lJr)
        a.dt = float(s.dt)
        a.conductive = int(self.conductive)
        a.ixs, a.iys = int(beam.ixs), int(beam.iys)
        a.r0[:] = (self.i0, self.j0, self.k0)
        a.r1[:] = (self.i1, self.j1, self.k1)
        a.g_ieps, a.g_sigma = _lib.ptr(g_ieps), _lib.ptr(g_sig)
        a.gcap = _lib.ptr(gcap)
        a.cap_ix0, a.cap_iy0, a.cap_nx, a.cap_ny = ix0, iy0, nx, ny
        a.cap_samples = gcap.shape[1]
        # the coefficients stay those of the forward (the same solver)
        a.a, a.b = _lib.ptr(s._a), _lib.ptr(s._b)
        a.ieps, a.sigma = _lib.ptr(s.ieps.array), _lib.ptr(s.sigma.array)
        a.metric2d = _lib.ptr(s._metric2d_all)
        lib, st = _lib.lib(), _lib.stream_ptr()
        row_bytes = table.shape[1] * table.element_size()
        W2 = lH2 = kap = None
        if adj_fused:
            W2, lH2 = torch.zeros_like(W), torch.zeros_like(lH)
            a.fused = int(self.fused_rows) or 1
            a.W2, a.lH2 = _lib.ptr(W2), _lib.ptr(lH2)
            a.fields_f32 = int(f32)
            a.compute_f32 = int(c32)
            # the forward's material codes, when it has them: one byte per cell
            # instead of the coefficient arrays (the same values)
            if self.use_codes and s.use_material_code and s._material_table():
                a.code, a.mat, a.code_metric2d = (_lib.ptr(s._code), _lib.ptr(s._mat),
                                                  _lib.ptr(s._metric2d))
                a.nmat, a.code_escape = s._nmat, int(s._code_escape)
            if self.use_kappa and self.conductive:
                kap = (-float(s.dt) * s.ieps.array.to(torch.float64)
                       * s.sigma.array.to(torch.float64))
                kz = 2 * s.N + int(beam.ixs) + s.Nx * int(beam.iys) + s.Nx * s.Ny * torch.arange(
                    s.Nz, device=dev)
                kap[kz] = 0.0  # the beam line's J is a source, not sigma E
                kap = kap.to(s.ieps.array.dtype)
                a.kappa = _lib.ptr(kap)

        def adjoint_step():
            _lib.check(lib.wf_adjoint_step(s._handle, ctypes.byref(a), st), "wf_adjoint_step")
            if adj_fused:  # the outputs are the next step's inputs
                a.W, a.W2 = a.W2, a.W
                a.lH, a.lH2 = a.lH2, a.lH

        # fused kernel: the steps store the region's W and lJ, and one pass per
        # run (or per segment) accumulates the gradients
        n3r = 3 * nreg
        ga = _lib.WfAdjointGradArgs()
        ga.sigma, ga.metric2d = a.sigma, a.metric2d
        ga.fields_f32 = int(f32)
        ga.dt, ga.conductive, ga.ixs, ga.iys = a.dt, a.conductive, a.ixs, a.iys
        ga.r0[:] = tuple(a.r0)
        ga.r1[:] = tuple(a.r1)
        ga.g_ieps, ga.g_sigma = _lib.ptr(g_ieps), _lib.ptr(g_sig)

        def grad_pass(E_, Es, H_, Hs, traj, beam_ptr, rW_, rJ_, nsteps):
            ga.E, ga.E_stride, ga.H, ga.H_stride, ga.traj = E_, Es, H_, Hs, traj
            ga.beam = beam_ptr
            ga.reg_W, ga.reg_lJ, ga.nsteps = _lib.ptr(rW_), _lib.ptr(rJ_), nsteps
            _lib.check(lib.wf_adjoint_grad(s._handle, ctypes.byref(ga), st), "wf_adjoint_grad")

        return dict(a=a, ga=ga, adjoint_step=adjoint_step, grad_pass=grad_pass, lib=lib, st=st,
                    fused=adj_fused, f32=f32, adt=adt, n3=n3, nreg=nreg, n3r=n3r, table=table,
                    row_bytes=row_bytes, g_ieps=g_ieps, g_sig=g_sig,
                    keep=[W, lH, lJr, gcap, W2, lH2, kap])

    def backward(self):
        """(dL/d eps_r, dL/d sigma) of the last forward, as FP64 torch tensors
        (dL/d sigma is None on a solver without conductivity)."""
        R = self._reverse_setup()
        s, f = self.s, self._fwd
        ctx, Nt = f["ctx"], f["Nt"]
        dev = s.device
        a, adjoint_step, grad_pass = R["a"], R["adjoint_step"], R["grad_pass"]
        adj_fused, adt, n3, n3r = R["fused"], R["adt"], R["n3"], R["n3r"]
        table, row_bytes, lib, st = R["table"], R["row_bytes"], R["lib"], R["st"]
        g_ieps, g_sig = R["g_ieps"], R["g_sig"]
        if "traj" in f:
            tE, tH = f["traj"]
            a.traj = 1
            pE, pH = tE.data_ptr(), tH.data_ptr()
            sE, sH = tE.stride(0) * tE.element_size(), tH.stride(0) * tH.element_size()
            pT = table.data_ptr()
            if adj_fused:
                rW = torch.empty((max(Nt, 1), n3r), dtype=adt, device=dev)
                rJ = torch.zeros((Nt + 1, n3r), dtype=torch.float64, device=dev)
                a.reg_W, a.reg_lJ, a.reg_slots = _lib.ptr(rW), _lib.ptr(rJ), max(Nt, 1)
                a.g_ieps = a.g_sigma = None
            for n in range(Nt - 1, -1, -1):
                a.E0, a.E1 = pE + n * sE, pE + (n + 1) * sE
                a.H1 = pH + (n + 1) * sH
                a.beam_row = pT + n * row_bytes
                a.cap_sample = n
                a.reg_slot = n
                adjoint_step()
            if adj_fused:
                grad_pass(pE, tE.stride(0), pH, tH.stride(0), 1, pT, rW, rJ, Nt)
            torch.cuda.synchronize(dev)
            return self._finish_grad(g_ieps, g_sig)
        K = f["K"]
        if min(self.i0, self.j0, self.k0) >= 1:
            return self._backward_segments(R)
        J_final = s.J.array.clone()
        # one step's forward arguments; each recomputed step reads state i of the
        # segment buffers and writes state i+1 (fused path), or steps a copy in place
        step = s._wake_args(ctx, 0, 1)
        fused = s._prepare_fused(step)
        Eb = torch.empty((K + 1, n3), dtype=s.E.array.dtype, device=dev)
        Hb = torch.empty_like(Eb)
        if adj_fused:
            rW = torch.empty((K, n3r), dtype=adt, device=dev)
            rJ = torch.zeros((K + 1, n3r), dtype=torch.float64, device=dev)
            carry = torch.zeros(n3r, dtype=torch.float64, device=dev)  # lJ after the segment
            a.reg_W, a.reg_lJ, a.reg_slots = _lib.ptr(rW), _lib.ptr(rJ), K
            a.g_ieps = a.g_sigma = None
        for (n0, E0, H0, J0) in reversed(f["ckpt"]):
            k = min(K, Nt - n0)
            if adj_fused:
                rJ[k].copy_(carry)
            Eb[0].copy_(E0)
            Hb[0].copy_(H0)
            s.J.array.copy_(J0)
            for i in range(k):
                n = n0 + i
                if fused:
                    step.E, step.H = _lib.ptr(Eb[i]), _lib.ptr(Hb[i])
                    step.E2, step.H2 = _lib.ptr(Eb[i + 1]), _lib.ptr(Hb[i + 1])
                else:
                    Eb[i + 1].copy_(Eb[i])
                    Hb[i + 1].copy_(Hb[i])
                    step.E, step.H = _lib.ptr(Eb[i + 1]), _lib.ptr(Hb[i + 1])
                step.nsteps = 1
                step.beam_row0, step.cap_sample0 = n, n + 1
                _lib.check(lib.wf_step(s._handle, ctypes.byref(step), st), "wf_step")
            for i in range(k - 1, -1, -1):
                n = n0 + i
                a.E0, a.E1, a.H1 = _lib.ptr(Eb[i]), _lib.ptr(Eb[i + 1]), _lib.ptr(Hb[i + 1])
                a.beam_row = ctypes.c_void_p(table.data_ptr() + n * row_bytes)
                a.cap_sample = n
                a.reg_slot = i
                adjoint_step()
            if adj_fused:
                grad_pass(Eb.data_ptr(), n3, Hb.data_ptr(), n3, 0,
                          table.data_ptr() + n0 * row_bytes, rW, rJ, k)
                carry.copy_(rJ[0])
        s.J.array.copy_(J_final)
        torch.cuda.synchronize(dev)
        del Eb, Hb
        return self._finish_grad(g_ieps, g_sig)

    def _backward_segments(self, R):
        """Checkpoint mode for a block away from the low faces: each segment is
        recomputed from its checkpoint by one multi-step wf_step call that
        captures only the block's trajectory (as the trajectory mode's forward
        does) and updates only the block's domain of dependence over the
        segment, then reversed from that trajectory. The solver's fields are
        restored to the end of the run afterwards."""
        s, f = self.s, self._fwd
        ctx, Nt, K = f["ctx"], f["Nt"], f["K"]
        dev = s.device
        a, adjoint_step, grad_pass = R["a"], R["adjoint_step"], R["grad_pass"]
        adj_fused, adt, n3r = R["fused"], R["adt"], R["n3r"]
        table, row_bytes = R["table"], R["row_bytes"]
        lib, st = R["lib"], R["st"]
        finals = [F.array.clone() for F in (s.E, s.H, s.J)]
        dt_ = s.E.array.dtype
        tE = torch.empty((K + 1, 3 * self._nE), dtype=dt_, device=dev)
        tH = torch.empty((K + 1, 3 * self._nH), dtype=dt_, device=dev)
        pE, pH, pT = tE.data_ptr(), tH.data_ptr(), table.data_ptr()
        sE, sH = tE.stride(0) * tE.element_size(), tH.stride(0) * tH.element_size()
        lo, hi = self._lohi()
        a.traj = 1
        if adj_fused:
            rW = torch.empty((K, n3r), dtype=adt, device=dev)
            rJ = torch.zeros((K + 1, n3r), dtype=torch.float64, device=dev)
            carry = torch.zeros(n3r, dtype=torch.float64, device=dev)  # lJ after the segment
            a.reg_W, a.reg_lJ, a.reg_slots = _lib.ptr(rW), _lib.ptr(rJ), K
            a.g_ieps = a.g_sigma = None
        box = f["box"]
        for (n0, E0, H0, J0) in reversed(f["ckpt"]):
            k = min(K, Nt - n0)
            s._load_run_state(E0, H0, J0, box=box)
            _lib.check(lib.wf_trajectory_gather(s._handle, _lib.ptr(s.E.array),
                                                _lib.ptr(s.H.array), pE, pH, lo, hi, K + 1, 0,
                                                st), "wf_trajectory_gather")
            args = s._wake_args(ctx, n0, k)
            args.capture = None  # the capture of the forward run stays as it is
            args.traj_E, args.traj_H = pE, pH
            args.traj_lo[:] = tuple(lo)
            args.traj_hi[:] = tuple(hi)
            args.traj_slots, args.traj_slot0 = K + 1, 0
            # only the block's domain of dependence over the k steps is recomputed:
            # a cell's value moves at most one cell per step, so the tiles within
            # k + 3 cells of the block (E box [lo, hi), H box [lo - 1, hi)) give
            # the block's trajectory exactly
            m = k + 3  # within the restored box (k <= K)
            args.box_lo[:] = tuple(int(v) - m for v in lo)
            args.box_hi[:] = tuple(int(v) + m for v in hi)
            if box is not None and not s._choose_fused(args):
                raise RuntimeError("WakeAdjoint: a box checkpoint must be recomputed on the "
                                   "fused path (the general path steps every cell)")
            s._run(args)
            if adj_fused:
                rJ[k].copy_(carry)
            for i in range(k - 1, -1, -1):
                n = n0 + i
                a.E0, a.E1 = pE + i * sE, pE + (i + 1) * sE
                a.H1 = pH + (i + 1) * sH
                a.beam_row = pT + n * row_bytes
                a.cap_sample = n
                a.reg_slot = i
                adjoint_step()
            if adj_fused:
                grad_pass(pE, tE.stride(0), pH, tH.stride(0), 1, pT + n0 * row_bytes, rW, rJ, k)
                carry.copy_(rJ[0])
        s._load_run_state(*finals)
        torch.cuda.synchronize(dev)
        return self._finish_grad(R["g_ieps"], R["g_sig"])

    def _finish_grad(self, g_ieps, g_sig):
        s = self.s
        sh = self.shape
        # per cell (x fastest): sum the three components; ieps = 1/(eps_r eps0)
        gi = g_ieps.view(3, sh[2], sh[1], sh[0]).sum(0).permute(2, 1, 0)
        gs = None
        if self.conductive:
            gs = g_sig.view(3, sh[2], sh[1], sh[0]).sum(0).permute(2, 1, 0).contiguous()
        ie = s.ieps.to_matrix("x")[self.i0:self.i1, self.j0:self.j1,
                                   self.k0:self.k1].to(torch.float64)
        # d ieps / d eps_r = -eps0 ieps^2
        g_eps = gi * (-eps_0 * ie * ie)
        return g_eps.contiguous(), gs


class _WakeLossFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eps_r, sigma, adj):
        sg = None if sigma is None else sigma.detach().cpu().numpy()
        L = adj.forward(eps_r.detach().cpu().numpy(), sg)
        ctx.adj = adj
        ctx.dev = eps_r.device
        return torch.tensor(L, dtype=torch.float64, device=eps_r.device)

    @staticmethod
    def backward(ctx, gL):
        ge, gs = ctx.adj.backward()
        return ((gL * ge).to(ctx.dev), None if gs is None else (gL * gs).to(ctx.dev), None)


def wake_loss(eps_r, sigma, adj):
    """L(eps_r, sigma) as a differentiable torch scalar (see WakeAdjoint);
    ``sigma`` None on a solver without conductivity."""
    return _WakeLossFn.apply(eps_r, sigma, adj)
