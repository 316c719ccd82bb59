"""SolverFIT3D on the MI355X engine.

Mirrors ``wakis.solverFIT3D.SolverFIT3D`` (solverFIT3D.py:40-354, 416-503,
820-1029) and ``wakis.boundaries.BCsMixin`` (boundaries.py:14-311): same
constructor, same exposed Fields (E, H, J, ieps, imu, sigma), same dt and
boundary semantics, ``one_step`` as the seam (solverFIT3D.py:142). The
3N x 3N CSR operators are never built: the curl runs as a HIP stencil on
coefficient arrays folded once from the same diagonals,

    a = fl(iA*imu) * (PMC row mask)      (rows of M1 = iDa*iDmu*C*Ds)
    b = fl(ieps*itA) * (PEC row mask)    (rows of M2 = iDeps*itDa*C^T*tDs)

with the edge lengths L, tL kept as per-axis vectors and the PEC/PMC column
masks evaluated from the face rule inside the kernel.
"""

import ctypes
import time

import numpy as np
import torch
from scipy.constants import c as c_light
from scipy.constants import epsilon_0 as eps_0
from scipy.constants import mu_0

from . import _lib
from .field import Field, torch_dtype
from .logger import Logger
from .materials import material_lib
from .routines import RoutinesMixin

PEC_KEYS = ("electric", "pec", "pml")


def _face_zero_bits(bc_low, bc_high, test):
    bits = 0
    for a in range(3):
        if test(bc_low[a]):
            bits |= 1 << a
        if test(bc_high[a]):
            bits |= 1 << (3 + a)
    return bits


def face_mask_kji(comp, zbits, Nx, Ny, Nz, device):
    """Face mask of one component as a (Nz,Ny,Nx) bool tensor.

    Same rule as the kernel's face_mask: for component d the two faces of the
    other axes are assigned in x, y, z order and the last assignment wins
    (boundaries.py:90-103). A cell is masked when the bit of its face is set
    in ``zbits``; the faces are written as planes in that order, so a later
    face overrides an earlier one on their common edge.
    """
    keep = torch.ones((Nz, Ny, Nx), dtype=torch.bool, device=device)
    if zbits == 0:
        return keep
    planes = {0: [(1, 0, 1), (1, Ny - 1, 4), (0, 0, 2), (0, Nz - 1, 5)],
              1: [(2, 0, 0), (2, Nx - 1, 3), (0, 0, 2), (0, Nz - 1, 5)],
              2: [(2, 0, 0), (2, Nx - 1, 3), (1, 0, 1), (1, Ny - 1, 4)]}[comp]
    for axis, val, face in planes:
        keep.select(axis, val).fill_(not (zbits >> face) & 1)
    return keep


def material_codes(cols, kmax=256):
    """Dense material codes of N cells from per-cell value columns (no
    reference counterpart): the code of a cell is the rank of its tuple of
    column values in lexicographic order among the distinct tuples. Returns
    (code int64 tensor, number of codes), or (None, kmax + 1) when a column
    has more than kmax distinct values. Per-column ranks are combined in mixed
    radix and ranked again only when the radix product would pass 2^62, and
    once at the end; a column with one value (rank 0 everywhere) is skipped,
    found for all columns with one device sync.
    """
    code = torch.zeros(cols[0].numel(), dtype=torch.int64, device=cols[0].device)
    span = 1
    single = torch.stack([(c == c[:1]).all() for c in cols]).tolist()
    for col, one in zip(cols, single):
        if one:
            continue
        u, inv = torch.unique(col, return_inverse=True)
        if u.numel() > kmax:
            return None, kmax + 1
        if span * u.numel() > 2 ** 62:
            code = torch.unique(code, return_inverse=True)[1]
            span = int(code.max().item()) + 1
        code, span = code * u.numel() + inv, span * u.numel()
    code = torch.unique(code, return_inverse=True)[1]
    return code, int(code.max().item()) + 1


class SolverFIT3D(RoutinesMixin):
    def __init__(self, grid, wake=None, cfln=0.5, dt=None,
                 bc_low=["Periodic", "Periodic", "Periodic"],
                 bc_high=["Periodic", "Periodic", "Periodic"],
                 use_stl=False, use_conductors=False, use_gpu=True, use_mpi=False,
                 use_sibc=True, fmax=1e9, dtype=np.float64, n_pml=10, bg=[1.0, 1.0],
                 verbose=1, path="auto", tau_clamp=True):
        t0 = time.time()
        self.logger = Logger()
        if use_mpi:
            raise NotImplementedError("wakis_amd runs the full domain on one GPU (use_mpi=False)")
        if use_conductors:
            raise NotImplementedError("legacy conductors are not supported; use solids")
        self.verbose = verbose
        self.step_0 = True
        self.nstep = 0
        self.use_stl, self.use_gpu, self.use_mpi = use_stl, True, False
        self.use_sibc, self.fmax = use_sibc, fmax
        self.activate_abc = False
        self.activate_pml = False
        self.use_conductivity = False
        self.path = path
        lib = _lib.lib()  # fail loudly before allocating anything if the engine is missing

        self.grid = grid
        self.background = bg
        self.Nx, self.Ny, self.Nz = grid.Nx, grid.Ny, grid.Nz
        self.N = N = self.Nx * self.Ny * self.Nz
        self.dims = (self.Nx, self.Ny, self.Nz)
        self.dx, self.dy, self.dz = grid.dx, grid.dy, grid.dz
        self.x = grid.x[:-1] + self.dx / 2
        self.y = grid.y[:-1] + self.dy / 2
        self.z = grid.z[:-1] + self.dz / 2

        self.wake = wake
        if wake is not None and fmax == 1e9:
            self.fmax = wake.fmax

        self.dtype = np.dtype(dtype).type
        self.E = Field(self.Nx, self.Ny, self.Nz, dtype=self.dtype)
        self.H = Field(self.Nx, self.Ny, self.Nz, dtype=self.dtype)
        self.J = Field(self.Nx, self.Ny, self.Nz, dtype=self.dtype)
        self.device = self.E.device

        # boundaries (boundaries.py:14-188)
        self.bc_low, self.bc_high = list(bc_low), list(bc_high)
        self._setup_bcs()

        # materials (solverFIT3D.py:237-273, 416-480) on the host
        if type(bg) is str:
            bg = material_lib[bg.lower()]
        if len(bg) == 3:
            self.eps_bg, self.mu_bg, self.sigma_bg = bg[0] * eps_0, bg[1] * mu_0, bg[2]
            self.use_conductivity = True
        else:
            self.eps_bg, self.mu_bg, self.sigma_bg = bg[0] * eps_0, bg[1] * mu_0, 0.0
        ieps = np.ones(3 * N, dtype=self.dtype) * (1.0 / self.eps_bg)
        imu = np.ones(3 * N, dtype=self.dtype) * (1.0 / self.mu_bg)
        sigma = np.ones(3 * N, dtype=self.dtype) * self.sigma_bg
        self._h = {"ieps": ieps, "imu": imu, "sigma": sigma}
        if self.use_stl:
            self._apply_stl_materials()
        if self.activate_pml:
            self.n_pml = n_pml
            self._initialize_PML()

        # timestep (solverFIT3D.py:283-310)
        self.cfln = cfln
        if dt is None:
            dt = cfln / (c_light * np.sqrt(1 / np.min(self.grid.dx) ** 2
                                           + 1 / np.min(self.grid.dy) ** 2
                                           + 1 / np.min(self.grid.dz) ** 2))
        self.dt = self.dtype(dt)
        # tau_clamp=False keeps dt independent of the conductivities (needed when
        # sigma is a parameter to differentiate); the caller then keeps
        # dt*sigma/eps well below 2 for stability
        if self.use_conductivity and tau_clamp:
            h = self._h
            m = np.logical_and(h["sigma"] != 0, h["ieps"] != 0)
            self.tau = (1 / h["ieps"][m]) / h["sigma"][m]
            if self.tau.size and self.dt > self.tau.min():
                self.dt = self.tau.min()

        # material Fields on the device
        self.ieps = Field(self.Nx, self.Ny, self.Nz, dtype=self.dtype)
        self.imu = Field(self.Nx, self.Ny, self.Nz, dtype=self.dtype)
        self.sigma = Field(self.Nx, self.Ny, self.Nz, dtype=self.dtype)
        self.ieps.fromarray(self._h["ieps"])
        self.imu.fromarray(self._h["imu"])
        self.sigma.fromarray(self._h["sigma"])
        del self._h

        if path not in ("auto", "general", "fused"):
            raise ValueError("path must be 'auto', 'general' or 'fused'")
        self._create_handle(lib)
        self._fold_count = 0
        self._wrap_key = None
        self._E2 = self._H2 = None
        self.fused_kz = 0  # z-planes per workgroup of the fused kernel (0: automatic)
        self.fused_variant = 0  # 0 auto, 1 scalar, 2/3/7/8: vectorised, rows per workgroup
        self.fused_cells_per_lane = 0  # x-vectorised kernel: 0 auto, 1, 2 (FP32), 16/itemsize
        # whole-row workgroups: N >= 1 rows; 0 row segments (tune_fused() measures and
        # sets -1 = segments or N), -1 row segments
        self.fused_full_rows = 0
        # two leapfrog steps per launch where that kernel applies (the wakefield
        # step with material codes, FP32 or FP64, Nx <= 256): 0 off, 1 on (8 rows
        # per workgroup, FP64 2), 4 or 8 rows (FP32); bit-identical to one step
        # per launch.
        # tune_fused() sets 0 when one step per launch measures faster
        self.fused_two_step = 1
        # {full_rows: ms per step} of the last tune_fused(); -2: two steps per launch
        self.fused_tuning = None
        # energy(): keep E^n and H^(n-1/2) on the exact path too (one copy per step)
        self.keep_previous_fields = False
        self._prev_key = None
        # fused kernel without in-kernel boundary masks when the masked values are zero
        self.skip_masks = True
        self._mz_key = self._mz_ok = None
        self.use_material_code = True  # fused path: uint8 code + coefficient table
        # more than 256 materials: code 255 for the rarest cells, read from the
        # arrays, when they are at most this fraction of the cells
        self.use_code_escape = True
        self.code_escape_max = 0.05
        self._code_escape = False
        self._lut_key = None
        self._code = self._mat = self._metric2d = None
        self._nmat = 0
        self.update_tensors("all", _initial=True)
        self._beam_line = None
        self._J_clean_version = self.J.version
        self._sigmaE_versions = (self.E.version, self.J.version)  # J == sigma*E (all zero)
        self.last_path = None
        self.last_two_step = False  # the last step call ran two steps per launch
        self.solverInitializationTime = time.time() - t0
        # the reference's solver log keys (solverFIT3D.py:149, 171-176, 234, 281, 299, 350-354)
        self.update_logger(["use_gpu", "use_mpi", "grid", "background", "bc_low", "bc_high"])
        if self.activate_pml:
            self.update_logger(["n_pml"])
        self.update_logger(["dt", "solverInitializationTime"])
        if self.wake is not None:
            self.logger.wakeSolver = self.wake.logger.wakeSolver
        if verbose:
            print(f"Total solver initialization time: {self.solverInitializationTime} s")

    # ------------------------------------------------------------------ BCs
    def _setup_bcs(self):
        lo, hi = self.bc_low, self.bc_high
        self.pec_active = any(b.lower() in PEC_KEYS for b in lo) or any(
            b.lower() in PEC_KEYS for b in hi)
        self.pmc_active = any(b.lower() in ("magnetic", "pmc") for b in lo) or any(
            b.lower() in ("magnetic", "pmc") for b in hi)
        self.pec_zero = _face_zero_bits(lo, hi, lambda b: b.lower() in PEC_KEYS) \
            if self.pec_active else 0
        self.pmc_zero = 0
        if self.pmc_active:
            self.pmc_zero = self.pec_zero | _face_zero_bits(
                lo, hi, lambda b: b.lower() == "magnetic" or b == "pmc")
        # scipy's csr_matmat reverses a row's column order once per diagonal
        # product; the M1 chain ends descending iff exactly one mask was applied
        self.m1_descending = int(self.pec_active) ^ int(self.pmc_active)
        # metric copies (boundaries.py:35-70, 152-181), in the reference's order
        self._metric_edits = []
        if any(b.lower() == "periodic" for b in lo):
            for a in range(3):
                if lo[a].lower() == "periodic" and hi[a].lower() == "periodic":
                    self._metric_edits.append(a)
        if any(b.lower() == "abc" for b in lo) or any(b.lower() == "abc" for b in hi):
            for a in range(3):
                if hi[a].lower() == "abc":
                    self._metric_edits.append(a)
            self.activate_abc = True
        if any(b.lower() == "pml" for b in lo) or any(b.lower() == "pml" for b in hi):
            self.activate_pml = True
            self.use_conductivity = True
        g = self.grid
        self._L1 = [np.array(g.dx, dtype=np.float64), np.array(g.dy, dtype=np.float64),
                    np.array(g.dz, dtype=np.float64)]
        self._tL1_base = [np.array(g.tdx, dtype=np.float64), np.array(g.tdy, dtype=np.float64),
                          np.array(g.tdz, dtype=np.float64)]
        self._tL1 = [v.copy() for v in self._tL1_base]
        for a in self._metric_edits:
            self._tL1[a][-1] = self._L1[a][0]

    def _create_handle(self, lib):
        self._vec_keep = [np.ascontiguousarray(v, dtype=np.float64)
                          for v in self._L1 + self._tL1]
        P = ctypes.POINTER(ctypes.c_double)
        ptrs = [v.ctypes.data_as(P) for v in self._vec_keep]
        g = _lib.WfGrid(self.Nx, self.Ny, self.Nz,
                        _lib.WF_F64 if self.dtype == np.float64 else _lib.WF_F32,
                        self.pec_zero, self.pmc_zero, self.m1_descending, *ptrs)
        h = ctypes.c_void_p()
        _lib.check(lib.wf_create(ctypes.byref(g), ctypes.byref(h)), "wf_create")
        self._handle = h

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            try:
                _lib.lib().wf_destroy(h)
            except Exception:
                pass
            self._handle = None

    # ------------------------------------------------------------- materials
    def _comp_h(self, arr, d):
        c = {"x": 0, "y": 1, "z": 2}[d]
        return np.reshape(arr[c * self.N:(c + 1) * self.N], self.dims, order="F")

    def _apply_stl_materials(self):
        """solverFIT3D.py:416-480 with analytic cell masks; SIBC keeps the solid's
        surface cells (``GridFIT3D._mark_cells_in_surface``) and gives them the surface impedance as
        conductivity and (absolute) permittivity (solverFIT3D.py:455-466)."""
        h = self._h
        for key in self.grid.stl_solids.keys():
            mask = np.asarray(self.grid.grid[key]).astype(int)
            mat = self.grid.stl_materials[key]
            if isinstance(mat, str):
                mat = material_lib[mat.lower()]
            eps, mu = mat[0] * eps_0, mat[1] * mu_0

            def scaled_add(arr, scale3d):
                tmp = np.zeros_like(arr)
                for d in ("x", "y", "z"):
                    tmp[{"x": 0, "y": 1, "z": 2}[d] * self.N:][:self.N] = np.reshape(
                        self._comp_h(arr, d) * scale3d, self.N, order="F")
                return arr + tmp

            def mat_add(arr, m3d):
                out = np.empty_like(arr)
                for d in ("x", "y", "z"):
                    out[{"x": 0, "y": 1, "z": 2}[d] * self.N:][:self.N] = np.reshape(
                        self._comp_h(arr, d) + m3d, self.N, order="F")
                return out

            if len(mat) == 3:
                sig = mat[2]
                dn = np.sqrt(2) * min(self.dx.min(), self.dy.min(), self.dz.min())
                sigma_max = 10 / (np.pi * self.fmax * mu * dn ** 2)
                if self.use_sibc and sig > sigma_max:
                    mask = self.grid._mark_cells_in_surface(key).astype(int)
                    imp = np.sqrt(np.pi * self.fmax * mu / sig)
                    sig = 1 / imp  # SIBC surface conductivity
                    eps = 1 / imp
                h["sigma"] = scaled_add(h["sigma"], -1.0 * mask)
                h["sigma"] = mat_add(h["sigma"], mask * sig)
                self.use_conductivity = True
            elif self.sigma_bg > 0.0:
                h["sigma"] = scaled_add(h["sigma"], -1.0 * mask)
            h["ieps"] = scaled_add(h["ieps"], -1.0 * mask)
            h["imu"] = scaled_add(h["imu"], -1.0 * mask)
            h["ieps"] = mat_add(h["ieps"], mask * 1.0 / eps)
            h["imu"] = mat_add(h["imu"], mask * 1.0 / mu)

    def _initialize_PML(self):
        """Graded PML conductivity (boundaries.py:190-311): sigma only, no aux field."""
        h = self._h
        n = self.n_pml
        lo, hi = 5.0e-3, 10.0
        self.pml_lo, self.pml_hi, self.pml_func, self.pml_eps_r = lo, hi, np.geomspace, 1.0
        for side in ("low", "high"):
            for axis in range(3):
                bc = self.bc_low[axis] if side == "low" else self.bc_high[axis]
                if bc.lower() != "pml":
                    continue
                prof = np.zeros(self.dims[axis])
                if side == "low":
                    prof[0:n] = np.geomspace(hi, lo, n)
                else:
                    prof[-n:] = np.geomspace(lo, hi, n)
                for d in ("x", "y", "z"):
                    centre = [self.Nx // 2, self.Ny // 2, self.Nz // 2]
                    centre[axis] = (n + 1) if side == "low" else -(n + 1)
                    ieps0 = self._comp_h(h["ieps"], d)[tuple(centre)]
                    sig0 = self._comp_h(h["sigma"], d)[tuple(centre)]
                    mult = 1 if sig0 < 1 else sig0
                    for m in range(n):
                        idx = m if side == "low" else -(m + 1)
                        sl = [slice(None)] * 3
                        sl[axis] = idx
                        self._comp_h(h["ieps"], d)[tuple(sl)] = ieps0
                        self._comp_h(h["sigma"], d)[tuple(sl)] = sig0 + mult * prof[idx]

    # --------------------------------------------------------- coefficients
    def update_tensors(self, tensor="all", _initial=False):
        """Re-fold the stencil coefficients from ieps/imu (solverFIT3D.py:356-414).

        Unlike the reference, this stays valid after the first step.
        """
        if self.verbose and not _initial:
            print(f'Re-computing tensor "{tensor}"...')
        Nx, Ny, Nz, N = self.Nx, self.Ny, self.Nz, self.N
        dev, tdt = self.device, torch_dtype(self.dtype)
        f64 = torch.float64
        L = [torch.as_tensor(v, dtype=f64, device=dev) for v in self._L1]
        tL = [torch.as_tensor(v, dtype=f64, device=dev) for v in self._tL1_base]
        shape = [(1, 1, Nx), (1, Ny, 1), (Nz, 1, 1)]
        L = [L[a].view(shape[a]) for a in range(3)]
        tL = [tL[a].view(shape[a]) for a in range(3)]
        if not hasattr(self, "_a"):
            self._a = torch.empty(3 * N, dtype=tdt, device=dev)
            self._b = torch.empty(3 * N, dtype=tdt, device=dev)
        others = {0: (1, 2), 1: (0, 2), 2: (0, 1)}
        self._iA2d, self._itA2d = [None] * 3, [None] * 3
        plane = {0: (slice(None), slice(None), 0), 1: (slice(None), 0, slice(None)),
                 2: (0, slice(None), slice(None))}
        for c in range(3):
            u, w = others[c]
            iA = torch.broadcast_to(1.0 / (L[u] * L[w]), (Nz, Ny, Nx))
            self._iA2d[c] = iA[plane[c]].to(tdt).contiguous().reshape(-1)
            if tensor in ("all", "imu"):
                a_c = iA.to(tdt) * self.imu._block(c).view(Nz, Ny, Nx)
                if self.pmc_zero:
                    a_c = a_c * face_mask_kji(c, self.pmc_zero, Nx, Ny, Nz, dev).to(tdt)
                self._a[c * N:(c + 1) * N].copy_(a_c.reshape(-1))
            aux = torch.broadcast_to(tL[u] * tL[w], (Nz, Ny, Nx))
            itA = torch.where(aux != 0, 1.0 / torch.where(aux != 0, aux, 1.0),
                              torch.zeros((), dtype=f64, device=dev)).clone()
            for e in self._metric_edits:  # itA[-1 along e] = iA[0 along e]
                if e == c:
                    continue
                ax = 2 - e  # (Nz,Ny,Nx) axis of grid axis e
                itA.select(ax, itA.shape[ax] - 1).copy_(iA.select(ax, 0))
            self._itA2d[c] = itA[plane[c]].to(tdt).contiguous().reshape(-1)
            if tensor in ("all", "ieps"):
                b_c = self.ieps._block(c).view(Nz, Ny, Nx) * itA.to(tdt)
                if self.pec_zero:
                    b_c = b_c * face_mask_kji(c, self.pec_zero, Nx, Ny, Nz, dev).to(tdt)
                self._b[c * N:(c + 1) * N].copy_(b_c.reshape(-1))
        self._metric2d_all = torch.cat(self._iA2d + self._itA2d).contiguous()
        torch.cuda.synchronize(dev)
        self._fold_count = getattr(self, "_fold_count", 0) + 1
        self._fold_versions = (self.ieps.version, self.imu.version)
        if not _initial:
            self.step_0 = False  # as the reference (solverFIT3D.py:414): no ghost zeroing after

    # ------------------------------------------------------------ stepping
    def _set_ghosts_to_0(self):
        _lib.check(_lib.lib().wf_zero_ghosts(self._handle, _lib.ptr(self.E.array),
                                             _lib.ptr(self.H.array), _lib.stream_ptr()),
                   "wf_zero_ghosts")

    def _compute_f64(self):
        return (self.dtype == np.float32
                and np.result_type(self.dt, np.float32) == np.dtype(np.float64))

    def _jmode(self):
        if self.use_conductivity or self.J.version != self._J_clean_version:
            return _lib.WF_J_FULL
        if self._beam_line is not None:
            return _lib.WF_J_LINE
        return _lib.WF_J_NONE

    def _write_beam_line(self, ixs, iys, values):
        """J[ixs, iys, :, 'z'] = values without marking J as user-written."""
        line = (int(ixs), int(iys))
        v = self.J.version
        self.J[ixs, iys, :, "z"] = values
        if self._beam_line is None or self._beam_line == line:
            self.J.version = v  # J stays zero outside this one line
            if v == self._J_clean_version:
                self._beam_line = line

    def _step_args(self, nsteps):
        jm = self._jmode()
        a = _lib.WfStepArgs()
        a.E, a.H, a.J = _lib.ptr(self.E.array), _lib.ptr(self.H.array), _lib.ptr(self.J.array)
        a.a, a.b = _lib.ptr(self._a), _lib.ptr(self._b)
        a.ieps, a.sigma = _lib.ptr(self.ieps.array), _lib.ptr(self.sigma.array)
        a.dt = float(self.dt)
        a.nsteps = nsteps
        a.jmode = jm
        a.conductive = 1 if self.use_conductivity else 0
        bl = self._beam_line if self._beam_line is not None else (-1, -1)
        a.ixs, a.iys = bl
        a.path = 0
        a.compute_f64 = 1 if self._compute_f64() else 0
        return a

    # -------------------------------------------------------- fast path
    def _sigmaE_holds(self):
        """J == fl(sigma*E) everywhere except the registered beam line."""
        return (self.E.version, self.J.version) == self._sigmaE_versions

    def _choose_fused(self, args):
        if self.path == "general":
            return False
        why = None
        if self.use_conductivity:
            if not self._sigmaE_holds():
                why = "J is not sigma*E (user write since the last conductive step)"
        elif args.jmode == _lib.WF_J_FULL:
            why = "J was written outside the beam line"
        if why is None and args.compute_f64:
            why = "FP64 arithmetic on FP32 storage runs on the exact general path"
        if why is None and not self._wraps_dead():
            why = "flat-index wrap terms are live for this boundary/material set-up"
        if why is not None:
            if self.path == "fused":
                raise RuntimeError(f"fused path not applicable: {why}")
            return False
        return True

    def _material_table(self):
        """Compressed coefficients for the fused kernel, if the set-up allows.

        Per cell a uint8 material code into a table of (imu, ieps, sigma) per
        component (<= 256 materials), plus 2-D metric tables iA_c, itA_c over
        the transverse axes of each component. The kernel recomputes
        a = fl(iA*imu) and b = fl(ieps*itA) with the row masks; this method
        checks that the recomputation reproduces the folded arrays exactly
        before enabling it. Valid until coefficients are re-folded or
        ieps/imu/sigma change.
        """
        key = (self._fold_count, self.ieps.version, self.imu.version, self.sigma.version)
        if key == self._lut_key:
            return self._code is not None
        self._lut_key = key
        self._code = self._mat = self._metric2d = None
        self._nmat = 0
        self._code_escape = False
        if (self.ieps.version, self.imu.version) != self._fold_versions:
            return False  # materials edited without update_tensors: use the arrays
        N, (Nx, Ny, Nz), dev = self.N, self.dims, self.device
        cols = ([self.imu.array[c * N:(c + 1) * N] for c in range(3)]
                + [self.ieps.array[c * N:(c + 1) * N] for c in range(3)]
                + [self.sigma.array[c * N:(c + 1) * N] for c in range(3)])
        code, K = material_codes(cols)
        esc = None  # cells whose coefficients come from the arrays (code 255)
        if K <= 256:
            rep = torch.empty(K, dtype=torch.int64, device=dev)
            rep.scatter_(0, code, torch.arange(N, device=dev))
            mat = torch.stack([col[rep] for col in cols], dim=1).contiguous()
        else:
            # more than 256 materials: the 255 most frequent get codes, the rest
            # (e.g. a block of cells with their own parameters) the escape code,
            # if they are few enough for the per-cell reads to pay off
            if (not self.use_code_escape or self.fused_variant == 1
                    or 3 * N * np.dtype(self.dtype).itemsize >= 2 ** 32):
                return False
            u, inv, cnt = torch.unique(torch.stack(cols, dim=1), dim=0, return_inverse=True,
                                       return_counts=True)
            order = torch.argsort(cnt, descending=True, stable=True)
            # materials of a single cell are escaped too: a short table leaves
            # room in LDS (the FP64 reverse kernel holds it beside its planes)
            K = max(1, min(255, int((cnt >= 2).sum().item())))
            if int(cnt[order[K:]].sum().item()) > self.code_escape_max * N:
                return False
            rank = torch.empty_like(order)
            rank[order] = torch.arange(order.numel(), device=dev)
            code = rank[inv].clamp(max=K)
            code[code == K] = 255
            esc = code == 255
            mat = u[order[:K]].contiguous()
        # 2-D metric tables with the row masks folded in (iA zero on PMC-masked
        # faces, itA zero on PEC-masked edges): a mask of component c depends
        # only on the two transverse indices, like the table
        plane = {0: (slice(None), slice(None), 0), 1: (slice(None), 0, slice(None)),
                 2: (0, slice(None), slice(None))}
        iA, itA = [], []
        for c in range(3):
            ia, ita = self._iA2d[c], self._itA2d[c]
            zero = torch.zeros((), dtype=ia.dtype, device=dev)
            if self.pmc_zero:
                m2 = face_mask_kji(c, self.pmc_zero, Nx, Ny, Nz, dev)[plane[c]].reshape(-1)
                ia = torch.where(m2, ia, zero)
            if self.pec_zero:
                m2 = face_mask_kji(c, self.pec_zero, Nx, Ny, Nz, dev)[plane[c]].reshape(-1)
                ita = torch.where(m2, ita, zero)
            iA.append(ia)
            itA.append(ita)
        metric2d = torch.cat(iA + itA).contiguous()
        # exactness check against the folded arrays
        shp = {0: (Nz, Ny, 1), 1: (Nz, 1, Nx), 2: (1, Ny, Nx)}
        c3 = code.clamp(max=K - 1).view(Nz, Ny, Nx)
        for c in range(3):
            a_c = (iA[c].view(shp[c]) * mat[:, c][c3]).reshape(-1)
            b_c = (mat[:, 3 + c][c3] * itA[c].view(shp[c])).reshape(-1)
            a_ref, b_ref = self._a[c * N:(c + 1) * N], self._b[c * N:(c + 1) * N]
            if esc is not None:  # escaped cells read the arrays themselves
                a_c, b_c = torch.where(esc, a_ref, a_c), torch.where(esc, b_ref, b_c)
            if not (torch.equal(a_c, a_ref) and torch.equal(b_c, b_ref)):
                return False
        self._code = code.to(torch.uint8).contiguous()
        self._mat, self._nmat, self._metric2d = mat, K, metric2d
        self._code_escape = esc is not None
        return True

    def _masked_zero(self):
        """True if E is zero on every PEC-masked edge, H on every PMC-masked
        face, and the beam line is not on a masked edge.

        Those values then stay zero (their rows have zero coefficients and no
        source), so the fast kernel may skip the boundary masks.
        """
        key = (self.E.version, self.H.version, self.J.version, self._fold_count,
               self._beam_line)
        if key == self._mz_key:
            return self._mz_ok
        Nx, Ny, Nz, N, dev = self.Nx, self.Ny, self.Nz, self.N, self.device
        ok = True
        for c in range(3):
            if self.pec_zero:
                keep = face_mask_kji(c, self.pec_zero, Nx, Ny, Nz, dev)
                Ec = self.E.array[c * N:(c + 1) * N].view(Nz, Ny, Nx)
                ok = ok and not bool((Ec[~keep] != 0).any().item())
                if c == 2 and self._beam_line is not None:
                    ixs, iys = self._beam_line
                    ok = ok and bool(keep[:, iys, ixs].all().item())
            if self.pmc_zero:
                keep = face_mask_kji(c, self.pmc_zero, Nx, Ny, Nz, dev)
                Hc = self.H.array[c * N:(c + 1) * N].view(Nz, Ny, Nx)
                ok = ok and not bool((Hc[~keep] != 0).any().item())
        self._mz_key, self._mz_ok = key, ok
        return ok

    def _wraps_dead(self):
        """True if every flat-index wrap term of M1/M2 contributes exactly zero.

        The reference operator couples i=Nx-1 to (0, j+1, k) and j=Ny-1 to
        (i, 0, k+1) (and the transposes). A term is dead if its weight is zero
        (row coefficient or column mask) or its column is frozen at zero:
        E with b=0, E=0 and no live source (ieps=0 or J=0, not on the beam
        line); H with a=0 and H=0. Frozen columns never change, so the result
        holds until E/H/J are written from outside or coefficients re-folded.
        """
        key = (self.E.version, self.H.version, self.J.version, self._fold_count,
               self._beam_line)
        if key == self._wrap_key:
            return self._wrap_ok
        self._wrap_key, self._wrap_ok = key, self._wraps_dead_now()
        return self._wrap_ok

    def _wraps_dead_now(self):
        """The test of ``_wraps_dead`` on the boundary planes. Each wrap term
        pairs a row with the column one flat index away (+-1 across a row end,
        +-Nx across the last row of a plane), so rows and columns are two views
        of the same boundary plane shifted by one element."""
        Nx, Ny, Nz = self.Nx, self.Ny, self.Nz
        dev = self.device
        a3 = self._a.view(3, Nz, Ny, Nx)
        b3 = self._b.view(3, Nz, Ny, Nx)
        E3 = self.E.array.view(3, Nz, Ny, Nx)
        H3 = self.H.array.view(3, Nz, Ny, Nx)
        J3 = self.J.array.view(3, Nz, Ny, Nx)
        ie3 = self.ieps.array.view(3, Nz, Ny, Nx)
        bl = self._beam_line
        beam = None
        if bl is not None:
            beam = torch.zeros((Nz, Ny, Nx), dtype=torch.bool, device=dev)
            beam[:, bl[1], bl[0]] = True

        def keep(zbits, c):  # columns not masked by the face masks
            if zbits == 0:
                return torch.ones((Nz, Ny, Nx), dtype=torch.bool, device=dev)
            return face_mask_kji(c, zbits, Nx, Ny, Nz, dev)

        def frozen_E(c, sl):  # E column frozen at zero (no coefficient, no source)
            fz = (b3[c][sl] == 0) & (E3[c][sl] == 0) & ((ie3[c][sl] == 0) | (J3[c][sl] == 0))
            if beam is not None and c == 2:
                fz = fz & ~beam[sl]
            return fz

        def frozen_H(c, sl):
            return (a3[c][sl] == 0) & (H3[c][sl] == 0)

        tL = [torch.as_tensor(self._tL1[0], device=dev).view(1, 1, Nx).expand(Nz, Ny, Nx),
              torch.as_tensor(self._tL1[1], device=dev).view(1, Ny, 1).expand(Nz, Ny, Nx),
              torch.as_tensor(self._tL1[2], device=dev).view(Nz, 1, 1).expand(Nz, Ny, Nx)]
        X0 = (slice(None), slice(None), 0)        # plane i = 0, flattened in (k, j)
        X1 = (slice(None), slice(None), Nx - 1)   # plane i = Nx-1
        Ylo = (slice(1, Nz), 0, slice(None))          # row j = 0 of planes 1..Nz-1
        Yhi = (slice(0, Nz - 1), Ny - 1, slice(None))  # row j = Ny-1 of planes 0..Nz-2
        terms = []
        # (1) H_y, H_z rows at i = Nx-1 take E_z, E_y at flat index + 1
        #     ((0, j+1, k), or (0, 0, k+1) after the last row of a plane)
        for rc, cc in ((1, 2), (2, 1)):
            w = (a3[rc][X1] != 0).reshape(-1)[:-1] & keep(self.pec_zero, cc)[X0].reshape(-1)[1:]
            terms.append(w & ~frozen_E(cc, X0).reshape(-1)[1:])
        # (2) H_x, H_z rows at j = Ny-1 take E_z, E_x at flat index + Nx: (i, 0, k+1)
        for rc, cc in ((0, 2), (2, 0)):
            w = (a3[rc][Yhi] != 0) & keep(self.pec_zero, cc)[Ylo]
            terms.append((w & ~frozen_E(cc, Ylo)).reshape(-1))
        # (3) E_y, E_z rows at i = 0 take H_z, H_y at flat index - 1 (weight tLz, tLy
        #     of the column)
        for rc, cc, ax in ((1, 2, 2), (2, 1, 1)):
            w = ((b3[rc][X0] != 0).reshape(-1)[1:] & (tL[ax][X1] != 0).reshape(-1)[:-1]
                 & keep(self.pmc_zero, cc)[X1].reshape(-1)[:-1])
            terms.append(w & ~frozen_H(cc, X1).reshape(-1)[:-1])
        # (4) E_x, E_z rows at j = 0 take H_z, H_x at flat index - Nx: (i, Ny-1, k-1)
        #     (weight tLz, tLx of the column)
        for rc, cc, ax in ((0, 2, 2), (2, 0, 0)):
            w = (b3[rc][Ylo] != 0) & (tL[ax][Yhi] != 0) & keep(self.pmc_zero, cc)[Yhi]
            terms.append((w & ~frozen_H(cc, Yhi)).reshape(-1))
        bad = torch.stack([t.any() for t in terms]).any()
        return not bool(bad.item())

    def _prepare_fused(self, args):
        """Choose the path for a batch and fill the fused-path fields of args.

        Returns True when the batch runs on the fused path (E2/H2 ping-pong,
        compressed coefficients when exact, no in-kernel masks when the masked
        values are zero), False for the exact general path.
        """
        fused = self._choose_fused(args)
        if not fused:
            return False
        if self._E2 is None:
            self._E2 = torch.empty_like(self.E.array)
            self._H2 = torch.empty_like(self.H.array)
        args.path = 1
        args.E2, args.H2 = _lib.ptr(self._E2), _lib.ptr(self._H2)
        args.kz = int(self.fused_kz)
        args.variant = int(self.fused_variant)
        args.cells_per_lane = int(self.fused_cells_per_lane)
        args.full_rows = int(self.fused_full_rows)
        args.two_step = int(self.fused_two_step)
        args.zero_masked = 1 if (self.skip_masks and self._masked_zero()) else 0
        if (self.use_material_code and self._material_table()
                and not (self._code_escape and self.fused_variant == 1)):
            args.code, args.mat, args.metric2d = (_lib.ptr(self._code), _lib.ptr(self._mat),
                                                  _lib.ptr(self._metric2d))
            args.nmat = self._nmat
            args.code_escape = int(self._code_escape)
        if args.jmode == _lib.WF_J_FULL and not args.conductive:
            raise RuntimeError("internal: fused path without a J mode")
        if args.conductive:
            args.jmode = _lib.WF_J_FULL
        return True

    def _finish_run(self, fused, swapped, kept=False, pair_last=False):
        """Bookkeeping after a batch: buffer swap and the invariants it keeps.
        ``pair_last``: the batch ended with a two-step launch, whose inputs in
        the spare buffers are the fields two steps back."""
        self.last_path = "fused" if fused else "general"
        if fused and swapped:
            self.E.array, self._E2 = self._E2, self.E.array
            self.H.array, self._H2 = self._H2, self.H.array
        # the spare buffers now hold E^n, H^(n-1/2) of the last step (fused: the
        # ping-pong inputs; exact path: the copies made before the step)
        self._prev_key = ((self.E.version, self.H.version)
                          if ((fused and not pair_last) or kept) else None)
        if self.use_conductivity:
            self._sigmaE_versions = (self.E.version, self.J.version)

    def _load_run_state(self, E, H, J, box=None):
        """Restore fields saved earlier in a run of this solver (checkpoints).

        With ``box`` (slices of z, y, x), E, H, J hold only those cells and only
        they are written; the cells outside keep what they hold (a recompute
        confined to the box's domain of dependence never carries them into the
        cells it keeps). The fast-path invariants checked for that run (J =
        sigma E, zero-valued masked edges and faces, dead flat-index wrap
        terms) hold for every state the run produced, so their cached results
        are carried over to the restored state instead of being checked again.
        """
        mz = (self._mz_key is not None, self._mz_ok)
        wr = (self._wrap_key is not None, self._wrap_ok)
        Nx, Ny, Nz = self.dims
        for F, X in ((self.E, E), (self.H, H), (self.J, J)):
            if box is None:
                F.array.copy_(X)
            else:
                F.array.view(3, Nz, Ny, Nx)[(slice(None),) + tuple(box)].copy_(X)
            F.version += 1
        if self.use_conductivity:
            self._sigmaE_versions = (self.E.version, self.J.version)
        key = (self.E.version, self.H.version, self.J.version, self._fold_count,
               self._beam_line)
        if mz[0]:
            self._mz_key, self._mz_ok = key, mz[1]
        if wr[0]:
            self._wrap_key, self._wrap_ok = key, wr[1]
        self._prev_key = None

    def _run(self, args):
        """Run one wf_step batch on the fast or the exact path."""
        fused = self._prepare_fused(args)
        keep = (not fused) and self.keep_previous_fields and args.nsteps == 1
        if keep:
            if self._E2 is None:
                self._E2 = torch.empty_like(self.E.array)
                self._H2 = torch.empty_like(self.H.array)
            self._E2.copy_(self.E.array)
            self._H2.copy_(self.H.array)
        _lib.check(_lib.lib().wf_step(self._handle, ctypes.byref(args), _lib.stream_ptr()),
                   "wf_step")
        # two steps per launch with an even count: the last launch was a pair
        self.last_two_step = bool(args.two_step_used)
        self._finish_run(fused, bool(args.result_swapped), keep,
                         pair_last=self.last_two_step and args.nsteps % 2 == 0)

    def tune_fused(self, make_args=None, steps=3):
        """Choose the fused kernel's workgroup shape by timing it (no reference
        counterpart). Candidates: the 64-lane row segments, whole-row
        workgroups with the most rows for each wave count that fits and, with
        ``fused_two_step`` on and where that kernel applies, two steps per
        launch (key -2 of ``fused_tuning``). Each runs one untimed and `steps`
        timed steps of ``make_args(k)`` (default: plain steps; two and an even
        count with two steps per launch on), in two rounds of which the faster
        time counts; E, H and J are restored
        afterwards, and every shape gives bitwise the same fields. Sets
        ``fused_full_rows`` to the fastest (-1 for both kinds of row segments),
        ``fused_two_step`` to 0 when one step per launch of row segments beat
        two, and returns ``fused_full_rows`` (None when the fused path would
        not run or the choice was fixed by the caller)."""
        if (self.fused_full_rows != 0 or self.fused_variant != 0 or self.fused_cells_per_lane
                or self.path == "general"):
            return None
        make = make_args if make_args is not None else self._step_args
        if not self._choose_fused(make(steps)):
            return None
        # with two steps per launch, warm up and time whole launches (even counts)
        warm = 2 if self.fused_two_step else 1
        if self.fused_two_step:
            steps += steps % 2
        size = np.dtype(self.dtype).itemsize
        vn = 16 // size
        if self.Nx % vn:
            return None
        q = self.Nx // vn
        table = (self._nmat * 9 * size) if self.use_material_code else 0
        by_waves = {}
        for T in range(1, self.Ny + 1):
            thr = -(-(T + 1) * q // 64) * 64
            if thr > 512 or (6 * T + 9) * (self.Nx + 2 * vn) * size + table > 65536:
                break
            by_waves[thr // 64] = T
        if not by_waves:
            return None
        two = int(self.fused_two_step)
        cands = ([-2] if two else []) + [-1] + sorted(by_waves.values())
        saved = [F.array.clone() for F in (self.E, self.H, self.J)]
        held = self.use_conductivity and self._sigmaE_holds()
        jclean, line = self.J.version == self._J_clean_version, self._beam_line
        key0 = (self.E.version, self.H.version, self.J.version, self._fold_count, line)
        mz = self._mz_ok if self._mz_key == key0 else None  # checks of the kept state
        wr = self._wrap_ok if self._wrap_key == key0 else None
        stream = torch.cuda.current_stream(self.device)
        times = {}
        try:
            # two rounds over the candidates, the faster time of each kept: one
            # trial of a few short steps is noisy
            for rnd in range(2):
                for c in cands:
                    if rnd and c not in times:
                        continue
                    self.fused_full_rows = max(c, -1)
                    self.fused_two_step = two if c == -2 else 0
                    try:
                        self._run(make(warm))
                    except ValueError:
                        continue
                    if c == -2 and not self.last_two_step:
                        continue  # two steps per launch do not apply: -1 times these shapes
                    e0 = torch.cuda.Event(enable_timing=True)
                    e1 = torch.cuda.Event(enable_timing=True)
                    e0.record(stream)
                    self._run(make(steps))
                    e1.record(stream)
                    e1.synchronize()
                    t = e0.elapsed_time(e1) / steps
                    times[c] = min(times.get(c, t), t)
        finally:
            self.fused_two_step = two
            for F, X in zip((self.E, self.H, self.J), saved):
                F.array.copy_(X)
                F.version += 1
            self._prev_key = None
            if jclean:  # J is again zero off the beam line
                self._J_clean_version = self.J.version
            self._beam_line = line
            # the restored state is the one checked before: carry the results over
            key1 = (self.E.version, self.H.version, self.J.version, self._fold_count, line)
            self._mz_key, self._mz_ok = (key1, mz) if mz is not None else (None, None)
            self._wrap_key, self._wrap_ok = (key1, wr) if wr is not None else (None, None)
            if self.use_conductivity:
                self._sigmaE_versions = ((self.E.version, self.J.version) if held
                                         else (-1, -1))
        best = min(times, key=times.get) if times else 0
        if best == -1 and -2 in times:
            self.fused_two_step = 0  # one step per launch measured faster
        self.fused_full_rows = max(best, -1)
        self.fused_tuning = times
        return self.fused_full_rows

    def energy(self, parts=False):
        """Discrete FIT energy of the last step, in joules (no reference
        counterpart; docs/physicsguide.md:181-182 sketches an energy density).

        W^n = 1/2 sum eps*tA*L*(E^n)^2 + 1/2 sum mu*A*tL*H^(n-1/2)*H^(n+1/2)
        over the live edges/faces: the quantity the lossless leapfrog conserves
        exactly (to rounding); with conductivity it decreases. Needs the fields
        before the last step: available after a fused step, or on the exact
        path with ``keep_previous_fields = True`` and single steps. FP64
        compensated reduction on the device. ``parts=True`` returns
        (electric, magnetic).
        """
        if self._prev_key is None or self._prev_key != (self.E.version, self.H.version):
            raise RuntimeError("energy(): the fields before the last step are not kept; step "
                               "again (fused path) or set keep_previous_fields = True")
        out = torch.zeros(2, dtype=torch.float64, device=self.device)
        _lib.check(_lib.lib().wf_energy(
            self._handle, _lib.ptr(self._E2), _lib.ptr(self._H2), _lib.ptr(self.H.array),
            _lib.ptr(self.ieps.array), _lib.ptr(self.imu.array), _lib.ptr(self._metric2d_all),
            _lib.ptr(out), _lib.stream_ptr()), "wf_energy")
        uE, uH = out.tolist()
        return (uE, uH) if parts else uE + uH

    def count_nonfinite(self, fields=("E", "H", "J")):
        """Number of Inf/NaN values in the named fields, counted on the device
        (no reference counterpart: the reference steps on with non-finite
        fields). Returns a dict {name: count}."""
        out = torch.zeros(len(fields), dtype=torch.int64, device=self.device)
        for i, name in enumerate(fields):
            F = getattr(self, name)
            _lib.check(_lib.lib().wf_count_nonfinite(
                self._handle, _lib.ptr(F.array), F.array.numel(), _lib.ptr(out[i:]),
                _lib.stream_ptr()), "wf_count_nonfinite")
        return dict(zip(fields, out.tolist()))

    def check_finite(self, fields=("E", "H", "J")):
        """True if the named fields hold no Inf or NaN (device check)."""
        return not any(self.count_nonfinite(fields).values())

    def one_step(self):
        """One leapfrog step (solverFIT3D.py:482-503)."""
        if self.step_0:
            self._set_ghosts_to_0()
            self.step_0 = False
        self._run(self._step_args(1))
        self.nstep += 1

    def _one_step(self):
        self.one_step()

    # ------------------------------------------------------- state helpers
    def update_logger(self, attrs):
        """Copy solver attributes into ``logger.solver``; 'grid' takes the
        grid's log (solverFIT3D.py:1015-1029)."""
        for atr in attrs:
            if atr == "grid":
                self.logger.grid = self.grid.logger.grid
            else:
                self.logger.solver[atr] = getattr(self, atr)

    def reset_fields(self):
        """solverFIT3D.py:995-1004."""
        for F in (self.E, self.H, self.J):
            F.array.zero_()
        self._J_clean_version = self.J.version
        self._beam_line = None

    def save_state(self, filename="solver_state.h5", close=True):
        """Dynamic state H, E, J as flat 3N datasets of an HDF5 file
        (solverFIT3D.py:886-934), in the field dtype. ``close=False`` returns
        the open file, as the reference does. A name ending in ``.npz`` writes
        the same three arrays with numpy instead."""
        H, E, J = self.H.to_numpy(), self.E.to_numpy(), self.J.to_numpy()
        if str(filename).endswith(".npz"):
            np.savez(filename, H=H, E=E, J=J)
            return None
        from . import h5io
        state = h5io.File(filename, "w")
        state.create_dataset("H", data=H)
        state.create_dataset("E", data=E)
        state.create_dataset("J", data=J)
        if close:
            state.close()
            return None
        return state

    def load_state(self, filename="solver_state.h5"):
        """solverFIT3D.py:936-980 (single device): E, H, J from a file written
        by save_state (HDF5, the reference's files included, or .npz)."""
        from . import h5io
        if h5io.is_hdf5(filename):
            with h5io.File(filename, "r") as st:
                E, H, J = st["E"][:], st["H"][:], st["J"][:]
        else:
            with np.load(filename, allow_pickle=False) as st:
                E, H, J = st["E"], st["H"], st["J"]
        self.E.fromarray(E)
        self.H.fromarray(H)
        self.J.fromarray(J)

    def read_state(self, filename="solver_state.h5"):
        """The state file opened read-only, not loaded (solverFIT3D.py:982-993)."""
        from . import h5io
        return h5io.File(filename, "r")

    def synchronize(self):
        torch.cuda.synchronize(self.device)
