"""CPU oracle: numpy/scipy restatement of the reference FIT leapfrog.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``). Every function names
the reference file:line it restates (paths relative to the Wakis v0.7.0
source tree). The arithmetic order is kept: the operators are assembled with
the same scipy ``diags``/``hstack``/``vstack`` calls and the same left-to-right
``*`` association, and the step uses the same ``H - dt*M1*E`` /
``E + dt*(M2*H - ieps*J)`` expressions, so that this oracle reproduces the
reference's numbers (including its FP32 mode under NumPy 2 promotion rules).

Field layout (``field.py:19-24, 57-91``): a flat array of 3N values, x|y|z
component blocks, each block in Fortran order (x fastest),
``n = i + Nx*j + Nx*Ny*k``.
"""

import numpy as np
from scipy.constants import c as c_light
from scipy.constants import epsilon_0 as eps_0
from scipy.constants import mu_0
from scipy.sparse import csc_matrix, diags, hstack, vstack

# materials.py:27-41 (library values, relative eps, mu [, sigma])
MATERIAL_LIB = {
    "pec": [np.inf, 1.0],
    "vacuum": [1.0, 1.0],
    "dielectric": [10.0, 1.0],
    "lossy metal": [10, 1.0, 10],
    "copper": [5.8e07, 1.0, 5.8e07],
    "berillium": [2.5e07, 1.0, 2.5e07],
    "stainless steel": [1.4e06, 1.0, 1.4e06],
}

_COMP = {"x": 0, "y": 1, "z": 2, 0: 0, 1: 1, 2: 2}


def comp(arr, dims, d):
    """F-order (Nx,Ny,Nz) view of component ``d`` of a flat 3N array.

    Restates ``Field.to_matrix`` (field.py:139-168); on numpy this is a view.
    """
    Nx, Ny, Nz = dims
    N = Nx * Ny * Nz
    c = _COMP[d]
    return np.reshape(arr[c * N:(c + 1) * N], (Nx, Ny, Nz), order="F")


def set_comp_from_matrix(arr, dims, d, mat):
    """``Field.from_matrix`` (field.py:170-190): F-order flatten, cast on store."""
    Nx, Ny, Nz = dims
    N = Nx * Ny * Nz
    c = _COMP[d]
    arr[c * N:(c + 1) * N] = np.reshape(mat, N, order="F")


# --------------------------------------------------------------------------
# Grid (gridFIT3D.py:141-180, 258-300)
# --------------------------------------------------------------------------
def uniform_axes(xmin, xmax, ymin, ymax, zmin, zmax, Nx, Ny, Nz):
    """Node axes of a uniform grid (gridFIT3D.py:164-166)."""
    return (np.linspace(xmin, xmax, Nx + 1), np.linspace(ymin, ymax, Ny + 1),
            np.linspace(zmin, zmax, Nz + 1))


def grid_metrics(x, y, z):
    """Primal/dual edge lengths and inverse areas (gridFIT3D.py:258-300).

    Returns (L, iA, tL, itA) as flat float64 3N arrays. The dual axis repeats
    its last midpoint, so tL is 0 on the last index of each direction and
    itA is 0 wherever the product of the two dual lengths is 0.
    """
    Nx, Ny, Nz = len(x) - 1, len(y) - 1, len(z) - 1
    dims = (Nx, Ny, Nz)
    N = Nx * Ny * Nz
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    L = np.zeros(3 * N)
    set_comp_from_matrix(L, dims, "x", X[1:, 1:, 1:] - X[:-1, :-1, :-1])
    set_comp_from_matrix(L, dims, "y", Y[1:, 1:, 1:] - Y[:-1, :-1, :-1])
    set_comp_from_matrix(L, dims, "z", Z[1:, 1:, 1:] - Z[:-1, :-1, :-1])
    Lx, Ly, Lz = L[:N], L[N:2 * N], L[2 * N:]
    iA = np.zeros(3 * N)
    iA[:N] = np.divide(1.0, Ly * Lz)
    iA[N:2 * N] = np.divide(1.0, Lx * Lz)
    iA[2 * N:] = np.divide(1.0, Lx * Ly)

    tx = np.append((x[1:] + x[:-1]) / 2, ((x[1:] + x[:-1]) / 2)[-1])
    ty = np.append((y[1:] + y[:-1]) / 2, ((y[1:] + y[:-1]) / 2)[-1])
    tz = np.append((z[1:] + z[:-1]) / 2, ((z[1:] + z[:-1]) / 2)[-1])
    tX, tY, tZ = np.meshgrid(tx, ty, tz, indexing="ij")
    tL = np.zeros(3 * N)
    set_comp_from_matrix(tL, dims, "x", tX[1:, 1:, 1:] - tX[:-1, :-1, :-1])
    set_comp_from_matrix(tL, dims, "y", tY[1:, 1:, 1:] - tY[:-1, :-1, :-1])
    set_comp_from_matrix(tL, dims, "z", tZ[1:, 1:, 1:] - tZ[:-1, :-1, :-1])
    tLx, tLy, tLz = tL[:N], tL[N:2 * N], tL[2 * N:]
    itA = np.zeros(3 * N)
    for c, aux in enumerate((tLy * tLz, tLx * tLz, tLx * tLy)):
        itA[c * N:(c + 1) * N] = np.divide(1.0, aux, out=np.zeros_like(aux),
                                           where=aux != 0)
    return L, iA, tL, itA


def cell_centres(x):
    """Solver cell-centre axis ``x[:-1] + dx/2`` (solverFIT3D.py:163-165)."""
    return x[:-1] + np.diff(x) / 2


# --------------------------------------------------------------------------
# Boundary conditions (boundaries.py:14-188)
# --------------------------------------------------------------------------
def _face_mask(dims, f):
    """int8 3N mask with tangential faces set in the reference's order.

    boundaries.py:90-103 / 130-143: for component d the two faces of every
    other axis are assigned in x, y, z order, later assignments winning.
    ``f`` = (xlo, ylo, zlo, xhi, yhi, zhi).
    """
    xlo, ylo, zlo, xhi, yhi, zhi = f
    Nx, Ny, Nz = dims
    BC = np.ones(3 * Nx * Ny * Nz, dtype=np.int8)
    for d in ("x", "y", "z"):
        V = comp(BC, dims, d)
        if d != "x":
            V[0, :, :] = xlo
            V[-1, :, :] = xhi
        if d != "y":
            V[:, 0, :] = ylo
            V[:, -1, :] = yhi
        if d != "z":
            V[:, :, 0] = zlo
            V[:, :, -1] = zhi
    return BC


def bc_column_row_masks(bc_low, bc_high, dims):
    """PEC column mask and PMC row mask of the curl (boundaries.py:72-150).

    Returns ``(pec_mask or None, pmc_mask or None)``. The PMC face flags start
    from the PEC-modified ones (boundaries.py:24-25 are not reset), and the PMC
    test compares ``.lower() == "magnetic"`` or the raw string ``== "pmc"``.
    """
    fl = [1.0] * 6  # xlo, ylo, zlo, xhi, yhi, zhi
    pec = None
    pmc = None
    pec_keys = ("electric", "pec", "pml")
    if any(b.lower() in pec_keys for b in bc_low) or any(
            b.lower() in pec_keys for b in bc_high):
        for a in range(3):
            if bc_low[a].lower() in pec_keys:
                fl[a] = 0
            if bc_high[a].lower() in pec_keys:
                fl[3 + a] = 0
        pec = _face_mask(dims, fl)
    if any(b.lower() in ("magnetic", "pmc") for b in bc_low) or any(
            b.lower() in ("magnetic", "pmc") for b in bc_high):
        for a in range(3):
            if bc_low[a].lower() == "magnetic" or bc_low[a] == "pmc":
                fl[a] = 0
            if bc_high[a].lower() == "magnetic" or bc_high[a] == "pmc":
                fl[3 + a] = 0
        pmc = _face_mask(dims, fl)
    return pec, pmc


def bc_metric_edits(bc_low, bc_high, dims, L, iA, tL, itA):
    """Periodic and ABC metric copies (boundaries.py:35-70, 152-181).

    Edits copies of tL/itA in place of the reference's Field writes and
    returns (tL, itA, activate_abc, activate_pml).
    """
    tL = tL.copy()
    itA = itA.copy()

    def copy_axis(a):
        # high-side dual metric <- low-side primal metric along axis a
        sl_hi = [slice(None)] * 3
        sl_lo = [slice(None)] * 3
        sl_hi[a] = -1
        sl_lo[a] = 0
        sl_hi, sl_lo = tuple(sl_hi), tuple(sl_lo)
        names = ("x", "y", "z")
        comp(tL, dims, names[a])[sl_hi] = comp(L, dims, names[a])[sl_lo]
        for b in range(3):
            if b != a:
                comp(itA, dims, names[b])[sl_hi] = comp(iA, dims, names[b])[sl_lo]

    if any(b.lower() == "periodic" for b in bc_low):
        for a in range(3):
            if bc_low[a].lower() == "periodic" and bc_high[a].lower() == "periodic":
                copy_axis(a)
    activate_abc = False
    if any(b.lower() == "abc" for b in bc_low) or any(
            b.lower() == "abc" for b in bc_high):
        for a in range(3):
            if bc_high[a].lower() == "abc":
                copy_axis(a)
        activate_abc = True
    activate_pml = any(b.lower() == "pml" for b in bc_low) or any(
        b.lower() == "pml" for b in bc_high)
    return tL, itA, activate_abc, activate_pml


# --------------------------------------------------------------------------
# Geometry fills (stand-in for gridFIT3D.py:470-517 "any corner inside")
# --------------------------------------------------------------------------
def corner_any_inside(x, y, z, inside):
    """Cell mask: a cell is inside if any of its 8 corner nodes is inside.

    ``inside(X, Y, Z) -> bool array`` on node coordinates. This is the
    reference's point_data_to_cell_data(...) > tol rule (gridFIT3D.py:504-506)
    with an analytic point test instead of VTK's enclosed-points filter.
    """
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    P = inside(X, Y, Z)
    return (P[:-1, :-1, :-1] | P[1:, :-1, :-1] | P[:-1, 1:, :-1] | P[1:, 1:, :-1]
            | P[:-1, :-1, 1:] | P[1:, :-1, 1:] | P[:-1, 1:, 1:] | P[1:, 1:, 1:])


# --------------------------------------------------------------------------
# Solver (solverFIT3D.py:40-354, 416-503, 820-884)
# --------------------------------------------------------------------------
def sibc_surface(mask):
    """Surface cells of a solid for the SIBC substitution (gridFIT3D.py:518-533):
    the cells where the gradient of the solid's 0/1 cell field is nonzero, the
    gradient restated as index-space finite differences (central in the
    interior, one-sided on the first and last cell of each axis). Parity
    unpinned: the reference computes it with pyvista's compute_derivative,
    which is not available to check against; the tests pin only the rule's
    invariants (a shell inside the solid, one cell thick on flat faces).
    """
    m = np.asarray(mask).astype(np.int64)
    out = np.zeros(m.shape, dtype=bool)
    for ax in range(3):
        n = m.shape[ax]
        if n < 2:
            continue
        d = np.zeros_like(m)
        sl = lambda a, b: tuple(slice(a, b) if i == ax else slice(None) for i in range(3))
        d[sl(1, n - 1)] = m[sl(2, n)] - m[sl(0, n - 2)]
        d[sl(0, 1)] = m[sl(1, 2)] - m[sl(0, 1)]
        d[sl(n - 1, n)] = m[sl(n - 1, n)] - m[sl(n - 2, n - 1)]
        out |= d != 0
    return out


class OracleSolver:
    """Restatement of ``SolverFIT3D`` on the CPU scipy path (no MKL/GPU/MPI).

    ``solids`` is a list of ``(mask3d_bool, material)`` applied in order with
    the arithmetic of ``_apply_stl_materials`` (solverFIT3D.py:416-480),
    including the surface-impedance substitution when ``use_sibc`` (the surface
    rule of ``sibc_surface``, parity unpinned).
    """

    def __init__(self, x, y, z, bc_low, bc_high, bg=(1.0, 1.0), solids=(),
                 dtype=np.float64, cfln=0.5, dt=None, n_pml=10, use_sibc=False, fmax=1e9):
        self.use_sibc, self.fmax = use_sibc, fmax
        self.x_nodes, self.y_nodes, self.z_nodes = x, y, z
        self.Nx, self.Ny, self.Nz = len(x) - 1, len(y) - 1, len(z) - 1
        self.dims = (self.Nx, self.Ny, self.Nz)
        self.N = N = self.Nx * self.Ny * self.Nz
        self.dx, self.dy, self.dz = np.diff(x), np.diff(y), np.diff(z)
        self.x, self.y, self.z = cell_centres(x), cell_centres(y), cell_centres(z)
        self.dtype = dtype
        self.bc_low, self.bc_high = list(bc_low), list(bc_high)
        self.step_0 = True
        self.use_conductivity = False

        self.E = np.zeros(3 * N, dtype=dtype)
        self.H = np.zeros(3 * N, dtype=dtype)
        self.J = np.zeros(3 * N, dtype=dtype)

        # metrics and the curl (solverFIT3D.py:203-227)
        L, iA, tL, itA = grid_metrics(x, y, z)
        self.L, self.iA = L, iA
        Px = diags([-1, 1], [0, 1], shape=(N, N), dtype=np.int8)
        Py = diags([-1, 1], [0, self.Nx], shape=(N, N), dtype=np.int8)
        Pz = diags([-1, 1], [0, self.Nx * self.Ny], shape=(N, N), dtype=np.int8)
        Z0 = csc_matrix((N, N))
        C = vstack([hstack([Z0, -Pz, Py]), hstack([Pz, Z0, -Px]),
                    hstack([-Py, Px, Z0])], dtype=np.int8)

        # boundaries (boundaries.py:14-188)
        self.tL, self.itA, self.activate_abc, self.activate_pml = bc_metric_edits(
            self.bc_low, self.bc_high, self.dims, L, iA, tL, itA)
        self.pec_mask, self.pmc_mask = bc_column_row_masks(
            self.bc_low, self.bc_high, self.dims)
        if self.pec_mask is not None:
            C = C * diags(self.pec_mask, shape=(3 * N, 3 * N), dtype=np.int8)
        if self.pmc_mask is not None:
            C = diags(self.pmc_mask, shape=(3 * N, 3 * N), dtype=np.int8) * C
        if self.activate_pml:
            self.use_conductivity = True
        self.C = C

        # materials (solverFIT3D.py:237-273)
        if isinstance(bg, str):
            bg = MATERIAL_LIB[bg.lower()]
        if len(bg) == 3:
            eps_bg, mu_bg, sigma_bg = bg[0] * eps_0, bg[1] * mu_0, bg[2]
            self.use_conductivity = True
        else:
            eps_bg, mu_bg, sigma_bg = bg[0] * eps_0, bg[1] * mu_0, 0.0
        self.sigma_bg = sigma_bg
        self.ieps = np.ones(3 * N, dtype=dtype) * (1.0 / eps_bg)
        self.imu = np.ones(3 * N, dtype=dtype) * (1.0 / mu_bg)
        self.sigma = np.ones(3 * N, dtype=dtype) * sigma_bg
        for mask, mat in solids:
            self._apply_solid(np.asarray(mask), mat)

        if self.activate_pml:
            self.n_pml = n_pml
            init_pml(self)

        # timestep (solverFIT3D.py:283-310)
        self.cfln = cfln
        if dt is None:
            dt = cfln / (c_light * np.sqrt(1 / np.min(self.dx) ** 2
                                           + 1 / np.min(self.dy) ** 2
                                           + 1 / np.min(self.dz) ** 2))
        self.dt = dtype(dt)
        if self.use_conductivity:
            m = np.logical_and(self.sigma != 0, self.ieps != 0)
            tau = (1 / self.ieps[m]) / self.sigma[m]
            if tau.size and self.dt > tau.min():
                self.dt = tau.min()
        self.assemble()

    def _apply_solid(self, mask, mat):
        """Material override arithmetic of solverFIT3D.py:416-480."""
        if isinstance(mat, str):
            mat = MATERIAL_LIB[mat.lower()]
        maski = mask.astype(int)
        eps = mat[0] * eps_0
        mu = mat[1] * mu_0
        if len(mat) == 3:
            # conductivity the grid resolves without SIBC (solverFIT3D.py:445-447)
            dn = np.sqrt(2) * min(self.dx.min(), self.dy.min(), self.dz.min())
            if self.use_sibc and mat[2] > 10 / (np.pi * self.fmax * mu * dn ** 2):
                # surface cells, surface impedance as conductivity and as the
                # (absolute) permittivity (solverFIT3D.py:455-466)
                maski = sibc_surface(mask).astype(int)
                imp = np.sqrt(np.pi * self.fmax * mu / mat[2])
                mat = (mat[0], mat[1], 1 / imp)
                eps = 1 / imp

        def scaled_add(field, scale3d):
            # Field.__mul__ with a 3-D matrix, then Field.__add__ (field.py:350-460)
            tmp = np.zeros_like(field)
            for d in ("x", "y", "z"):
                set_comp_from_matrix(tmp, self.dims, d, comp(field, self.dims, d) * scale3d)
            return field + tmp

        def mat_add(field, m3d):
            out = np.empty_like(field)
            for d in ("x", "y", "z"):
                set_comp_from_matrix(out, self.dims, d, comp(field, self.dims, d) + m3d)
            return out

        if len(mat) == 3:
            sigma = mat[2]
            self.sigma = scaled_add(self.sigma, -1.0 * maski)
            self.sigma = mat_add(self.sigma, maski * sigma)
            self.use_conductivity = True
        elif self.sigma_bg > 0.0:
            self.sigma = scaled_add(self.sigma, -1.0 * maski)
        self.ieps = scaled_add(self.ieps, -1.0 * maski)
        self.imu = scaled_add(self.imu, -1.0 * maski)
        self.ieps = mat_add(self.ieps, maski * 1.0 / eps)
        self.imu = mat_add(self.imu, maski * 1.0 / mu)

    def assemble(self):
        """Operators M1=iDa*iDmu*C*Ds, M2=iDeps*itDa*C^T*tDs (solverFIT3D.py:312-322)."""
        n3 = 3 * self.N
        dt_ = self.dtype
        iDeps = diags(self.ieps, shape=(n3, n3), dtype=dt_)
        iDmu = diags(self.imu, shape=(n3, n3), dtype=dt_)
        Ds = diags(self.L, shape=(n3, n3), dtype=dt_)
        iDa = diags(self.iA, shape=(n3, n3), dtype=dt_)
        tDs = diags(self.tL, shape=(n3, n3), dtype=dt_)
        itDa = diags(self.itA, shape=(n3, n3), dtype=dt_)
        self.M1 = iDa * iDmu * self.C * Ds
        self.M2 = iDeps * itDa * self.C.transpose() * tDs

    def set_ghosts_to_0(self):
        """solverFIT3D.py:820-840 (first step only)."""
        for d in ("x", "y", "z"):
            Hd = comp(self.H, self.dims, d)
            if d != "x":
                Hd[-1, :, :] = 0.0
            if d != "y":
                Hd[:, -1, :] = 0.0
            if d != "z":
                Hd[:, :, -1] = 0.0
        comp(self.E, self.dims, "x")[-1, :, :] = 0.0
        comp(self.E, self.dims, "y")[:, -1, :] = 0.0
        comp(self.E, self.dims, "z")[:, :, -1] = 0.0

    def one_step(self):
        """``_one_step`` (solverFIT3D.py:482-503)."""
        if self.step_0:
            self.set_ghosts_to_0()
            self.step_0 = False
        self.H[:] = self.H - self.dt * self.M1 * self.E
        self.E[:] = self.E + self.dt * (self.M2 * self.H - self.ieps * self.J)
        if self.use_conductivity:
            self.J[:] = self.sigma * self.E


def init_pml(s):
    """Graded-conductivity PML (boundaries.py:190-311).

    Writes ``ieps``/``sigma`` of the outer ``n_pml`` layers from the values
    one layer inside, sampled at the face centre; no auxiliary field.
    """
    dims = s.dims
    Nx, Ny, Nz = dims
    n = s.n_pml
    lo, hi = 5.0e-3, 10.0
    prof_lo = np.geomspace(hi, lo, n)   # low faces: 10 -> 5e-3 inward
    prof_hi = np.geomspace(lo, hi, n)   # high faces: 5e-3 -> 10 outward

    def sample(arr, d, idx):
        return comp(arr, dims, d)[idx]

    # faces in the reference's order: x-, y-, z-low, then x-, y-, z-high
    for side in ("low", "high"):
        for axis in range(3):
            bc = s.bc_low[axis] if side == "low" else s.bc_high[axis]
            if bc.lower() != "pml":
                continue
            prof = np.zeros(dims[axis])
            if side == "low":
                prof[0:n] = prof_lo
            else:
                prof[-n:] = prof_hi
            for d in ("x", "y", "z"):
                centre = [Nx // 2, Ny // 2, Nz // 2]
                centre[axis] = (n + 1) if side == "low" else -(n + 1)
                ieps0 = sample(s.ieps, d, tuple(centre))
                sig0 = sample(s.sigma, d, tuple(centre))
                mult = 1 if sig0 < 1 else sig0
                for m in range(n):
                    idx = m if side == "low" else -(m + 1)
                    sl = [slice(None)] * 3
                    sl[axis] = idx
                    comp(s.ieps, dims, d)[tuple(sl)] = ieps0
                    comp(s.sigma, dims, d)[tuple(sl)] = sig0 + mult * prof[idx]


# --------------------------------------------------------------------------
# Sources (sources.py:22-110, 142-282)
# --------------------------------------------------------------------------
class OracleBeam:
    """Gaussian beam current on the z line through the source (sources.py:22-110)."""

    def __init__(self, xsource=0.0, ysource=0.0, beta=1.0, q=1e-9, sigmaz=None, ti=None):
        self.xsource, self.ysource = xsource, ysource
        self.sigmaz, self.q, self.beta = sigmaz, q, beta
        self.v = c_light * beta
        self.ti = ti if ti is not None else 8.548921333333334 * self.sigmaz / self.v
        self.is_first_update = True

    def profile(self, solver, t):
        if self.is_first_update:
            self.ixs = np.abs(solver.x - self.xsource).argmin()
            self.iys = np.abs(solver.y - self.ysource).argmin()
            self.zmin = solver.z.min()
            self.is_first_update = False
        s0 = self.zmin - self.v * self.ti
        s = solver.z - self.v * t
        prof = 1 / np.sqrt(2 * np.pi * self.sigmaz ** 2) * np.exp(
            -((s - s0) ** 2) / (2 * self.sigmaz ** 2))
        return self.q * self.v * prof / solver.dx[self.ixs] / solver.dy[self.iys]

    def update(self, solver, t):
        line = self.profile(solver, t)
        comp(solver.J, solver.dims, "z")[self.ixs, self.iys, :] = line


class OraclePlaneWave:
    """z-travelling plane wave written on a z plane (sources.py:142-282)."""

    def __init__(self, xs=None, ys=None, zs=0, nodes=None, f=None, amplitude=1.0,
                 beta=1.0, phase=0):
        self.xs, self.ys, self.zs = xs, ys, zs
        self.f, self.amplitude, self.beta, self.phase = f, amplitude, beta, phase
        self.vp = beta * c_light
        self.w = 2 * np.pi * f
        self.kz = self.w / c_light
        self.tmax = nodes / f if nodes is not None else np.inf

    def update(self, solver, t):
        xs = self.xs if self.xs is not None else slice(0, solver.Nx)
        ys = self.ys if self.ys is not None else slice(0, solver.Ny)
        Hy = comp(solver.H, solver.dims, "y")
        Ex = comp(solver.E, solver.dims, "x")
        if t <= self.tmax:
            Hy[xs, ys, self.zs] = -self.amplitude * np.cos(self.w * t + self.phase)
            Ex[xs, ys, self.zs] = (self.amplitude * np.cos(self.w * t + self.phase)
                                   / (self.kz / (mu_0 * self.vp)))
        else:
            Hy[xs, ys, self.zs] = 0.0
            Ex[xs, ys, self.zs] = 0.0


class OracleWavePacket:
    """Gaussian wave packet on a z plane (sources.py:285-455): H_y and E_x of
    the plane are -A cos(w t + phi) g_xy g_t and A mu0 c cos(w t + phi) g_xy g_t
    with g_xy over meshgrid(x[xs], y[ys]) and g_t of s = z_min - beta c t
    around s0 = z_min - t_inj (sources.py:397-429)."""

    def __init__(self, xs=None, ys=None, zs=0, sigmaz=None, sigmaxy=None, tinj=None,
                 wavelength=None, f=None, amplitude=1.0, beta=1.0, phase=0):
        self.xs, self.ys, self.zs = xs, ys, zs
        self.sigmaz, self.sigmaxy = sigmaz, sigmaxy
        self.amplitude, self.beta, self.phase = amplitude, beta, phase
        self.f = f if f is not None else c_light / wavelength
        self.w = 2 * np.pi * self.f
        self.tinj = tinj if tinj is not None else 6 * sigmaz

    def update(self, solver, t):
        xs = self.xs if self.xs is not None else slice(0, solver.Nx)
        ys = self.ys if self.ys is not None else slice(0, solver.Ny)
        X, Y = np.meshgrid(solver.x[xs], solver.y[ys])
        gxy = np.exp(-(X ** 2 + Y ** 2) / (2 * self.sigmaxy ** 2))
        s0 = solver.z.min() - self.tinj
        s = solver.z.min() - self.beta * c_light * t
        gt = np.exp(-((s - s0) ** 2) / (2 * self.sigmaz ** 2))
        cw = np.cos(self.w * t + self.phase)
        comp(solver.H, solver.dims, "y")[xs, ys, self.zs] = -self.amplitude * cw * gxy * gt
        comp(solver.E, solver.dims, "x")[xs, ys, self.zs] = (self.amplitude * mu_0 * c_light
                                                              * cw * gxy * gt)


def _oracle_field(solver, field):
    return {"E": solver.E, "H": solver.H, "J": solver.J}[field]


class OracleDipole:
    """A sin(w t + phi) on one component of E, H or J at one cell
    (sources.py:458-566); f = nodes c / (z_max - z_min) when not given."""

    def __init__(self, field="E", component="z", xs=None, ys=None, zs=None, nodes=10, f=None,
                 amplitude=1.0, phase=0):
        if len(field) == 2:
            field, component = field[0], field[1]
        self.field, self.component = field, component
        self.xs, self.ys, self.zs, self.nodes, self.f = xs, ys, zs, nodes, f
        self.amplitude, self.phase = amplitude, phase

    def update(self, solver, t):
        xs = self.xs if self.xs is not None else int(solver.Nx / 2)
        ys = self.ys if self.ys is not None else int(solver.Ny / 2)
        zs = self.zs if self.zs is not None else int(solver.Nz / 2)
        f = self.f if self.f is not None else self.nodes / ((solver.z.max() - solver.z.min())
                                                             / c_light)
        comp(_oracle_field(solver, self.field), solver.dims, self.component)[xs, ys, zs] = (
            self.amplitude * np.sin(2 * np.pi * f * t + self.phase))


class OraclePulse:
    """Pulse of tau = c t - delay on one component (sources.py:569-762):
    Harris (10 - 15 cos(2 pi tau / L) + 6 cos(4 pi tau / L) - cos(6 pi tau / L))
    / 32 for tau < L, Gaussian exp(-(tau - L/2)^2 / (2 (L/10)^2)) with L/2
    written 5 (L/10), rectangular 1 on 0 < tau < L; L = 50 dt by default."""

    def __init__(self, field="E", component="z", xs=None, ys=None, zs=None, shape="Harris",
                 L=None, amplitude=1.0, delay=0.0):
        if len(field) == 2:
            field, component = field[0], field[1]
        self.field, self.component, self.shape = field, component, shape.lower()
        self.xs, self.ys, self.zs, self.L = xs, ys, zs, L
        self.amplitude, self.delay = amplitude, delay

    def profile(self, t, L):
        tau = t * c_light - self.delay
        if self.shape == "harris":
            if tau >= L:
                return 0.0
            return (10 - 15 * np.cos(2 * np.pi / L * tau) + 6 * np.cos(4 * np.pi / L * tau)
                    - np.cos(6 * np.pi / L * tau)) / 32
        if self.shape == "gaussian":
            return np.exp(-((tau - 5 * (L / 10)) ** 2) / (2 * (L / 10) ** 2))
        return 1.0 if 0.0 < tau < L else 0.0

    def update(self, solver, t):
        xs = self.xs if self.xs is not None else int(solver.Nx / 2)
        ys = self.ys if self.ys is not None else int(solver.Ny / 2)
        zs = self.zs if self.zs is not None else int(solver.Nz / 2)
        L = self.L if self.L is not None else 50 * solver.dt
        comp(_oracle_field(solver, self.field), solver.dims, self.component)[xs, ys, zs] = (
            self.amplitude * self.profile(t, L))


# --------------------------------------------------------------------------
# Drivers (routines.py:17-229, 231-482)
# --------------------------------------------------------------------------
def abc_copy_low_faces(s):
    """CPU ``get_abc``/``update_abc`` effect (boundaries.py:313-410).

    On CPU the saved slices are live views, so ``update_abc`` copies the
    current index-1 plane into index 0 on every low ABC face (x, then y,
    then z; E then H per component) and leaves high faces unchanged.
    """
    for axis in range(3):
        if s.bc_low[axis].lower() != "abc":
            continue
        for d in ("x", "y", "z"):
            for F in (s.E, s.H):
                V = comp(F, s.dims, d)
                sl0 = [slice(None)] * 3
                sl1 = [slice(None)] * 3
                sl0[axis], sl1[axis] = 0, 1
                V[tuple(sl0)] = V[tuple(sl1)]


def emsolve(s, Nt, source=None, callback=None, record=None):
    """Time loop of ``emsolve`` without file output (routines.py:183-225).

    ``record(s, n)``, if given, is called where the reference saves the fields
    of step n: after the source update, before the step (routines.py:193-204).
    """
    for n in range(Nt):
        if source is not None:
            source.update(s, n * s.dt)
        if record is not None:
            record(s, n)
        s.one_step()
        if s.activate_abc:
            abc_copy_low_faces(s)
        if callback is not None:
            callback(s, n * s.dt)


def wakesolve_fields(s, wakelength, q, sigmaz, beta, xsource, ysource, xtest, ytest,
                     ti_wake, compute_plane="both", nsteps=None):
    """Field part of ``wakesolve`` (routines.py:319-459) with in-memory capture.

    Returns a dict with the per-step Ez window (what ``Ez.h5`` would hold),
    shape (Nt, 3, 3, Nz) or (Nt, Nz) for ``compute_plane='longitudinal'``,
    in the field dtype, the J_z window at the same point of each step (what
    ``Jz.h5`` would hold with ``save_J``, routines.py:441-442), and the ``t``
    axis ``np.arange(0, Nt*dt, dt)``.
    ``nsteps`` truncates the loop (the time axis keeps the full length).
    """
    ixt = np.abs(s.x - xtest).argmin()
    iyt = np.abs(s.y - ytest).argmin()
    if compute_plane.lower() == "longitudinal":
        xx, yy = ixt, iyt
    else:
        xx, yy = slice(ixt - 1, ixt + 2), slice(iyt - 1, iyt + 2)
    v = beta * c_light
    tmax = (wakelength + ti_wake * v + (s.z.max() - s.z.min())) / v
    Nt = int(tmax / s.dt)
    beam = OracleBeam(q=q, sigmaz=sigmaz, beta=beta, xsource=xsource, ysource=ysource)
    t = np.arange(0, Nt * s.dt, s.dt)
    nrun = Nt if nsteps is None else min(Nt, nsteps)
    Ez = comp(s.E, s.dims, "z")
    samples = np.zeros((nrun,) + Ez[xx, yy, :].shape, dtype=s.dtype)
    jsamples = np.zeros_like(samples)
    for n in range(nrun):
        beam.update(s, n * s.dt)
        samples[n] = Ez[xx, yy, :]
        jsamples[n] = comp(s.J, s.dims, "z")[xx, yy, :]
        s.one_step()
    return dict(Ez=samples, Jz=jsamples, t=t, Nt=Nt, tmax=tmax, ixt=ixt, iyt=iyt,
                xf=s.x[xx], yf=s.y[yy], zf=s.z, dx=s.dx, dy=s.dy, dz=s.dz,
                ixs=beam.ixs, iys=beam.iys)
