"""CPU: pins of the oracle (oracle/) against what the mathematics fixes.

The oracle restates the reference algorithm; these tests check it against
facts that do not come from the oracle's own formulas: closed forms, discrete
identities of the curl, the leapfrog's conserved energy, the reference's
resonator test (test_002) and the reference's golden vectors (test_001).
Where a test could pass with a plausible mistake in the oracle, a second test
plants that mistake (in a copy, on the CPU) and checks that the pin fails.
"""

import math

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.constants import c as c_light
from scipy.constants import mu_0

from oracle import cases
from oracle import fit_oracle as fo
from oracle import wake_oracle as wo


# ---------------------------------------------------------------- metrics
def test_grid_metrics_uniform_closed_form():
    Nx, Ny, Nz = 5, 4, 6
    x, y, z = fo.uniform_axes(0.0, 1.0, -0.5, 0.5, 0.0, 3.0, Nx, Ny, Nz)
    dx, dy, dz = 1.0 / Nx, 1.0 / Ny, 3.0 / Nz
    L, iA, tL, itA = fo.grid_metrics(x, y, z)
    dims = (Nx, Ny, Nz)
    for c, d in enumerate((dx, dy, dz)):
        np.testing.assert_allclose(fo.comp(L, dims, c), d, rtol=1e-14)
    np.testing.assert_allclose(fo.comp(iA, dims, 0), 1 / (dy * dz), rtol=1e-13)
    np.testing.assert_allclose(fo.comp(iA, dims, 1), 1 / (dx * dz), rtol=1e-13)
    np.testing.assert_allclose(fo.comp(iA, dims, 2), 1 / (dx * dy), rtol=1e-13)
    # dual lengths: the primal spacing, except 0 on the last index of the axis
    tLx, tLy, tLz = (fo.comp(tL, dims, c) for c in range(3))
    np.testing.assert_allclose(tLx[:-1], dx, rtol=1e-13)
    assert np.all(tLx[-1] == 0)
    np.testing.assert_allclose(tLy[:, :-1], dy, rtol=1e-13)
    assert np.all(tLy[:, -1] == 0)
    np.testing.assert_allclose(tLz[:, :, :-1], dz, rtol=1e-13)
    assert np.all(tLz[:, :, -1] == 0)
    # itA_x = 1/(tLy tLz): zero on the last y or z index, 1/(dy dz) elsewhere
    itAx = fo.comp(itA, dims, 0)
    assert np.all(itAx[:, -1, :] == 0) and np.all(itAx[:, :, -1] == 0)
    np.testing.assert_allclose(itAx[:, :-1, :-1], 1 / (dy * dz), rtol=1e-13)


def test_grid_metrics_dual_axis_nonuniform():
    """Dual nodes are the primal midpoints; tL is their spacing."""
    rng = np.random.default_rng(1)
    x = np.cumsum(np.r_[0.0, rng.uniform(0.5, 1.5, 6)])
    y = np.array([0.0, 1.0, 3.0])
    z = np.array([0.0, 2.0, 2.5, 4.0])
    L, iA, tL, itA = fo.grid_metrics(x, y, z)
    dims = (6, 2, 3)
    mid = (x[1:] + x[:-1]) / 2
    expect = np.append(np.diff(mid), 0.0)
    for j in range(2):
        for k in range(3):
            np.testing.assert_allclose(fo.comp(tL, dims, 0)[:, j, k], expect, rtol=1e-14)
            np.testing.assert_allclose(fo.comp(L, dims, 0)[:, j, k], np.diff(x), rtol=1e-14)


# ---------------------------------------------------------- boundary masks
def _literal_face_mask(dims, f):
    """Per-cell replay of the reference's slice assignments, in order
    (boundaries.py:90-103): the last statement that covers a cell wins."""
    Nx, Ny, Nz = dims
    xlo, ylo, zlo, xhi, yhi, zhi = f
    out = np.ones((3, Nx, Ny, Nz), dtype=np.int8)
    for c, d in enumerate("xyz"):
        for i in range(Nx):
            for j in range(Ny):
                for k in range(Nz):
                    v = 1
                    stmts = []
                    if d != "x":
                        stmts += [(i == 0, xlo), (i == Nx - 1, xhi)]
                    if d != "y":
                        stmts += [(j == 0, ylo), (j == Ny - 1, yhi)]
                    if d != "z":
                        stmts += [(k == 0, zlo), (k == Nz - 1, zhi)]
                    for hit, val in stmts:
                        if hit:
                            v = val
                    out[c, i, j, k] = v
    return out


@pytest.mark.parametrize("f", [(0, 0, 0, 0, 0, 0), (0, 1, 0, 1, 0, 1), (1, 1, 0, 0, 1, 1),
                               (0, 1, 1, 1, 1, 0)])
def test_face_mask_matches_literal_assignment_order(f):
    dims = (4, 5, 3)
    got = fo._face_mask(dims, f)
    ref = _literal_face_mask(dims, f)
    for c in range(3):
        assert np.array_equal(fo.comp(got, dims, c), ref[c])


def test_pmc_mask_inherits_pec_faces():
    """PEC flags are not reset before the PMC pass (boundaries.py:24-25, 116-127)."""
    pec, pmc = fo.bc_column_row_masks(["pec", "pmc", "pec"], ["pmc", "pec", "pec"], (4, 4, 4))
    assert pec is not None and pmc is not None
    assert np.array_equal(pmc, fo._face_mask((4, 4, 4), (0, 0, 0, 0, 0, 0)))
    # "PMC" in upper case passes the lower-cased guard but sets no face (the face
    # test compares the raw string with "pmc"): the row mask is the PEC one
    pec, pmc = fo.bc_column_row_masks(["PMC", "pec", "pec"], ["pec"] * 3, (4, 4, 4))
    assert np.array_equal(pmc, pec)
    assert np.array_equal(pec, fo._face_mask((4, 4, 4), (1, 0, 0, 0, 0, 0)))
    # "Magnetic" sets the face in any case
    _, mag = fo.bc_column_row_masks(["Magnetic", "pec", "pec"], ["pec"] * 3, (4, 4, 4))
    assert np.array_equal(mag, fo._face_mask((4, 4, 4), (0, 0, 0, 0, 0, 0)))


# ---------------------------------------------------------------- the curl
def _unmasked_curl(dims):
    s = fo.OracleSolver(*fo.uniform_axes(0, 1, 0, 1, 0, 1, *dims), ["periodic"] * 3,
                        ["periodic"] * 3)
    return s.C.tocsr().astype(np.int64), s


def test_curl_is_a_discrete_curl():
    """div(curl) = 0 and curl(grad) = 0 with the flat-stride differences."""
    dims = (4, 3, 5)
    C, s = _unmasked_curl(dims)
    N = s.N
    Px = sp.diags([-1, 1], [0, 1], shape=(N, N), dtype=np.int64)
    Py = sp.diags([-1, 1], [0, dims[0]], shape=(N, N), dtype=np.int64)
    Pz = sp.diags([-1, 1], [0, dims[0] * dims[1]], shape=(N, N), dtype=np.int64)
    D = sp.hstack([Px, Py, Pz]).tocsr()
    G = sp.vstack([Px, Py, Pz]).tocsr()
    assert abs(D @ C).sum() == 0
    assert abs(C @ G).sum() == 0
    assert C.nnz > 0


def test_curl_stokes_on_one_face():
    """Hx on a face equals the circulation of E around it (Stokes, right-hand rule)."""
    dims = (4, 4, 4)
    C, s = _unmasked_curl(dims)
    N = s.N
    Nx, Ny = dims[0], dims[1]
    i, j, k = 1, 2, 1
    n = i + Nx * j + Nx * Ny * k
    e = np.zeros(3 * N)
    # unit voltage on the four edges bounding the x-face at (i,j,k), counter-clockwise
    # seen from +x: +y edge at z=k, +z edge at y=j+1, -y edge at z=k+1, -z edge at y=j
    e[N + n] = 1.0                          # E_y(i,j,k)
    e[2 * N + n + Nx] = 1.0                 # E_z(i,j+1,k)
    e[N + n + Nx * Ny] = -1.0               # -E_y(i,j,k+1)
    e[2 * N + n] = -1.0                     # -E_z(i,j,k)
    h = C @ e
    # (curl E)_x = dEz/dy - dEy/dz: (1 - (-1)) - ((-1) - 1) = 4, the circulation
    assert h[n] == 4
    # the edges are shared with neighbouring x-faces, which see part of the loop
    assert h[n - Nx] != 0 and h[n + Nx] != 0


# ------------------------------------------------ leapfrog energy invariant
def _energy(so, E, Hold, Hnew):
    """Staggered FIT energy sum(eps*tA*L*E^2) + sum(mu*A*tL*Hold*Hnew) over live edges."""
    live = (so.ieps > 0) & (so.itA > 0)
    we = np.zeros_like(so.ieps, dtype=np.float64)
    we[live] = (1.0 / so.ieps[live]) * (1.0 / so.itA[live]) * so.L[live]
    wh = (1.0 / so.imu) * (1.0 / so.iA) * so.tL
    return float(np.sum(we * E.astype(np.float64) ** 2)
                 + np.sum(wh * Hold.astype(np.float64) * Hnew.astype(np.float64)))


def _energy_run(so, steps):
    so.set_ghosts_to_0()
    so.step_0 = False
    W = []
    for _ in range(steps):
        E0, H0 = so.E.copy(), so.H.copy()
        so.one_step()
        W.append(_energy(so, E0, H0, so.H))
    return np.array(W)


def _dielectric_box(seed, sigma_bg=None, bc=("pec", "pec", "pec")):
    rng = np.random.default_rng(seed)
    Nx, Ny, Nz = 9, 8, 10
    x = np.cumsum(np.r_[0.0, rng.uniform(0.8, 1.2, Nx)]) * 1e-3
    y = np.cumsum(np.r_[0.0, rng.uniform(0.8, 1.2, Ny)]) * 1e-3
    z = np.cumsum(np.r_[0.0, rng.uniform(0.8, 1.2, Nz)]) * 1e-3
    diel = rng.random((Nx, Ny, Nz)) < 0.4
    metal = rng.random((Nx, Ny, Nz)) < 0.1
    bg = [1.0, 1.0] if sigma_bg is None else [1.0, 1.0, sigma_bg]
    so = fo.OracleSolver(x, y, z, list(bc), list(bc), bg=bg,
                         solids=[(diel, [4.0, 2.0]), (metal, "pec")], dtype=np.float64)
    live = (so.ieps > 0) & (so.itA > 0)
    so.E[:] = np.where(live, rng.standard_normal(3 * so.N), 0.0)
    so.H[:] = rng.standard_normal(3 * so.N)
    return so


def test_leapfrog_conserves_energy_lossless():
    so = _dielectric_box(3)
    W = _energy_run(so, 300)
    drift = np.max(np.abs(W - W[0])) / W[0]
    assert drift < 1e-11, drift


def test_energy_pin_catches_a_transposed_or_dropped_term():
    """The invariant fails when the oracle carries a plausible mistake."""
    # (1) one M2 entry with the wrong sign
    so = _dielectric_box(3)
    M2 = so.M2.tocsr().copy()
    r = int(np.argmax(np.diff(M2.indptr) == 4))
    M2.data[M2.indptr[r]] *= -1
    so.M2 = M2
    W = _energy_run(so, 300)
    assert np.max(np.abs(W - W[0])) / W[0] > 1e-6
    # (2) the dual and primal lengths swapped in M1 (Ds -> tDs)
    so = _dielectric_box(3)
    n3 = 3 * so.N
    so.M1 = (sp.diags(so.iA, shape=(n3, n3)) * sp.diags(so.imu, shape=(n3, n3)) * so.C
             * sp.diags(np.where(so.tL > 0, so.tL, so.L) * 1.01, shape=(n3, n3)))
    W = _energy_run(so, 300)
    assert np.max(np.abs(W - W[0])) / W[0] > 1e-6


def test_energy_decreases_with_conductivity():
    so = _dielectric_box(5, sigma_bg=2.0)
    so.J[:] = so.sigma * so.E
    W = _energy_run(so, 200)
    assert np.all(np.diff(W) <= 1e-12 * W[0]), "energy must not grow with sigma > 0"
    assert W[-1] < 0.999 * W[0]


# --------------------------------------------------- test_002 closed form
def test_resonator_te101_frequency():
    """reference tests/test_002_analytic_resonator.py: PEC cube L=1 m, N=30,
    H of the TE101 mode at t=-dt/2, 5000 steps; the Ey spectrum peaks at
    the analytic 2.121e8 Hz within 5%."""
    Lx = Ly = Lz = 1.0
    N = 30
    d = Lx / N
    x, y, z = fo.uniform_axes(-Lx / 2 + d / 2, Lx / 2 + d / 2, -Ly / 2 + d / 2, Ly / 2 + d / 2,
                              -Lz / 2 + d / 2, Lz / 2 + d / 2, N, N, N)
    so = fo.OracleSolver(x, y, z, ["pec"] * 3, ["pec"] * 3)
    m, n, p = 1, 0, 1
    ii = np.arange(N)
    X, Y, Z = np.meshgrid(ii * d + x[0], ii * d + y[0], ii * d + z[0], indexing="ij")
    t = -0.5 * so.dt
    h2 = (m * np.pi / Lx) ** 2 + (n * np.pi / Ly) ** 2 + (p * np.pi / Lz) ** 2
    wt = np.cos(np.sqrt(2) * np.pi / Lx * c_light * t)
    Hz = (np.cos(m * np.pi / Lx * (X - Lx / 2)) * np.cos(n * np.pi / Ly * (Y - Ly / 2))
          * np.sin(p * np.pi / Lz * (Z - Lz / 2)) * wt)
    Hx = (-2 / h2 * (m * np.pi / Lx) * (p * np.pi / Lz) * np.sin(m * np.pi / Lx * (X - Lx / 2))
          * np.cos(n * np.pi / Ly * (Y - Ly / 2)) * np.cos(p * np.pi / Lz * (Z - Lz / 2)) * wt)
    fo.set_comp_from_matrix(so.H, so.dims, "x", Hx)
    fo.set_comp_from_matrix(so.H, so.dims, "z", Hz)
    Ey = []
    for _ in range(5000):
        Ey.append(fo.comp(so.E, so.dims, "y")[N // 2, N // 2, N // 2])
        so.one_step()
    Ey = np.array(Ey)
    F = np.fft.fftshift(np.fft.fft(Ey))[len(Ey) // 2:]
    f = np.fft.fftshift(np.fft.fftfreq(len(Ey), d=so.dt))[len(Ey) // 2:]
    fr = f[np.argmax(F)]
    assert fr == pytest.approx(2.121e8, 0.05)
    # analytic TE101 of a 1 m cube: c/(2) * sqrt(1/Lx^2 + 1/Lz^2)
    assert fr == pytest.approx(c_light / 2 * math.sqrt(2), 0.05)


# -------------------------------------------------------------- sources
def test_beam_line_integrates_to_q():
    x, y, z = fo.uniform_axes(-1e-2, 1e-2, -1e-2, 1e-2, -0.2, 0.2, 4, 4, 400)
    so = fo.OracleSolver(x, y, z, ["pec"] * 3, ["pec"] * 3)
    b = fo.OracleBeam(q=2e-9, sigmaz=10e-3, beta=0.9)
    # bunch centre at z = 0: s - s0 = 0 at z = v t + zmin - v ti
    zmin = so.z.min()
    t = (0.0 - zmin + b.v * b.ti) / b.v
    b.update(so, t)
    Jz = fo.comp(so.J, so.dims, "z")[b.ixs, b.iys, :]
    charge = np.sum(Jz * so.dx[b.ixs] * so.dy[b.iys] * so.dz) / b.v
    assert charge == pytest.approx(2e-9, rel=1e-9)
    assert so.z[np.argmax(Jz)] == pytest.approx(0.0, abs=so.dz[0])


@pytest.mark.parametrize("bcs", [
    (["abc", "pec", "pec"], ["abc", "pec", "pec"]),          # x high ABC
    (["pec", "periodic", "abc"], ["pec", "periodic", "pec"]),  # y periodic, z low ABC only
    (["pec", "pec", "pec"], ["pec", "pec", "abc"]),          # z high ABC
    (["periodic", "pec", "periodic"], ["periodic", "pec", "periodic"]),
])
def test_high_face_metric_copies_closed_form(bcs):
    """Periodic (both faces of an axis) and ABC (high face) set the dual length
    of the last layer along axis a to the first primal length, and the dual
    areas of the other two components there to the first layer's primal
    inverse areas (boundaries.py:35-70, 152-181):
    tL_a[-1] = d_a[0], itA_b[-1] = 1/(d_a[0] d_c), everything else as the
    grid metrics give. Non-uniform axes, so d[0] differs from d[-1]."""
    rng = np.random.default_rng(3)
    nodes = [np.cumsum(np.r_[0.0, rng.uniform(0.5, 1.5, n)]) for n in (5, 4, 6)]
    dims = (5, 4, 6)
    d = [np.diff(a) for a in nodes]
    L, iA, tL0, itA0 = fo.grid_metrics(*nodes)
    tL, itA, abc, _ = fo.bc_metric_edits(bcs[0], bcs[1], dims, L, iA, tL0, itA0)
    assert abc == any(b == "abc" for b in bcs[0] + bcs[1])
    copied = [bcs[1][a] == "abc" or (bcs[0][a] == "periodic" and bcs[1][a] == "periodic")
              for a in range(3)]
    exp_tL = [fo.comp(tL0, dims, c).copy() for c in range(3)]
    exp_itA = [fo.comp(itA0, dims, c).copy() for c in range(3)]
    for a in range(3):
        if not copied[a]:
            continue
        hi = [slice(None)] * 3
        hi[a] = -1
        hi = tuple(hi)
        exp_tL[a][hi] = d[a][0]
        for b in range(3):
            if b == a:
                continue
            c = 3 - a - b  # the third axis: the face of component b spans axes a and c
            sh = [1, 1]
            sh[c if c < a else c - 1] = dims[c]  # position of axis c in the slab at a = -1
            area = d[a][0] * d[c]
            exp_itA[b][hi] = np.broadcast_to((1.0 / area).reshape(sh), exp_itA[b][hi].shape)
    for c in range(3):
        np.testing.assert_allclose(fo.comp(tL, dims, c), exp_tL[c], rtol=1e-15, atol=0)
        np.testing.assert_allclose(fo.comp(itA, dims, c), exp_itA[c], rtol=1e-14, atol=0)
    # low-face ABC alone copies nothing
    if bcs == (["pec", "periodic", "abc"], ["pec", "periodic", "pec"]):
        assert np.array_equal(fo.comp(tL, dims, 2), fo.comp(tL0, dims, 2))


def test_pml_eps_and_sigma_layers_closed_form():
    """PML layers (boundaries.py:190-311) on a z-layered lossy dielectric
    (eps_r = 1 + k, sigma = 0.15 k S/m in layer k, varying across x so that the
    face-centre sample matters): the n low layers take the values of layer
    n + 1 at (Nx//2, Ny//2) (layer n keeps its own), the n high layers those of
    layer Nz - n - 1; sigma = sigma0 + max(1, sigma0) * geomspace(10, 5e-3, n)
    on the low side and geomspace(5e-3, 10, n) on the high side."""
    from scipy.constants import epsilon_0 as eps_0
    Nx, Ny, Nz, n = 6, 5, 14, 3
    x, y, z = fo.uniform_axes(0, 1, 0, 1, 0, 2, Nx, Ny, Nz)
    solids = []
    for k in range(Nz):
        for i in range(Nx):
            m = np.zeros((Nx, Ny, Nz), dtype=bool)
            m[i, :, k] = True
            solids.append((m, [1.0 + k + 0.25 * i, 1.0, 0.15 * k + 0.01 * i]))
    so = fo.OracleSolver(x, y, z, ["pec", "pec", "pml"], ["pec", "pec", "pml"],
                         solids=solids, n_pml=n)
    ic, jc = Nx // 2, Ny // 2

    def eps_r(i, k):
        return 1.0 + k + 0.25 * i

    def sig(i, k):
        return 0.15 * k + 0.01 * i

    prof_lo = np.geomspace(10.0, 5e-3, n)
    prof_hi = np.geomspace(5e-3, 10.0, n)
    for c in range(3):
        ie = fo.comp(so.ieps, so.dims, c)
        sg = fo.comp(so.sigma, so.dims, c)
        s_lo, s_hi = sig(ic, n + 1), sig(ic, Nz - n - 1)
        assert s_lo < 1 < s_hi  # both branches of max(1, sigma0)
        for m in range(n):
            np.testing.assert_allclose(ie[:, :, m], 1.0 / (eps_r(ic, n + 1) * eps_0), rtol=1e-15)
            np.testing.assert_allclose(sg[:, :, m], s_lo + 1.0 * prof_lo[m], rtol=1e-15)
            kh = Nz - 1 - m
            np.testing.assert_allclose(ie[:, :, kh], 1.0 / (eps_r(ic, Nz - n - 1) * eps_0),
                                       rtol=1e-15)
            np.testing.assert_allclose(sg[:, :, kh], s_hi + s_hi * prof_hi[n - 1 - m],
                                       rtol=1e-15)
        # layer n and the layers inside keep their own values
        for k in range(n, Nz - n):
            for i in range(Nx):
                np.testing.assert_allclose(ie[i, :, k], 1.0 / (eps_r(i, k) * eps_0), rtol=1e-15)
                np.testing.assert_allclose(sg[i, :, k], sig(i, k), rtol=1e-15)


def test_pml_profile_closed_form():
    x, y, z = fo.uniform_axes(0, 1, 0, 1, 0, 1, 6, 6, 40)
    n = 8
    so = fo.OracleSolver(x, y, z, ["pec", "pec", "pml"], ["pec", "pec", "pml"], n_pml=n)
    for c in range(3):
        sig = fo.comp(so.sigma, so.dims, c)[3, 3, :]
        lo, hi = sig[:n], sig[-n:]
        assert lo[0] == pytest.approx(10.0) and lo[-1] == pytest.approx(5e-3)
        assert hi[0] == pytest.approx(5e-3) and hi[-1] == pytest.approx(10.0)
        ratio = (5e-3 / 10.0) ** (1.0 / (n - 1))
        np.testing.assert_allclose(lo[1:] / lo[:-1], ratio, rtol=1e-12)
        assert np.all(sig[n:-n] == 0)
    assert so.use_conductivity


# --------------------------------------------------- wake post-processing
def _synthetic_capture(Nz=40, nt=300, dtype=np.float64, fill=None):
    z_nodes = np.linspace(-0.05, 0.05, Nz + 1)
    zf = z_nodes[:-1] + np.diff(z_nodes) / 2
    dt = 1e-12
    t = np.arange(0, nt * dt, dt)
    Ez = np.zeros((nt, 3, 3, Nz), dtype=dtype) if fill is None else fill
    return dict(Ez=Ez, t=t, zf=zf, dz=np.diff(z_nodes), xf=np.array([-1e-3, 0.0, 1e-3]),
                yf=np.array([-1e-3, 0.0, 1e-3]))


def test_wake_constant_field_closed_form():
    cap = _synthetic_capture(nt=1000)
    E0 = 3.0
    cap["Ez"][:] = E0
    q, v, ti = 1e-9, c_light, 2e-11
    s, WP, WP3 = wo.long_WP_3d(cap, q, v, ti, wakelength=0.01)
    Q = q * 1e12
    dt = cap["t"][2] - cap["t"][1]
    zmin = cap["zf"].min()
    # every (n,k) with a sample index inside the record contributes E0*dz[k]
    full = np.array([np.all(((cap["zf"] + sn) / v - zmin / v + ti) / dt < len(cap["t"]))
                     for sn in s])
    assert full.sum() > 10
    S = E0 * np.sum(cap["dz"])
    np.testing.assert_allclose(WP3[0, 0, full], S / Q, rtol=1e-13)
    # the running WP is carried from one line to the next (wakeSolver.py:485, 536-537)
    np.testing.assert_allclose(WP3[0, 1, full], (S / Q + S) / Q, rtol=1e-13)


def test_wake_impulse_lands_at_its_retarded_position():
    """An Ez impulse at (z_k, t_m) contributes to the one s with
    s in [v(t_m - ti) + zmin - z_k, v(t_{m+1} - ti) + zmin - z_k)."""
    cap = _synthetic_capture()
    k0, m0 = 17, 250
    cap["Ez"][m0, 1, 1, k0] = 1.0
    q, v, ti = 1e-9, c_light, 2e-11
    s, WP, WP3 = wo.long_WP_3d(cap, q, v, ti, wakelength=0.05)
    hits = np.flatnonzero(WP)
    assert len(hits) == 1
    dt = cap["t"][2] - cap["t"][1]
    zmin = cap["zf"].min()
    lo = v * (m0 * dt - ti) + zmin - cap["zf"][k0]
    assert lo - 1e-9 <= s[hits[0]] < lo + v * dt + 1e-9
    assert WP[hits[0]] == pytest.approx(cap["dz"][k0] / (q * 1e12), rel=1e-14)


def test_wake_counter_moving_constant_field_closed_form():
    """Counter-moving (wakeSolver.py:503-517): a constant Ez = E0 gives
    -E0 * sum(dz) / (q*1e12) for every s whose samples all lie in the record,
    with the line-to-line carry; the time of row z is (zmax - z + s)/v + ti."""
    cap = _synthetic_capture(nt=1000)
    E0 = 3.0
    cap["Ez"][:] = E0
    q, v, ti = 1e-9, c_light, 2e-11
    s, WP, WP3 = wo.long_WP_3d(cap, q, v, ti, wakelength=0.01, counter_moving=True)
    Q = q * 1e12
    dt = cap["t"][2] - cap["t"][1]
    zmax = cap["zf"].max()
    full = np.array([np.all(((zmax - cap["zf"] + sn) / v + ti) / dt < len(cap["t"]))
                     for sn in s])
    assert full.sum() > 10
    S = -E0 * np.sum(cap["dz"])
    np.testing.assert_allclose(WP3[0, 0, full], S / Q, rtol=1e-13)
    np.testing.assert_allclose(WP3[0, 1, full], (S / Q + S) / Q, rtol=1e-13)


def test_wake_counter_moving_impulse_position_and_weight():
    """An impulse at (z_k, t_m) reaches the one s with
    (zmax - z_k + s)/v + ti in [t_m, t_m+1); its weight is -dz[nz-1-k] (the
    reference weights row -k-1 with dz[k]), distinct from -dz[k] on this
    non-uniform axis."""
    rng = np.random.default_rng(4)
    Nz, nt = 40, 300
    zn = np.cumsum(np.r_[-0.02, rng.uniform(0.5e-3, 1.5e-3, Nz)])
    cap = _synthetic_capture(Nz=Nz, nt=nt)
    cap["zf"], cap["dz"] = zn[:-1] + np.diff(zn) / 2, np.diff(zn)
    k0, m0 = 11, 180
    cap["Ez"][m0, 1, 1, k0] = 1.0
    q, v, ti = 1e-9, c_light, 2e-11
    s, WP, _ = wo.long_WP_3d(cap, q, v, ti, wakelength=0.05, counter_moving=True)
    hits = np.flatnonzero(WP)
    assert len(hits) == 1
    dt = cap["t"][2] - cap["t"][1]
    lo = v * (m0 * dt - ti) - cap["zf"].max() + cap["zf"][k0]
    assert lo - 1e-9 <= s[hits[0]] < lo + v * dt + 1e-9
    w = -cap["dz"][Nz - 1 - k0] / (q * 1e12)
    assert WP[hits[0]] == pytest.approx(w, rel=1e-14)
    assert abs(cap["dz"][k0] - cap["dz"][Nz - 1 - k0]) > 1e-5


def test_long_wp_of_a_window_reads_the_plus_one_line():
    """calc_long_WP on a 3x3 window integrates line [2, 2]
    (Ez.shape//2 + 1, wakeSolver.py:385-388): constant fields of distinct
    value per line give that line's closed form."""
    cap = _synthetic_capture(nt=1000)
    vals = np.arange(1.0, 10.0).reshape(3, 3)
    cap["Ez"][:] = vals[None, :, :, None]
    q, v, ti = 1e-9, c_light, 2e-11
    s, WP = wo.long_WP_1d(cap, q, v, ti, wakelength=0.01)
    dt = cap["t"][2] - cap["t"][1]
    zmin = cap["zf"].min()
    full = np.array([np.all(((cap["zf"] + sn) / v - zmin / v + ti) / dt < len(cap["t"]))
                     for sn in s])
    assert full.sum() > 10
    np.testing.assert_allclose(WP[full], vals[2, 2] * np.sum(cap["dz"]) / (q * 1e12),
                               rtol=1e-13)


def test_long_wp_from_ezt_equals_the_line_capture():
    """The integral over a given Ezt equals calc_long_WP over a one-line
    capture holding the same samples, and a constant Ezt gives the closed form."""
    rng = np.random.default_rng(2)
    cap = _synthetic_capture(nt=400)
    line = rng.standard_normal((400, cap["zf"].size))
    q, v, ti = 1e-9, c_light, 2e-11
    s1, WP1 = wo.long_WP_1d(dict(cap, Ez=line), q, v, ti, wakelength=0.02)
    s2, WP2 = wo.long_WP_from_Ezt(line.T, cap["t"], cap["zf"], cap["dz"], q, v, ti, 0.02)
    assert np.array_equal(s1, s2) and np.array_equal(WP1, WP2)
    s3, WP3 = wo.long_WP_from_Ezt(np.full_like(line.T, 2.0), cap["t"], cap["zf"], cap["dz"], q,
                                  v, ti, 0.02)
    dt = cap["t"][2] - cap["t"][1]
    zmin = cap["zf"].min()
    full = np.array([np.all(((cap["zf"] + sn) / v - zmin / v + ti) / dt < 400) for sn in s3])
    np.testing.assert_allclose(WP3[full], 2.0 * np.sum(cap["dz"]) / (q * 1e12), rtol=1e-13)


def test_trans_wp_of_a_linear_wake_is_constant_slope():
    """WP_3d = a*x + b*y + c(s) gives WPx[n] = -a * n * ds, WPy[n] = -b * n * ds."""
    s = np.linspace(-0.01, 0.05, 200)
    xf = np.array([-2e-3, 0.0, 2e-3])
    yf = np.array([-1e-3, 0.0, 1e-3])
    a, b = 3.0e3, -5.0e2
    c = np.sin(s * 100)
    WP3 = a * xf[:, None, None] + b * yf[None, :, None] + c[None, None, :]
    WPx, WPy = wo.trans_WP(s, WP3, xf, yf)
    ds = s[2] - s[1]
    n = np.arange(len(s))
    np.testing.assert_allclose(WPx, -a * n * ds, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(WPy, -b * n * ds, rtol=1e-10, atol=1e-12)


def test_impedance_bins_are_the_textbook_dft():
    rng = np.random.default_rng(2)
    v = c_light
    s = np.arange(-0.02, 0.03, 2e-4)
    WP = rng.standard_normal(len(s)) * 1e-12
    lam = wo.lambdas_analytic(s, 5e-3)
    fmax = v / (3 * 5e-3)
    f, Z, lamf = wo.long_Z(s, WP, lam, v, fmax, samples=101)
    ds, N = wo.dft_setup(s, v, fmax, samples=101)
    n = np.arange(len(s))
    kk = np.arange(len(f))
    tw = np.exp(-2j * np.pi * np.outer(kk, n) / N)
    np.testing.assert_allclose(f, kk * v / (N * ds), rtol=1e-12)
    ref_lf = tw @ (lam * v) * ds
    ref_Z = -(tw @ (WP * 1e12) * ds) / ref_lf
    np.testing.assert_allclose(lamf, ref_lf, rtol=1e-10, atol=1e-10 * np.abs(ref_lf).max())
    np.testing.assert_allclose(Z, ref_Z, rtol=1e-9, atol=1e-9 * np.abs(ref_Z).max())


def test_gaussian_spectrum_closed_form():
    """|lambda(f)| of a normalised Gaussian bunch is v*exp(-(2 pi f sigma/v)^2/2)."""
    v = c_light
    sig = 5e-3
    s = np.arange(-0.05, 0.1, 1e-4)
    lam = wo.lambdas_analytic(s, sig)
    fmax = v / (3 * sig)
    f, _, lamf = wo.long_Z(s, np.zeros_like(s), lam, v, fmax, samples=101)
    ref = v * np.exp(-((2 * np.pi * f * sig / v) ** 2) / 2)
    np.testing.assert_allclose(np.abs(lamf), ref, rtol=1e-6, atol=1e-9 * v)


# ------------------------------------------------ test_001 golden vectors
@pytest.mark.slow
def test_oracle_reproduces_test001_golden():
    """The reference's test_001 (FP32 PEC cubic cavity, 5184 steps) golden
    WP[::50], Z[::20] and cumulative sums, through the oracle."""
    s, g = cases.test001_solver(np.float32)
    b = g["beam"]
    ti = wo.wake_ti(b["sigmaz"], b["beta"])
    cap = fo.wakesolve_fields(s, b["wakelength"], b["q"], b["sigmaz"], b["beta"],
                              b["xs"], b["ys"], b["xt"], b["yt"], ti)
    assert cap["Nt"] == 5184
    out = wo.wake_solve(cap, b["q"], b["sigmaz"], b["beta"], b["wakelength"],
                        skip_cells=b["skip_cells"])
    WPg = np.array(g["WP_every50"])
    Zg = np.array(g["Z_every20_re_im"])
    Zg = Zg[:, 0] + 1j * Zg[:, 1]
    assert np.allclose(out["WP"][::50], WPg, **g["tol_WP"])
    for part in (np.abs, np.real, np.imag):
        assert np.allclose(part(out["Z"])[::20], part(Zg), **g["tol_Z"])
    assert np.cumsum(np.abs(out["WP"]))[-1] == pytest.approx(g["cumsum_abs_WP"],
                                                             g["cumsum_rel"])
    assert np.cumsum(np.abs(out["Z"]))[-1] == pytest.approx(g["cumsum_abs_Z"],
                                                            g["cumsum_rel"])
