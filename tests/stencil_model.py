"""Numpy model of the HIP general-path stencil (tests only).

Restates, vectorised over cells, the exact term tables, face-mask rule,
flat-index wrap/truncation and summation order that
``wakis_amd/csrc/step_general.h`` implements, so the stencil definition can be
checked bit for bit against the oracle's scipy CSR product on the CPU.
"""

import numpy as np

# (component of the neighbour, neighbour kind, sign) in ascending column order
H_TERMS = {0: [(1, "self", +1), (1, "fz", -1), (2, "self", -1), (2, "fy", +1)],
           1: [(0, "self", -1), (0, "fz", +1), (2, "self", +1), (2, "fx", -1)],
           2: [(0, "self", +1), (0, "fy", -1), (1, "self", -1), (1, "fx", +1)]}
E_TERMS = {0: [(1, "bz", +1), (1, "self", -1), (2, "by", -1), (2, "self", +1)],
           1: [(0, "bz", -1), (0, "self", +1), (2, "bx", +1), (2, "self", -1)],
           2: [(0, "by", +1), (0, "self", -1), (1, "bx", -1), (1, "self", +1)]}


def cells(dims):
    Nx, Ny, Nz = dims
    n = np.arange(Nx * Ny * Nz)
    k = n // (Nx * Ny)
    r = n - k * Nx * Ny
    j = r // Nx
    i = r - j * Nx
    return i, j, k


def neighbour(kind, i, j, k, dims):
    """Flat-index neighbour of diags([-1,1],[0,s]) (forward) or its transpose."""
    Nx, Ny, Nz = dims
    i, j, k = i.copy(), j.copy(), k.copy()
    ok = np.ones(i.shape, bool)
    if kind == "self":
        return i, j, k, ok
    if kind == "fz":
        ok = k < Nz - 1
        k = k + 1
    elif kind == "bz":
        ok = k > 0
        k = k - 1
    elif kind == "fy":
        wrap = j == Ny - 1
        ok = ~wrap | (k < Nz - 1)
        j = np.where(wrap, 0, j + 1)
        k = np.where(wrap, k + 1, k)
    elif kind == "by":
        wrap = j == 0
        ok = ~wrap | (k > 0)
        j = np.where(wrap, Ny - 1, j - 1)
        k = np.where(wrap, k - 1, k)
    elif kind == "fx":
        wx = i == Nx - 1
        wy = wx & (j == Ny - 1)
        ok = ~wy | (k < Nz - 1)
        i = np.where(wx, 0, i + 1)
        j = np.where(wx, np.where(wy, 0, j + 1), j)
        k = np.where(wy, k + 1, k)
    elif kind == "bx":
        wx = i == 0
        wy = wx & (j == 0)
        ok = ~wy | (k > 0)
        i = np.where(wx, Nx - 1, i - 1)
        j = np.where(wx, np.where(wy, Ny - 1, j - 1), j)
        k = np.where(wy, k - 1, k)
    ok &= (i >= 0) & (i < Nx) & (j >= 0) & (j < Ny) & (k >= 0) & (k < Nz)
    return np.clip(i, 0, Nx - 1), np.clip(j, 0, Ny - 1), np.clip(k, 0, Nz - 1), ok


def face_mask(zbits, c, i, j, k, dims):
    Nx, Ny, Nz = dims
    f = np.full(i.shape, -1)
    order = {0: [(j, 0, 1), (j, Ny - 1, 4), (k, 0, 2), (k, Nz - 1, 5)],
             1: [(i, 0, 0), (i, Nx - 1, 3), (k, 0, 2), (k, Nz - 1, 5)],
             2: [(i, 0, 0), (i, Nx - 1, 3), (j, 0, 1), (j, Ny - 1, 4)]}[c]
    for idx, val, face in order:
        f = np.where(idx == val, face, f)
    zero = np.zeros(i.shape, bool)
    for face in range(6):
        if (zbits >> face) & 1:
            zero |= f == face
    return ~zero


def step(E, H, J, a, b, ieps, sigma, dt, dims, L1, tL1, pecz, pmcz, m1desc, jfull, cond):
    """One step of the general path in place (returns nothing)."""
    Nx, Ny, Nz = dims
    N = Nx * Ny * Nz
    S = E.dtype.type
    i, j, k = cells(dims)
    dtS = S(dt)
    for comp in range(3):
        terms = H_TERMS[comp][::-1] if m1desc else H_TERMS[comp]
        ar = a[comp * N:(comp + 1) * N]
        acc = np.zeros(N, dtype=S)
        for (cc, kind, sign) in terms:
            ii, jj, kk, ok = neighbour(kind, i, j, k, dims)
            Lc = (L1[cc].astype(S))[(ii, jj, kk)[cc]]
            m = ar * Lc
            valid = ok & face_mask(pecz, cc, ii, jj, kk, dims) & (m != 0)
            idx = cc * N + ii + Nx * jj + Nx * Ny * kk
            p = (dtS * m) * E[idx]
            acc = np.where(valid, acc + p if sign > 0 else acc - p, acc)
        H[comp * N:(comp + 1) * N] = H[comp * N:(comp + 1) * N] - acc
    for comp in range(3):
        br = b[comp * N:(comp + 1) * N]
        acc = np.zeros(N, dtype=S)
        for (cc, kind, sign) in E_TERMS[comp]:
            ii, jj, kk, ok = neighbour(kind, i, j, k, dims)
            tLc = (tL1[cc].astype(S))[(ii, jj, kk)[cc]]
            m = br * tLc
            valid = ok & face_mask(pmcz, cc, ii, jj, kk, dims) & (m != 0)
            idx = cc * N + ii + Nx * jj + Nx * Ny * kk
            p = m * H[idx]
            acc = np.where(valid, acc + p if sign > 0 else acc - p, acc)
        sl = slice(comp * N, (comp + 1) * N)
        d = acc - ieps[sl] * J[sl] if jfull else acc
        E[sl] = E[sl] + dtS * d
        if cond:
            J[sl] = sigma[sl] * E[sl]
