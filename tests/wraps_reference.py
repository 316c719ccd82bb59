"""The fused path's wrap-term check written with gathered index vectors (tests
only): every wrap row's column is built from explicit (i, j, k) coordinates.
``SolverFIT3D._wraps_dead_now`` forms the same test on boundary planes shifted
by one flat index; the GPU and CPU tests compare the two."""

import torch


def wraps_dead_gather(s):
    Nx, Ny, Nz, N = s.Nx, s.Ny, s.Nz, s.N
    dev = s.device
    ar = torch.arange
    a, b = s._a, s._b
    E, H, J, ie = s.E.array, s.H.array, s.J.array, s.ieps.array
    bl = s._beam_line

    def lin(i, j, k):
        return i + Nx * j + Nx * Ny * k

    def fmask(zbits, comp, i, j, k):
        if zbits == 0:
            return torch.ones_like(i, dtype=torch.bool)
        f = torch.full_like(i, -1)
        order = {0: [(j, 0, 1), (j, Ny - 1, 4), (k, 0, 2), (k, Nz - 1, 5)],
                 1: [(i, 0, 0), (i, Nx - 1, 3), (k, 0, 2), (k, Nz - 1, 5)],
                 2: [(i, 0, 0), (i, Nx - 1, 3), (j, 0, 1), (j, Ny - 1, 4)]}[comp]
        for idx, val, face in order:
            f = torch.where(idx == val, torch.full_like(f, face), f)
        zero = torch.zeros_like(i, dtype=torch.bool)
        for face in range(6):
            if (zbits >> face) & 1:
                zero |= f == face
        return ~zero

    def frozen_E(c, i, j, k):
        t = c * N + lin(i, j, k)
        fz = (b[t] == 0) & (E[t] == 0) & ((ie[t] == 0) | (J[t] == 0))
        if bl is not None and c == 2:
            fz &= ~((i == bl[0]) & (j == bl[1]))
        return fz

    def frozen_H(c, i, j, k):
        t = c * N + lin(i, j, k)
        return (a[t] == 0) & (H[t] == 0)

    bad = torch.zeros((), dtype=torch.bool, device=dev)
    # (1) H_y, H_z rows at i=Nx-1 use E_z, E_y at (0, j+1, k) / (0, 0, k+1)
    jj, kk = torch.meshgrid(ar(Ny, device=dev), ar(Nz, device=dev), indexing="ij")
    jj, kk = jj.reshape(-1), kk.reshape(-1)
    valid = ~((jj == Ny - 1) & (kk == Nz - 1))
    tj = torch.where(jj < Ny - 1, jj + 1, torch.zeros_like(jj))
    tk = torch.where(jj < Ny - 1, kk, kk + 1).clamp(max=Nz - 1)
    ti = torch.zeros_like(jj)
    rn = lin(torch.full_like(jj, Nx - 1), jj, kk)
    for rc, cc in ((1, 2), (2, 1)):
        w = valid & (a[rc * N + rn] != 0) & fmask(s.pec_zero, cc, ti, tj, tk)
        bad |= (w & ~frozen_E(cc, ti, tj, tk)).any()
    # (2) H_x, H_z rows at j=Ny-1 use E_z, E_x at (i, 0, k+1)
    ii, kk = torch.meshgrid(ar(Nx, device=dev), ar(Nz - 1, device=dev), indexing="ij")
    ii, kk = ii.reshape(-1), kk.reshape(-1)
    rn = lin(ii, torch.full_like(ii, Ny - 1), kk)
    tj, tk = torch.zeros_like(ii), kk + 1
    for rc, cc in ((0, 2), (2, 0)):
        w = (a[rc * N + rn] != 0) & fmask(s.pec_zero, cc, ii, tj, tk)
        bad |= (w & ~frozen_E(cc, ii, tj, tk)).any()
    # (3) E_y, E_z rows at i=0 use H_z, H_y at (Nx-1, j-1, k) / (Nx-1, Ny-1, k-1)
    jj, kk = torch.meshgrid(ar(Ny, device=dev), ar(Nz, device=dev), indexing="ij")
    jj, kk = jj.reshape(-1), kk.reshape(-1)
    valid = ~((jj == 0) & (kk == 0))
    sj = torch.where(jj > 0, jj - 1, torch.full_like(jj, Ny - 1))
    sk = torch.where(jj > 0, kk, kk - 1).clamp(min=0)
    si = torch.full_like(jj, Nx - 1)
    rn = lin(torch.zeros_like(jj), jj, kk)
    tLz = torch.as_tensor(s._tL1[2], device=dev)
    tLy = torch.as_tensor(s._tL1[1], device=dev)
    tLx = torch.as_tensor(s._tL1[0], device=dev)
    for rc, cc, tl in ((1, 2, tLz[sk]), (2, 1, tLy[sj])):
        w = valid & (b[rc * N + rn] != 0) & (tl != 0) & fmask(s.pmc_zero, cc, si, sj, sk)
        bad |= (w & ~frozen_H(cc, si, sj, sk)).any()
    # (4) E_x, E_z rows at j=0 use H_z, H_x at (i, Ny-1, k-1)
    ii, kk = torch.meshgrid(ar(Nx, device=dev), ar(1, Nz, device=dev), indexing="ij")
    ii, kk = ii.reshape(-1), kk.reshape(-1)
    rn = lin(ii, torch.zeros_like(ii), kk)
    sj, sk = torch.full_like(ii, Ny - 1), kk - 1
    for rc, cc, tl in ((0, 2, tLz[sk]), (2, 0, tLx[ii])):
        w = (b[rc * N + rn] != 0) & (tl != 0) & fmask(s.pmc_zero, cc, ii, sj, sk)
        bad |= (w & ~frozen_H(cc, ii, sj, sk)).any()
    return not bool(bad.item())
