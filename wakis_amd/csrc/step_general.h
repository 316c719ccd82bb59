// Exact general leapfrog path: one thread per cell, H sweep then E sweep.
//
// Restates SolverFIT3D._one_step (solverFIT3D.py:482-503) on the flat index
// without materialising M1/M2. Row r of M1 = iDa*iDmu*C*Ds has entries
// +-fl(a_r*L_c), r's stencil columns in the CSR order scipy produces
// (ascending, or descending when exactly one of the PEC column mask and the
// PMC row mask was applied); row r of M2 = iDeps*itDa*C^T*tDs has entries
// +-fl(b_r*tL_c) in ascending order. Terms are accumulated from 0 in that
// order without contraction, H <- H - sum fl(fl(dt*m)*E), E <- E + fl(dt*(S2
// - fl(ieps*J))), so results match scipy's csr_matvec bit for bit.
#pragma once
#include "fit_common.h"

namespace wf {

struct Term {
  Cell c;
  int comp;
  int sign;
  bool ok;
};

__device__ __forceinline__ Cell cell_of(int64_t n, const GridDev &g) {
  Cell c;
  c.k = (int)(n / g.sz);
  int64_t r = n - (int64_t)c.k * g.sz;
  c.j = (int)(r / g.Nx);
  c.i = (int)(r - (int64_t)c.j * g.Nx);
  return c;
}

// Row terms of C (H update, forward differences), ascending column order.
__device__ __forceinline__ void h_terms(int comp, const Cell &c0, const GridDev &g,
                                        Term (&t)[4]) {
  Cell a = c0, b = c0;
  if (comp == 0) {  // [0, -Pz, Py]: Ey[n]+, Ey[n+sz]-, Ez[n]-, Ez[n+sy]+
    bool oa = fwd_z(a, g), ob = fwd_y(b, g);
    t[0] = {c0, 1, +1, true}; t[1] = {a, 1, -1, oa};
    t[2] = {c0, 2, -1, true}; t[3] = {b, 2, +1, ob};
  } else if (comp == 1) {  // [Pz, 0, -Px]: Ex[n]-, Ex[n+sz]+, Ez[n]+, Ez[n+1]-
    bool oa = fwd_z(a, g), ob = fwd_x(b, g);
    t[0] = {c0, 0, -1, true}; t[1] = {a, 0, +1, oa};
    t[2] = {c0, 2, +1, true}; t[3] = {b, 2, -1, ob};
  } else {  // [-Py, Px, 0]: Ex[n]+, Ex[n+sy]-, Ey[n]-, Ey[n+1]+
    bool oa = fwd_y(a, g), ob = fwd_x(b, g);
    t[0] = {c0, 0, +1, true}; t[1] = {a, 0, -1, oa};
    t[2] = {c0, 1, -1, true}; t[3] = {b, 1, +1, ob};
  }
}

// Row terms of C^T (E update, backward differences), ascending column order.
__device__ __forceinline__ void e_terms(int comp, const Cell &c0, const GridDev &g,
                                        Term (&t)[4]) {
  Cell a = c0, b = c0;
  if (comp == 0) {  // [0, Pz^T, -Py^T]: Hy[n-sz]+, Hy[n]-, Hz[n-sy]-, Hz[n]+
    bool oa = bwd_z(a, g), ob = bwd_y(b, g);
    t[0] = {a, 1, +1, oa}; t[1] = {c0, 1, -1, true};
    t[2] = {b, 2, -1, ob}; t[3] = {c0, 2, +1, true};
  } else if (comp == 1) {  // [-Pz^T, 0, Px^T]: Hx[n-sz]-, Hx[n]+, Hz[n-1]+, Hz[n]-
    bool oa = bwd_z(a, g), ob = bwd_x(b, g);
    t[0] = {a, 0, -1, oa}; t[1] = {c0, 0, +1, true};
    t[2] = {b, 2, +1, ob}; t[3] = {c0, 2, -1, true};
  } else {  // [Py^T, -Px^T, 0]: Hx[n-sy]+, Hx[n]-, Hy[n-1]-, Hy[n]+
    bool oa = bwd_y(a, g), ob = bwd_x(b, g);
    t[0] = {a, 0, +1, oa}; t[1] = {c0, 0, -1, true};
    t[2] = {b, 1, -1, ob}; t[3] = {c0, 1, +1, true};
  }
}

template <typename S, typename Cm>
__global__ void __launch_bounds__(256)
k_h_general(GridDev g, MetricDev<S> m, S *__restrict__ H, const S *__restrict__ E,
            const S *__restrict__ a, Cm dt) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.N) return;
  const Cell c0 = cell_of(n, g);
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    Term t[4];
    h_terms(comp, c0, g, t);
    const S ar = a[comp * g.N + n];
    Cm acc = Cm(0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const Term &tt = t[g.m1desc ? 3 - q : q];
      if (!tt.ok) continue;
      if (!face_mask(g.pecz, tt.comp, tt.c.i, tt.c.j, tt.c.k, g.Nx, g.Ny, g.Nz)) continue;
      const S mm = ar * len_of(m.Lx, m.Ly, m.Lz, tt.comp, tt.c);
      if (mm == S(0)) continue;
      const Cm w = dt * (Cm)mm;
      const Cm p = w * (Cm)E[tt.comp * g.N + flat(tt.c, g)];
      acc = (tt.sign > 0) ? acc + p : acc - p;
    }
    const int64_t r = comp * g.N + n;
    H[r] = (S)((Cm)H[r] - acc);
  }
}

struct EStepAux {
  int jmode, ixs, iys;
  int cap_ix0, cap_iy0, cap_nx, cap_ny;
  int64_t cap_samples, cap_sample;
};

template <typename S, typename Cm, int JM, bool COND>
__global__ void __launch_bounds__(256)
k_e_general(GridDev g, MetricDev<S> m, S *__restrict__ E, const S *__restrict__ H,
            S *__restrict__ J, const S *__restrict__ b, const S *__restrict__ ieps,
            const S *__restrict__ sigma, Cm dt, EStepAux x, const S *__restrict__ beam_row,
            S *__restrict__ cap) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.N) return;
  const Cell c0 = cell_of(n, g);
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    Term t[4];
    e_terms(comp, c0, g, t);
    const int64_t r = comp * g.N + n;
    const S br = b[r];
    S acc = S(0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const Term &tt = t[q];
      if (!tt.ok) continue;
      if (!face_mask(g.pmcz, tt.comp, tt.c.i, tt.c.j, tt.c.k, g.Nx, g.Ny, g.Nz)) continue;
      const S mm = br * len_of(m.tLx, m.tLy, m.tLz, tt.comp, tt.c);
      if (mm == S(0)) continue;
      const S p = mm * H[tt.comp * g.N + flat(tt.c, g)];
      acc = (tt.sign > 0) ? acc + p : acc - p;
    }
    const bool online = (comp == 2) && (c0.i == x.ixs) && (c0.j == x.iys);
    S d = acc;
    S jv = S(0);
    if (JM == 2) {
      jv = (online && beam_row) ? beam_row[c0.k] : J[r];
      d = acc - ieps[r] * jv;
    } else if (JM == 1 && online) {
      jv = beam_row ? beam_row[c0.k] : J[r];
      d = acc - ieps[r] * jv;
    }
    const S en = (S)((Cm)E[r] + dt * (Cm)d);
    E[r] = en;
    if (COND) {
      J[r] = sigma[r] * en;
    } else if (online && beam_row) {
      J[r] = jv;
    }
    if (cap && comp == 2) {
      const int li = c0.i - x.cap_ix0, lj = c0.j - x.cap_iy0;
      if (li >= 0 && li < x.cap_nx && lj >= 0 && lj < x.cap_ny && x.cap_sample < x.cap_samples) {
        const int64_t l = (int64_t)li * x.cap_ny + lj;
        cap[(l * x.cap_samples + x.cap_sample) * g.Nz + c0.k] = en;
      }
    }
  }
}

template <typename S>
__global__ void k_zero_ghosts(GridDev g, S *__restrict__ E, S *__restrict__ H) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.N) return;
  const Cell c = cell_of(n, g);
  // solverFIT3D.py:820-840: H tangential and E normal on the -1 layers
  if (c.i == g.Nx - 1) { H[g.N + n] = S(0); H[2 * g.N + n] = S(0); E[n] = S(0); }
  if (c.j == g.Ny - 1) { H[n] = S(0); H[2 * g.N + n] = S(0); E[g.N + n] = S(0); }
  if (c.k == g.Nz - 1) { H[n] = S(0); H[g.N + n] = S(0); E[2 * g.N + n] = S(0); }
}

template <typename S>
__global__ void k_capture(GridDev g, const S *__restrict__ E, S *__restrict__ cap, int ix0,
                          int iy0, int nx, int ny, int64_t samples, int64_t sample) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nl = (int64_t)nx * ny;
  if (t >= nl * g.Nz) return;
  const int64_t l = t / g.Nz;
  const int k = (int)(t - l * g.Nz);
  const int i = ix0 + (int)(l / ny), j = iy0 + (int)(l % ny);
  cap[(l * samples + sample) * g.Nz + k] = E[2 * g.N + i + g.sy * j + g.sz * k];
}

// Copy plane 1 -> plane 0 of all components of E and H across one axis.
template <typename S>
__global__ void k_abc_face(GridDev g, int axis, S *__restrict__ E, S *__restrict__ H) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t na, nb;
  if (axis == 0) { na = g.Ny; nb = g.Nz; }
  else if (axis == 1) { na = g.Nx; nb = g.Nz; }
  else { na = g.Nx; nb = g.Ny; }
  if (t >= na * nb) return;
  const int64_t u = t % na, w = t / na;
  int64_t n0, n1;
  if (axis == 0) { n0 = g.sy * u + g.sz * w; n1 = n0 + 1; }
  else if (axis == 1) { n0 = u + g.sz * w; n1 = n0 + g.sy; }
  else { n0 = u + g.sy * w; n1 = n0 + g.sz; }
  for (int c = 0; c < 3; ++c) {
    E[c * g.N + n0] = E[c * g.N + n1];
    H[c * g.N + n0] = H[c * g.N + n1];
  }
}

}  // namespace wf
