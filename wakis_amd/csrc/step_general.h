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

// Launch geometry of the exact kernels: 64 x 4 cells of one z plane per
// workgroup, the grid's z dimension striding over the planes (so no cell
// index is divided out of a flat index).
constexpr int EX_TX = 64, EX_TY = 4;

// One term of a row: its sign s, the neighbour value v (0 when the column is
// absent or masked), the metric weight w = fl(coef * L) of its column, and
// whether the column exists in the CSR row at all. A term whose weight is an
// exact zero is skipped, as the stored-entry loop of the reference skips it.
template <typename S, typename Cm>
__device__ __forceinline__ void add_h_term(Cm &acc, Cm dt, S w, S v, bool ok, int sign) {
  if (!ok || w == S(0)) return;
  const Cm p = (dt * (Cm)w) * (Cm)v;
  acc = (sign > 0) ? acc + p : acc - p;
}
template <typename S>
__device__ __forceinline__ void add_e_term(S &acc, S w, S v, bool ok, int sign) {
  if (!ok || w == S(0)) return;
  const S p = w * v;
  acc = (sign > 0) ? acc + p : acc - p;
}

// H sweep of the exact path: every row of M1 with the reference operator's
// flat-index neighbours (n+1, n+sy, n+sz; the x and y ones wrap into the next
// row/plane, all are dropped past the array end), the PEC column mask, terms
// in CSR column order (ascending, or descending when m1desc).
template <typename S, typename Cm>
__global__ void __launch_bounds__(EX_TX * EX_TY)
k_h_exact(GridDev g, MetricDev<S> m, S *__restrict__ H, const S *__restrict__ E,
          const S *__restrict__ a, Cm dt) {
  const int i = blockIdx.x * EX_TX + threadIdx.x, j = blockIdx.y * EX_TY + threadIdx.y;
  if (i >= g.Nx || j >= g.Ny) return;
  const int64_t N = g.N;
  // x+ neighbour cell (flat n+1) and y+ neighbour cell (flat n+sy) of row j
  const bool xin = i < g.Nx - 1, yin = j < g.Ny - 1;
  const int xi = xin ? i + 1 : 0, xj = xin ? j : (yin ? j + 1 : 0);
  const S Lxi = m.Lx[i], Lyj = m.Ly[j], Lyx = m.Ly[xj];
  for (int k = blockIdx.z; k < g.Nz; k += gridDim.z) {
    const int64_t n = (int64_t)i + g.sy * j + g.sz * k;
    const int xk = (xin || yin) ? k : k + 1, yk = yin ? k : k + 1;
    const bool xok = xk < g.Nz, yok = yk < g.Nz, zok = k + 1 < g.Nz;
    const int yj = yin ? j + 1 : 0;
    const S Lzk = m.Lz[k];
    const S Lzx = xok ? m.Lz[xk] : S(0), Lzy = yok ? m.Lz[yk] : S(0);
    // neighbour values with the PEC column mask (a masked column's term is
    // not stored; its value is taken as 0 and its weight test skipped below)
    const int pz = g.pecz;
    const bool mx0 = face_mask(pz, 0, i, j, k, g.Nx, g.Ny, g.Nz);
    const bool my0 = face_mask(pz, 1, i, j, k, g.Nx, g.Ny, g.Nz);
    const bool mz0 = face_mask(pz, 2, i, j, k, g.Nx, g.Ny, g.Nz);
    const bool mxz = zok && face_mask(pz, 0, i, j, k + 1, g.Nx, g.Ny, g.Nz);
    const bool myz = zok && face_mask(pz, 1, i, j, k + 1, g.Nx, g.Ny, g.Nz);
    const bool mxy = yok && face_mask(pz, 0, i, yj, yk, g.Nx, g.Ny, g.Nz);
    const bool mzy = yok && face_mask(pz, 2, i, yj, yk, g.Nx, g.Ny, g.Nz);
    const bool myx = xok && face_mask(pz, 1, xi, xj, xk, g.Nx, g.Ny, g.Nz);
    const bool mzx = xok && face_mask(pz, 2, xi, xj, xk, g.Nx, g.Ny, g.Nz);
    const S ex0 = E[n], ey0 = E[N + n], ez0 = E[2 * N + n];
    const S exz = mxz ? E[n + g.sz] : S(0), eyz = myz ? E[N + n + g.sz] : S(0);
    const S exy = mxy ? E[n + g.sy] : S(0), ezy = mzy ? E[2 * N + n + g.sy] : S(0);
    const S eyx = myx ? E[N + n + 1] : S(0), ezx = mzx ? E[2 * N + n + 1] : S(0);
    const S a0 = a[n], a1 = a[N + n], a2 = a[2 * N + n];
    Cm acc0 = Cm(0), acc1 = Cm(0), acc2 = Cm(0);
    // rows (h_terms order): x: +Ey, -Ey(z+), -Ez, +Ez(y+); y: -Ex, +Ex(z+), +Ez,
    // -Ez(x+); z: +Ex, -Ex(y+), -Ey, +Ey(x+)
    if (!g.m1desc) {
      add_h_term<S, Cm>(acc0, dt, (S)(a0 * Lyj), ey0, my0, +1);
      add_h_term<S, Cm>(acc0, dt, (S)(a0 * Lyj), eyz, myz, -1);
      add_h_term<S, Cm>(acc0, dt, (S)(a0 * Lzk), ez0, mz0, -1);
      add_h_term<S, Cm>(acc0, dt, (S)(a0 * Lzy), ezy, mzy, +1);
      add_h_term<S, Cm>(acc1, dt, (S)(a1 * Lxi), ex0, mx0, -1);
      add_h_term<S, Cm>(acc1, dt, (S)(a1 * Lxi), exz, mxz, +1);
      add_h_term<S, Cm>(acc1, dt, (S)(a1 * Lzk), ez0, mz0, +1);
      add_h_term<S, Cm>(acc1, dt, (S)(a1 * Lzx), ezx, mzx, -1);
      add_h_term<S, Cm>(acc2, dt, (S)(a2 * Lxi), ex0, mx0, +1);
      add_h_term<S, Cm>(acc2, dt, (S)(a2 * Lxi), exy, mxy, -1);
      add_h_term<S, Cm>(acc2, dt, (S)(a2 * Lyj), ey0, my0, -1);
      add_h_term<S, Cm>(acc2, dt, (S)(a2 * Lyx), eyx, myx, +1);
    } else {
      add_h_term<S, Cm>(acc0, dt, (S)(a0 * Lzy), ezy, mzy, +1);
      add_h_term<S, Cm>(acc0, dt, (S)(a0 * Lzk), ez0, mz0, -1);
      add_h_term<S, Cm>(acc0, dt, (S)(a0 * Lyj), eyz, myz, -1);
      add_h_term<S, Cm>(acc0, dt, (S)(a0 * Lyj), ey0, my0, +1);
      add_h_term<S, Cm>(acc1, dt, (S)(a1 * Lzx), ezx, mzx, -1);
      add_h_term<S, Cm>(acc1, dt, (S)(a1 * Lzk), ez0, mz0, +1);
      add_h_term<S, Cm>(acc1, dt, (S)(a1 * Lxi), exz, mxz, +1);
      add_h_term<S, Cm>(acc1, dt, (S)(a1 * Lxi), ex0, mx0, -1);
      add_h_term<S, Cm>(acc2, dt, (S)(a2 * Lyx), eyx, myx, +1);
      add_h_term<S, Cm>(acc2, dt, (S)(a2 * Lyj), ey0, my0, -1);
      add_h_term<S, Cm>(acc2, dt, (S)(a2 * Lxi), exy, mxy, -1);
      add_h_term<S, Cm>(acc2, dt, (S)(a2 * Lxi), ex0, mx0, +1);
    }
    H[n] = (S)((Cm)H[n] - acc0);
    H[N + n] = (S)((Cm)H[N + n] - acc1);
    H[2 * N + n] = (S)((Cm)H[2 * N + n] - acc2);
  }
}

struct EStepAux {
  int jmode, ixs, iys;
  int cap_ix0, cap_iy0, cap_nx, cap_ny;
  int64_t cap_samples, cap_sample;
};

// E sweep of the exact path: M2 rows with the flat-index backward neighbours
// (n-1, n-sy, n-sz, wrapping into the previous row/plane), the PMC column
// mask, ascending order; then the J term, the conduction update of J, the
// beam line and the Ez window capture.
template <typename S, typename Cm, int JM, bool COND>
__global__ void __launch_bounds__(EX_TX * EX_TY)
k_e_exact(GridDev g, MetricDev<S> m, S *__restrict__ E, const S *__restrict__ H,
          S *__restrict__ J, const S *__restrict__ b, const S *__restrict__ ieps,
          const S *__restrict__ sigma, Cm dt, EStepAux x, const S *__restrict__ beam_row,
          S *__restrict__ cap) {
  const int i = blockIdx.x * EX_TX + threadIdx.x, j = blockIdx.y * EX_TY + threadIdx.y;
  if (i >= g.Nx || j >= g.Ny) return;
  const int64_t N = g.N;
  const bool xin = i > 0, yin = j > 0;
  const int xi = xin ? i - 1 : g.Nx - 1, xj = xin ? j : (yin ? j - 1 : g.Ny - 1);
  const int yj = yin ? j - 1 : g.Ny - 1;
  const S tLxi = m.tLx[i], tLyj = m.tLy[j], tLxx = m.tLx[xi], tLyx = m.tLy[xj];
  const bool online = (i == x.ixs) && (j == x.iys);
  const int li = i - x.cap_ix0, lj = j - x.cap_iy0;
  const bool incap = cap && li >= 0 && li < x.cap_nx && lj >= 0 && lj < x.cap_ny &&
                     x.cap_sample < x.cap_samples;
  for (int k = blockIdx.z; k < g.Nz; k += gridDim.z) {
    const int64_t n = (int64_t)i + g.sy * j + g.sz * k;
    const int xk = (xin || yin) ? k : k - 1, yk = yin ? k : k - 1;
    const bool xok = xk >= 0, yok = yk >= 0, zok = k > 0;
    const S tLzk = m.tLz[k];
    const S tLzy = yok ? m.tLz[yk] : S(0);
    const int hz = g.pmcz;
    const bool mx0 = face_mask(hz, 0, i, j, k, g.Nx, g.Ny, g.Nz);
    const bool my0 = face_mask(hz, 1, i, j, k, g.Nx, g.Ny, g.Nz);
    const bool mz0 = face_mask(hz, 2, i, j, k, g.Nx, g.Ny, g.Nz);
    const bool mxz = zok && face_mask(hz, 0, i, j, k - 1, g.Nx, g.Ny, g.Nz);
    const bool myz = zok && face_mask(hz, 1, i, j, k - 1, g.Nx, g.Ny, g.Nz);
    const bool mxy = yok && face_mask(hz, 0, i, yj, yk, g.Nx, g.Ny, g.Nz);
    const bool mzy = yok && face_mask(hz, 2, i, yj, yk, g.Nx, g.Ny, g.Nz);
    const bool myx = xok && face_mask(hz, 1, xi, xj, xk, g.Nx, g.Ny, g.Nz);
    const bool mzx = xok && face_mask(hz, 2, xi, xj, xk, g.Nx, g.Ny, g.Nz);
    const S hx0 = H[n], hy0 = H[N + n], hz0 = H[2 * N + n];
    const S hxz = mxz ? H[n - g.sz] : S(0), hyz = myz ? H[N + n - g.sz] : S(0);
    const S hxy = mxy ? H[n - g.sy] : S(0), hzy = mzy ? H[2 * N + n - g.sy] : S(0);
    const S hyx = myx ? H[N + n - 1] : S(0), hzx = mzx ? H[2 * N + n - 1] : S(0);
    S acc[3];
    {
      const S b0 = b[n], b1 = b[N + n], b2 = b[2 * N + n];
      acc[0] = acc[1] = acc[2] = S(0);
      // rows (e_terms order): x: +Hy(z-), -Hy, -Hz(y-), +Hz; y: -Hx(z-), +Hx,
      // +Hz(x-), -Hz; z: +Hx(y-), -Hx, -Hy(x-), +Hy
      add_e_term<S>(acc[0], (S)(b0 * tLyj), hyz, myz, +1);
      add_e_term<S>(acc[0], (S)(b0 * tLyj), hy0, my0, -1);
      add_e_term<S>(acc[0], (S)(b0 * tLzy), hzy, mzy, -1);
      add_e_term<S>(acc[0], (S)(b0 * tLzk), hz0, mz0, +1);
      add_e_term<S>(acc[1], (S)(b1 * tLxi), hxz, mxz, -1);
      add_e_term<S>(acc[1], (S)(b1 * tLxi), hx0, mx0, +1);
      add_e_term<S>(acc[1], (S)(b1 * (xok ? m.tLz[xk] : S(0))), hzx, mzx, +1);
      add_e_term<S>(acc[1], (S)(b1 * tLzk), hz0, mz0, -1);
      add_e_term<S>(acc[2], (S)(b2 * tLxi), hxy, mxy, +1);
      add_e_term<S>(acc[2], (S)(b2 * tLxi), hx0, mx0, -1);
      add_e_term<S>(acc[2], (S)(b2 * tLyx), hyx, myx, -1);
      add_e_term<S>(acc[2], (S)(b2 * tLyj), hy0, my0, +1);
    }
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) {
      const int64_t r = comp * N + n;
      const bool line = (comp == 2) && online;
      S d = acc[comp];
      S jv = S(0);
      if (JM == 2) {
        jv = (line && beam_row) ? beam_row[k] : J[r];
        d = d - ieps[r] * jv;
      } else if (JM == 1 && line) {
        jv = beam_row ? beam_row[k] : J[r];
        d = d - ieps[r] * jv;
      }
      const S en = (S)((Cm)E[r] + dt * (Cm)d);
      E[r] = en;
      if (COND) {
        J[r] = sigma[r] * en;
      } else if (line && beam_row) {
        J[r] = jv;
      }
      if (comp == 2 && incap) {
        const int64_t l = (int64_t)li * x.cap_ny + lj;
        cap[(l * x.cap_samples + x.cap_sample) * g.Nz + k] = en;
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
