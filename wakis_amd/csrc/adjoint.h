// Discrete adjoint of the leapfrog step (reverse mode, FP64 adjoint fields).
//
// Forward step n (SolverFIT3D._one_step, solverFIT3D.py:482-503):
//   H' = H - dt M1 E,   E' = E + dt (M2 H' - ieps J),   J' = sigma E'  (conductive)
// with M1 = Dpmc iDa iDmu C Dpec Ds (row coefficient a = fl(iA imu) with the PMC
// row mask) and M2 = iDeps itDa Dpec C^T Dpmc tDs (row coefficient b = fl(ieps
// itA) with the PEC row mask). On the beam line J is a source, not a state.
// Reverse step n, from the adjoints (lE, lH, lJ) of step n+1, carried as
// W = lE + sigma lJ (what the step's E update sees, since J' = sigma E') and
// lH, with lJ itself kept only on the gradient region:
//   lH <- lH + dt M2^T W,   M2^T W = pmc . tL . C (b . W)
//   lE  = W - dt M1^T lH,   M1^T y = pec . L . C^T (a . y)
//   lJ  = -dt (1 - beam line) ieps . W          (conductive)
//   W  <- lE + sigma lJ
// and the parameter gradients of that step, with E', H' = E^{n+1}, H^{n+1}
// and J = J^n (the beam current on the beam line, fl(sigma E^n) elsewhere):
//   d/d ieps_q : dt z_q (K2 H'_q - J_q),  K2 = itDa Dpec C^T Dpmc tDs (M2 without ieps)
//   d/d sigma_q: lJ'_q E'_q
// The C and C^T rows (with the flat-index wrap and truncation of the
// reference operator) are h_terms / e_terms of the forward path.
#pragma once
#include "fit_common.h"
#include "step_general.h"
#include "step_fused4.h"

namespace wf {

struct AdjAux {
  int conductive, ixs, iys;
  int r0[3], r1[3];  // gradient region (cells), half-open
  int cap_ix0, cap_iy0, cap_nx, cap_ny;
  int64_t cap_samples, cap_sample;
  int traj;  // E0/E1/H1 in the trajectory box layout
};

// Region trajectory box: E on [lo, lo + e), H on [lo - 1, lo + e) (h = e + 1
// per axis), three components each, x fastest, then y, z, component.
struct TrajBox {
  int lo[3], e[3];
  int64_t nE, nH;  // cells per component
};

__host__ __device__ inline TrajBox traj_box(const int *lo, const int *hi) {
  TrajBox b;
  for (int d = 0; d < 3; ++d) { b.lo[d] = lo[d]; b.e[d] = hi[d] - lo[d]; }
  b.nE = (int64_t)b.e[0] * b.e[1] * b.e[2];
  b.nH = (int64_t)(b.e[0] + 1) * (b.e[1] + 1) * (b.e[2] + 1);
  return b;
}

// index of H component comp of cell (i, j, k) in the H box
__device__ __forceinline__ int64_t traj_h_index(const TrajBox &b, int comp, int i, int j, int k) {
  const int hx = b.e[0] + 1, hy = b.e[1] + 1;
  return comp * b.nH + (i - b.lo[0] + 1) + (int64_t)hx * ((j - b.lo[1] + 1) +
                                                           (int64_t)hy * (k - b.lo[2] + 1));
}

// the element t of a trajectory slot (E box, then H box) from the fields E, H
template <typename S>
__device__ __forceinline__ void traj_copy(const GridDev &g, const S *__restrict__ E,
                                          const S *__restrict__ H, const TrajBox &b, int64_t t,
                                          S *__restrict__ tE, S *__restrict__ tH) {
  if (t < 3 * b.nE) {
    const int comp = (int)(t / b.nE);
    const int64_t r = t - comp * b.nE;
    const int i = b.lo[0] + (int)(r % b.e[0]);
    const int j = b.lo[1] + (int)((r / b.e[0]) % b.e[1]);
    const int k = b.lo[2] + (int)(r / ((int64_t)b.e[0] * b.e[1]));
    tE[t] = E[comp * g.N + i + (int64_t)g.Nx * (j + (int64_t)g.Ny * k)];
  } else if (t < 3 * (b.nE + b.nH)) {
    const int64_t u = t - 3 * b.nE;
    const int comp = (int)(u / b.nH);
    const int64_t r = u - comp * b.nH;
    const int hx = b.e[0] + 1, hy = b.e[1] + 1;
    const int i = b.lo[0] - 1 + (int)(r % hx);
    const int j = b.lo[1] - 1 + (int)((r / hx) % hy);
    const int k = b.lo[2] - 1 + (int)(r / ((int64_t)hx * hy));
    tH[u] = H[comp * g.N + i + (int64_t)g.Nx * (j + (int64_t)g.Ny * k)];
  }
}

// one trajectory slot: tE[3 nE] from E, tH[3 nH] from H
template <typename S>
__global__ void __launch_bounds__(256)
k_traj_gather(GridDev g, const S *__restrict__ E, const S *__restrict__ H, TrajBox b,
              S *__restrict__ tE, S *__restrict__ tH) {
  traj_copy<S>(g, E, H, b, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, tE, tH);
}

// the trajectory slots of every grid of a batch after the step the batch
// kernel just ran (step number from device memory, as k_fused4_batch)
template <typename S>
__global__ void __launch_bounds__(256)
k_traj_gather_batch(const BatchDesc<S> *__restrict__ d, int ninst, int64_t total,
                    const int *__restrict__ step_ptr) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  int lo = 0, hi = ninst - 1;  // last descriptor with thr0 <= t
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].thr0 <= t) lo = mid;
    else hi = mid - 1;
  }
  const BatchDesc<S> &e = d[lo];
  const int step = *step_ptr;
  if (!e.tE || step >= e.nsteps) return;
  const TrajBox b = traj_box(e.tlo, e.thi);
  const int64_t slot = e.traj_slot0 + step + 1;
  const bool odd = step & 1;  // the step wrote E2/H2 when even, E/H when odd
  traj_copy<S>(e.g, odd ? e.E : e.E2, odd ? e.H : e.H2, b, t - e.thr0, e.tE + slot * 3 * b.nE,
               e.tH + slot * 3 * b.nH);
}

// parameter gradients of one step over the region (reads W and the region's lJ
// of step n+1); E0, E1, H1 are full fields, or trajectory slots (x.traj)
template <typename S, typename A = double>
__global__ void __launch_bounds__(256)
k_adj_grad(GridDev g, MetricDev<S> m, const A *__restrict__ W,
           const double *__restrict__ lJr, const S *__restrict__ E0, const S *__restrict__ E1,
           const S *__restrict__ H1, const S *__restrict__ beam_row, const S *__restrict__ sigma,
           const S *__restrict__ metric2d, double dt, AdjAux x, double *__restrict__ g_ieps,
           double *__restrict__ g_sigma) {
  const int rx = x.r1[0] - x.r0[0], ry = x.r1[1] - x.r0[1], rz = x.r1[2] - x.r0[2];
  const int64_t nreg = (int64_t)rx * ry * rz;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nreg) return;
  Cell c0;
  c0.i = x.r0[0] + (int)(t % rx);
  c0.j = x.r0[1] + (int)((t / rx) % ry);
  c0.k = x.r0[2] + (int)(t / ((int64_t)rx * ry));
  const int64_t n = flat(c0, g);
  const TrajBox tb = traj_box(x.r0, x.r1);
  const int64_t yz = (int64_t)g.Ny * g.Nz, xz = (int64_t)g.Nx * g.Nz, xy = (int64_t)g.Nx * g.Ny;
  const S *itA[3] = {metric2d + yz + xz + xy, metric2d + 2 * yz + xz + xy,
                     metric2d + 2 * yz + 2 * xz + xy};
  const int64_t t2[3] = {(int64_t)c0.j + (int64_t)g.Ny * c0.k,
                         (int64_t)c0.i + (int64_t)g.Nx * c0.k,
                         (int64_t)c0.i + (int64_t)g.Nx * c0.j};
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    const int64_t q = comp * g.N + n;
    double k2 = 0.0;
    if (face_mask(g.pecz, comp, c0.i, c0.j, c0.k, g.Nx, g.Ny, g.Nz)) {
      Term tt[4];
      e_terms(comp, c0, g, tt);
      double acc = 0.0;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (!tt[u].ok) continue;
        if (!face_mask(g.pmcz, tt[u].comp, tt[u].c.i, tt[u].c.j, tt[u].c.k, g.Nx, g.Ny, g.Nz))
          continue;
        const double w = (double)len_of(m.tLx, m.tLy, m.tLz, tt[u].comp, tt[u].c);
        const int64_t hq = x.traj ? traj_h_index(tb, tt[u].comp, tt[u].c.i, tt[u].c.j,
                                                 tt[u].c.k)
                                  : tt[u].comp * g.N + flat(tt[u].c, g);
        const double p = w * (double)H1[hq];
        acc = tt[u].sign > 0 ? acc + p : acc - p;
      }
      k2 = (double)itA[comp][t2[comp]] * acc;
    }
    // the J step n used: the beam current on the beam line, sigma*E^n elsewhere
    const bool online = (comp == 2) && (c0.i == x.ixs) && (c0.j == x.iys);
    const int64_t eq = x.traj ? comp * nreg + t : q;
    double j0 = 0.0;
    if (online && beam_row) j0 = (double)beam_row[c0.k];
    else if (x.conductive) j0 = (double)(S)(sigma[q] * E0[eq]);
    g_ieps[comp * nreg + t] += dt * (double)W[q] * (k2 - j0);
    if (x.conductive) g_sigma[comp * nreg + t] += lJr[comp * nreg + t] * (double)E1[eq];
  }
}

// Parameter gradients of nsteps reverse steps in one pass, from stored slots:
// slot s holds the forward E of steps n0+s (E at s) and n0+s+1 (E at s+1), H of
// step n0+s+1 (H at s+1), the region's W read by reverse step n0+s (rW at s) and
// lJ of step n0+s+1 (rJ at s+1); beam row n0+s at beam + s*Nz. E/H slots are
// trajectory boxes (x.traj) or full 3N fields (E_stride/H_stride elements per
// slot). Each (cell, component) accumulates over s = nsteps-1 .. 0, the order
// of the per-step kernel, so the sums are the same.
template <typename S, typename A>
__global__ void __launch_bounds__(256)
k_adj_grad_seq(GridDev g, MetricDev<S> m, const S *__restrict__ E, int64_t Es,
               const S *__restrict__ H, int64_t Hs, const A *__restrict__ rW,
               const double *__restrict__ rJ, const S *__restrict__ beam,
               const S *__restrict__ sigma, const S *__restrict__ metric2d, double dt, AdjAux x,
               int nsteps, double *__restrict__ g_ieps, double *__restrict__ g_sigma) {
  const int rx = x.r1[0] - x.r0[0], ry = x.r1[1] - x.r0[1], rz = x.r1[2] - x.r0[2];
  const int64_t nreg = (int64_t)rx * ry * rz;
  const int64_t tt3 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tt3 >= 3 * nreg) return;
  const int comp = (int)(tt3 / nreg);
  const int64_t t = tt3 - comp * nreg;
  Cell c0;
  c0.i = x.r0[0] + (int)(t % rx);
  c0.j = x.r0[1] + (int)((t / rx) % ry);
  c0.k = x.r0[2] + (int)(t / ((int64_t)rx * ry));
  const int64_t n = flat(c0, g);
  const int64_t q = comp * g.N + n;
  const TrajBox tb = traj_box(x.r0, x.r1);
  const int64_t yz = (int64_t)g.Ny * g.Nz, xz = (int64_t)g.Nx * g.Nz, xy = (int64_t)g.Nx * g.Ny;
  // the step-independent part of K2 H' = itA * sum(+-tL H) of this row
  double wgt[4] = {0, 0, 0, 0};
  int64_t hq[4] = {0, 0, 0, 0};
  double itA = 0.0;
  if (face_mask(g.pecz, comp, c0.i, c0.j, c0.k, g.Nx, g.Ny, g.Nz)) {
    Term tt[4];
    e_terms(comp, c0, g, tt);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (!tt[u].ok) continue;
      if (!face_mask(g.pmcz, tt[u].comp, tt[u].c.i, tt[u].c.j, tt[u].c.k, g.Nx, g.Ny, g.Nz))
        continue;
      const double w = (double)len_of(m.tLx, m.tLy, m.tLz, tt[u].comp, tt[u].c);
      wgt[u] = tt[u].sign > 0 ? w : -w;
      hq[u] = x.traj ? traj_h_index(tb, tt[u].comp, tt[u].c.i, tt[u].c.j, tt[u].c.k)
                     : tt[u].comp * g.N + flat(tt[u].c, g);
    }
    const S *itAc = metric2d + (comp == 0 ? yz + xz + xy : (comp == 1 ? 2 * yz + xz + xy
                                                                      : 2 * yz + 2 * xz + xy));
    const int64_t t2 = comp == 0 ? (int64_t)c0.j + (int64_t)g.Ny * c0.k
                                 : (comp == 1 ? (int64_t)c0.i + (int64_t)g.Nx * c0.k
                                              : (int64_t)c0.i + (int64_t)g.Nx * c0.j);
    itA = (double)itAc[t2];
  }
  const bool online = (comp == 2) && (c0.i == x.ixs) && (c0.j == x.iys);
  const int64_t eq = x.traj ? comp * nreg + t : q;
  const S sg = x.conductive ? sigma[q] : S(0);
  double gi = g_ieps[comp * nreg + t];
  double gs = x.conductive ? g_sigma[comp * nreg + t] : 0.0;
  for (int s = nsteps - 1; s >= 0; --s) {
    const S *H1 = H + (int64_t)(s + 1) * Hs;
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (wgt[u] == 0.0) continue;  // absent, masked or zero-length term
      const double p = fabs(wgt[u]) * (double)H1[hq[u]];
      acc = wgt[u] > 0 ? acc + p : acc - p;
    }
    const double k2 = itA * acc;
    double j0 = 0.0;
    if (online && beam) j0 = (double)beam[(int64_t)s * g.Nz + c0.k];
    else if (x.conductive) j0 = (double)(S)(sg * E[(int64_t)s * Es + eq]);
    gi += dt * (double)rW[(int64_t)s * 3 * nreg + comp * nreg + t] * (k2 - j0);
    if (x.conductive)
      gs += rJ[(int64_t)(s + 1) * 3 * nreg + comp * nreg + t] *
            (double)E[(int64_t)(s + 1) * Es + eq];
  }
  g_ieps[comp * nreg + t] = gi;
  if (x.conductive) g_sigma[comp * nreg + t] = gs;
}

// lH <- lH + dt M2^T W
template <typename S>
__global__ void __launch_bounds__(256)
k_adj_h(GridDev g, MetricDev<S> m, double *__restrict__ lH, const double *__restrict__ W,
        const S *__restrict__ b, double dt) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.N) return;
  const Cell c0 = cell_of(n, g);
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    const int64_t r = comp * g.N + n;
    if (!face_mask(g.pmcz, comp, c0.i, c0.j, c0.k, g.Nx, g.Ny, g.Nz)) continue;
    Term tt[4];
    h_terms(comp, c0, g, tt);
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (!tt[u].ok) continue;
      const int64_t q = tt[u].comp * g.N + flat(tt[u].c, g);
      const double p = (double)b[q] * W[q];
      acc = tt[u].sign > 0 ? acc + p : acc - p;
    }
    const double tl = (double)len_of(m.tLx, m.tLy, m.tLz, comp, c0);
    lH[r] = lH[r] + dt * (tl * acc);
  }
}

// lE = W - dt M1^T lH (+ capture gradient of sample n), lJ = -dt (1-B) ieps W,
// W <- lE + sigma lJ (lJ also stored on the gradient region)
template <typename S>
__global__ void __launch_bounds__(256)
k_adj_e(GridDev g, MetricDev<S> m, double *__restrict__ W, double *__restrict__ lJr,
        const double *__restrict__ lH, const S *__restrict__ a, const S *__restrict__ ieps,
        const S *__restrict__ sigma, double dt, AdjAux x, const double *__restrict__ gcap) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.N) return;
  const Cell c0 = cell_of(n, g);
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    const int64_t q = comp * g.N + n;
    const double z = W[q];
    double m1t = 0.0;
    if (face_mask(g.pecz, comp, c0.i, c0.j, c0.k, g.Nx, g.Ny, g.Nz)) {
      Term tt[4];
      e_terms(comp, c0, g, tt);
      double acc = 0.0;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (!tt[u].ok) continue;
        const int64_t r = tt[u].comp * g.N + flat(tt[u].c, g);
        const double p = (double)a[r] * lH[r];
        acc = tt[u].sign > 0 ? acc + p : acc - p;
      }
      m1t = (double)len_of(m.Lx, m.Ly, m.Lz, comp, c0) * acc;
    }
    double le = z - dt * m1t;
    if (gcap && comp == 2 && x.cap_sample < x.cap_samples) {
      const int li = c0.i - x.cap_ix0, lj = c0.j - x.cap_iy0;
      if (li >= 0 && li < x.cap_nx && lj >= 0 && lj < x.cap_ny) {
        const int64_t l = (int64_t)li * x.cap_ny + lj;
        le += gcap[(l * x.cap_samples + x.cap_sample) * g.Nz + c0.k];
      }
    }
    if (x.conductive) {
      const bool online = (comp == 2) && (c0.i == x.ixs) && (c0.j == x.iys);
      const double lj = online ? 0.0 : -dt * (double)ieps[q] * z;
      W[q] = le + (double)sigma[q] * lj;
      if (lJr && c0.i >= x.r0[0] && c0.i < x.r1[0] && c0.j >= x.r0[1] && c0.j < x.r1[1] &&
          c0.k >= x.r0[2] && c0.k < x.r1[2]) {
        const int rx = x.r1[0] - x.r0[0], ry = x.r1[1] - x.r0[1], rz = x.r1[2] - x.r0[2];
        const int64_t t = (int64_t)(c0.i - x.r0[0]) +
                          (int64_t)rx * ((c0.j - x.r0[1]) + (int64_t)ry * (c0.k - x.r0[2]));
        lJr[(int64_t)comp * rx * ry * rz + t] = lj;
      }
    } else {
      W[q] = le;
    }
  }
}

// Capture gradient of the line integrals: gcap[l][m][k0+kr] = sum over the s
// indices n with it(n,k) == m of gI[l][n] * w_k, with the row kr, the time
// index it and the weight w_k of k_wake_wp<CM> (forward: kr = k, w_k = dz[k];
// counter-moving: kr = nz-1-k, w_k = -dz[k]). In both cases it is
// non-decreasing in n (ts grows with s). One thread per (line, k): runs of
// equal m are summed in order, so the result is deterministic.
template <bool CM>
__global__ void k_wake_wp_adjoint(const double *__restrict__ gI, int nlines, int64_t cap_samples,
                                  int Nz, int k0, int nz, const double *__restrict__ z,
                                  const double *__restrict__ dz, const double *__restrict__ s,
                                  int64_t ns, double v, double zref_v, double t0, double ti,
                                  double dt, int64_t nt, double *__restrict__ gcap) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nlines * nz) return;
  const int l = (int)(t / nz), k = (int)(t - (int64_t)l * nz);
  const int kr = CM ? nz - 1 - k : k;
  const double zk = z[kr], wk = CM ? -dz[k] : dz[k];
  const double nv = -v;
  int64_t cur = -1;
  double acc = 0.0;
  for (int64_t n = 0; n < ns; ++n) {
    double ts = CM ? (zk - s[n]) / nv : (zk + s[n]) / v;
    ts = ts - zref_v;
    ts = ts - t0;
    ts = ts + ti;
    int64_t it = (int64_t)(ts / dt);
    if (it >= nt) break;
    if (it < 0) it += nt;  // python negative index (not reached for t0 = 0)
    if (it != cur) {
      if (cur >= 0) gcap[((int64_t)l * cap_samples + cur) * Nz + k0 + kr] += acc;
      cur = it;
      acc = 0.0;
    }
    acc += gI[(int64_t)l * ns + n] * wk;
  }
  if (cur >= 0) gcap[((int64_t)l * cap_samples + cur) * Nz + k0 + kr] += acc;
}

// out[n] = Re sum_f (w_re[f] + i w_im[f]) exp(-2 pi i f n / N), n < len: the
// transpose of the kept-bin DFT, phase reduced exactly as (f*n mod N).
static __global__ void k_idft_real(const double *__restrict__ w, int64_t nbins, int64_t N, int64_t len,
                            double *__restrict__ out) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= len) return;
  double s = 0.0, c = 0.0;
  for (int64_t f = 0; f < nbins; ++f) {
    const int64_t ph = (f * n) % N;  // exact phase reduction (f*n < 2^63 for kept bins)
    double sn, cs;
    sincospi(2.0 * (double)ph / (double)N, &sn, &cs);
    const double x = w[2 * f] * cs + w[2 * f + 1] * sn;  // Re((wr + i wi)(cs - i sn))
    const double tt = s + x;
    if (fabs(s) >= fabs(x)) c += (s - tt) + x;
    else c += (x - tt) + s;
    s = tt;
  }
  out[n] = s + c;
}

}  // namespace wf
