// Single-pass reverse leapfrog step (the two kernels k_adj_h and k_adj_e of
// adjoint.h in one z-march over the grid).
//
// Reverse step n, P = b.W and Q = a.lH', arithmetic and LDS staging in C
// (FP64; FP32 when the fields are stored FP32 and the caller asks for FP32
// arithmetic); W and lH are stored as A (FP64, or FP32 to halve their traffic
// when the forward run is FP32); lJ and the gradient slots stay FP64:
//   lH' = lH + dt tL . (C P)            (PMC-masked rows keep lH)
//   W'  = W - dt pec . L . (C^T Q) [+ capture gradient] [+ sigma lJ]
//   lJ  = -dt (1 - beam line) ieps . W  (conductive; stored on the region)
// C P uses forward differences of P (x+1, y+1, z+1), C^T Q backward
// differences of Q (x-1, y-1, z-1); the terms are summed in the order of
// h_terms / e_terms, so with A = double the results equal the two-kernel path
// bit for bit (C = double). The flat-index wrap couplings are dropped: valid when their
// operator entries are zero or act on frozen zero fields (the condition of the
// fused forward step), with the gradient region away from the low x and y
// faces. Offsets are 32-bit byte offsets (3N * sizeof(double) < 4 GiB).
//
// Layout: a workgroup covers rows j0-1 .. j0+TY with TY + 2 row waves (wave
// w <-> row j0-1+w, lane l <-> cell i0+l) plus one halo wave whose lanes hold
// the columns i0-1 (lanes 0..31) and i0+64 (lanes 32..63) of those rows. It
// owns rows 1..TY of its 64 columns and marches up a chunk of KZ planes. Each
// thread loads and stages its own cell, so the loads of a wave are uniform
// (out-of-grid cells read a valid address and are masked to zero): W and b two
// planes ahead, lH, a, ieps, sigma one plane ahead, in raw form. Per plane:
// stage P -> barrier -> lH', Q -> stage Q -> barrier -> W' of the owned cells.
// With kappa = -dt ieps sigma (host-computed, zero on the beam line's z edges)
// the conduction term is kappa W, one array instead of two. W' and lH' go to
// separate buffers (the caller alternates them), so every
// tile reads the step's inputs only.
#pragma once
#include "adjoint.h"
#include "fit_common.h"
#include "step_fused4.h"
#include "step_general.h"

namespace wf {

struct AdjTile {
  int ntx, nty, ntz, KZ;
};

constexpr int ADJ_TX = 64;  // owned cells per tile row

// is the component-c row of cell (i, j, k) kept by the zero-face bits (the
// face_mask of the general path)?
__device__ __forceinline__ bool keep_row(int bits, int c, int i, int j, int k, const GridDev &g) {
  return bits == 0 || face_mask(bits, c, i, j, k, g.Nx, g.Ny, g.Nz);
}

// Everything one reverse step of one grid reads (kernel argument by value).
template <typename S, typename A>
struct AdjStep {
  GridDev g;
  MetricDev<S> m;
  const A *Win, *lHin;
  A *Wout, *lHout;
  double *lJr;
  const S *a, *b, *ieps, *sigma, *kappa;
  double dt;
  AdjAux x;
  const double *gcap;
  AdjTile t;
  A *rW;
  double *rJ;
  int64_t rslot;
  // LUT kernels: per-cell material code into mat[nmat][9] (imu, ieps, sigma per
  // component, in LDS) and the masked 2-D metric tables of the forward step;
  // with esc, code 255 marks cells whose coefficients come from a, b, ieps and
  // kappa (or sigma)
  const uint8_t *code;
  const S *mat;
  int nmat, esc;
  const S *iAx, *iAy, *iAz, *itAx, *itAy, *itAz;
};

// One workgroup's z-march over tile `tile` of one grid's reverse step. LUT:
// the coefficients a, b, kappa (or sigma) and ieps are recomputed from the
// material table exactly as the forward step folds them (a = fl(iA imu),
// b = fl(ieps itA), kappa = fl((-dt ieps) sigma) in FP64, zero on the beam
// line's z edges), reading one byte per cell instead of nine values.
template <typename S, typename A, int TY, bool LUT, typename C>
__device__ __forceinline__ void adj_fused_tile(const AdjStep<S, A> &p, const int tile) {
  const GridDev &g = p.g;
  const MetricDev<S> &m = p.m;
  const A *__restrict__ Win = p.Win;
  const A *__restrict__ lHin = p.lHin;
  A *__restrict__ Wout = p.Wout;
  A *__restrict__ lHout = p.lHout;
  double *__restrict__ lJr = p.lJr;
  const S *__restrict__ a = p.a;
  const S *__restrict__ b = p.b;
  const S *__restrict__ ieps = p.ieps;
  const S *__restrict__ sigma = p.sigma;
  const S *__restrict__ kappa = p.kappa;
  const double dtd = p.dt;  // FP64 dt: the host's fold of kappa
  const C dt = (C)dtd;
  const AdjAux &x = p.x;
  const double *__restrict__ gcap = p.gcap;
  const AdjTile &t = p.t;
  A *__restrict__ rW = p.rW;
  double *__restrict__ rJ = p.rJ;
  const int64_t rslot = p.rslot;
  static_assert(TY + 3 <= 16, "at most 1024 threads per workgroup");
  using I = uint32_t;
  constexpr int NW = TY + 2;              // row waves
  constexpr int NS = ADJ_TX + 3;          // slots: 0 = i0-1, 1..64 cells, 65 = i0+64, 66 spare
  __shared__ C sP[3][NW + 1][NS];    // row NW: zeros (above the top row)
  __shared__ C sQ[3][NW][NS];

  const int tz = tile / (t.ntx * t.nty);
  const int rem = tile - tz * t.ntx * t.nty;
  const int tyb = rem / t.ntx, txb = rem - tyb * t.ntx;
  const int i0 = txb * ADJ_TX, j0 = tyb * TY, k0 = tz * t.KZ;
  const int kend = min(k0 + t.KZ, g.Nz);

  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const bool halo = (w == NW);
  const int side = lane >> 5;  // halo wave: 0 = column i0-1, 1 = column i0+64
  const int row = halo ? (lane & 31) : w;
  const int slot = halo ? (side ? ADJ_TX + 1 : 0) : 1 + lane;
  const int i = halo ? (side ? i0 + ADJ_TX : i0 - 1) : i0 + lane;
  const int j = j0 - 1 + row;
  const bool inrow = row < NW;
  const bool valid = inrow && (i >= 0) && (i < g.Nx) && (j >= 0) && (j < g.Ny);
  const bool own = valid && !halo && (w >= 1) && (w <= TY);
  const bool makes_q = inrow && !(halo && side);  // the column i0+64 needs P only
  // what a thread's cell needs besides W and b: lH and a where lH' is used (own
  // rows, the row below, the column i0-1), ieps/sigma/kappa on own rows only
  const bool need_lh = (!halo && w <= TY) || (halo && !side && inrow);
  const bool need_c = !halo && w >= 1 && w <= TY;
  const I N = (I)g.N, sz = (I)g.sz;
  const I col = valid ? (I)i + (I)g.sy * (I)j : 0;
  for (int u = threadIdx.x; u < 3 * (NW + 1) * NS; u += 64 * (NW + 1))
    (&sP[0][0][0])[u] = C(0);
  extern __shared__ unsigned char s_dyn[];
  S *sMat = reinterpret_cast<S *>(s_dyn);
  if (LUT) {
    for (int u = threadIdx.x; u < p.nmat * 9; u += 64 * (NW + 1)) sMat[u] = p.mat[u];
  }

  const C tLx = valid ? (C)m.tLx[i] : C(0), tLy = valid ? (C)m.tLy[j] : C(0);
  const C Lx = valid ? (C)m.Lx[i] : C(0), Ly = valid ? (C)m.Ly[j] : C(0);
  const bool online = own && (i == x.ixs) && (j == x.iys);
  const bool capcol = gcap && own && x.cap_sample < x.cap_samples && i >= x.cap_ix0 &&
                      i < x.cap_ix0 + x.cap_nx && j >= x.cap_iy0 && j < x.cap_iy0 + x.cap_ny;
  const int64_t capbase =
      capcol ? ((int64_t)(i - x.cap_ix0) * x.cap_ny + (j - x.cap_iy0)) * x.cap_samples : 0;
  const bool regcol = own && i >= x.r0[0] && i < x.r1[0] && j >= x.r0[1] && j < x.r1[1];
  const int rx = x.r1[0] - x.r0[0], ry = x.r1[1] - x.r0[1], rz = x.r1[2] - x.r0[2];
  const int64_t nreg = (int64_t)rx * ry * rz;
  const bool cond = x.conductive != 0;
  // kappa = -dt ieps sigma per edge replaces the sigma and ieps reads (ieps is
  // then read on the region only, for lJ); sg[] holds kappa in that case
  const bool kap = kappa != nullptr;
  const S *sgsrc = kap ? kappa : sigma;
  const auto in_region = [&](int p) { return regcol && p >= x.r0[2] && p < x.r1[2]; };

  // plane p's offset of this cell, clamped into the grid (masked at use)
  const auto off = [&](int p) { return col + sz * (I)min(p, g.Nz - 1); };
  const int kstart = (k0 > 0) ? k0 - 1 : k0;
  // LUT: material codes of planes k, k+1 (k+2 in the prefetch) and the
  // metric those planes need; escaped cells (code 255) load the arrays instead
  const int ii = valid ? i : 0, jj = valid ? j : 0;
  const auto kc = [&](int p) { return min(p, g.Nz - 1); };
  const auto esc_at = [&](int cd) { return LUT && p.esc && cd == 255; };
  int ck = 0, ck1 = 0;
  S bxk = 0, bxk1 = 0, byk = 0, byk1 = 0, axk = 0, ayk = 0, bz = 0, az = 0;
  if (LUT) {
    ck = p.code[off(kstart)];
    ck1 = p.code[off(kstart + 1)];
    bz = p.itAz[ii + g.Nx * jj];
    az = p.iAz[ii + g.Nx * jj];
    bxk = p.itAx[jj + g.Ny * kc(kstart)]; bxk1 = p.itAx[jj + g.Ny * kc(kstart + 1)];
    byk = p.itAy[ii + g.Nx * kc(kstart)]; byk1 = p.itAy[ii + g.Nx * kc(kstart + 1)];
    axk = p.iAx[jj + g.Ny * kc(kstart)];
    ayk = p.iAy[ii + g.Nx * kc(kstart)];
  }
  const bool arr0 = !LUT || esc_at(ck), arr1 = !LUT || esc_at(ck1);
  A Wk[3], Wk1[3], lh[3];
  S bk[3], bk1[3], ac[3], ie[3], sg[3];
  {
    const I n0 = off(kstart), n1 = off(kstart + 1);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      Wk[c] = *at(Win, (I)c * N + n0);
      bk[c] = arr0 ? *at(b, (I)c * N + n0) : S(0);
      Wk1[c] = *at(Win, (I)c * N + n1);
      bk1[c] = arr1 ? *at(b, (I)c * N + n1) : S(0);
      lh[c] = need_lh ? *at(lHin, (I)c * N + n0) : A(0);
      ac[c] = (need_lh && arr0) ? *at(a, (I)c * N + n0) : S(0);
      ie[c] = (need_c && arr0 && (!kap || in_region(kstart))) ? *at(ieps, (I)c * N + n0) : S(0);
      sg[c] = (need_c && arr0) ? *at(sgsrc, (I)c * N + n0) : S(0);
    }
  }
  // coefficients of a plane from its code (or the loaded arrays when escaped)
  const bool onl = valid && (i == x.ixs) && (j == x.iys);  // beam line: kappa_z = 0
  const auto bval = [&](int c, int cd, S bx, S by, S arr) -> S {
    if (!LUT || esc_at(cd)) return arr;
    return (S)(sMat[9 * cd + 3 + c] * (c == 0 ? bx : (c == 1 ? by : bz)));
  };
  const auto aval = [&](int c, int cd, S ax, S ay, S arr) -> S {
    if (!LUT || esc_at(cd)) return arr;
    return (S)((c == 0 ? ax : (c == 1 ? ay : az)) * sMat[9 * cd + c]);
  };
  const auto ieval = [&](int c, int cd, S arr) -> S {
    return (!LUT || esc_at(cd)) ? arr : sMat[9 * cd + 3 + c];
  };
  const auto sgval = [&](int c, int cd, S arr) -> S {  // kappa, or sigma
    if (!LUT || esc_at(cd)) return arr;
    if (!kap) return sMat[9 * cd + 6 + c];
    if (c == 2 && onl) return S(0);
    return (S)((-dtd * (double)sMat[9 * cd + 3 + c]) * (double)sMat[9 * cd + 6 + c]);
  };
  C Qpx = 0, Qpy = 0;  // Q(k-1) of this cell, x and y components
  __syncthreads();          // zeroed LDS

  for (int k = kstart; k < kend; ++k) {
    const bool body = (k >= k0);
    const bool kp1 = (k + 1 < g.Nz);
    // ---- prefetch: W, b of plane k+2; lH, a, ieps, sigma of plane k+1
    A Wn2[3], lhN[3];
    S bn2[3], acN[3], ieN[3], sgN[3];
    int cn2 = 0;
    S bxn2 = 0, byn2 = 0, axN = 0, ayN = 0;
    {
      const I n2 = off(k + 2), n1 = off(k + 1);
      if (LUT) {
        cn2 = p.code[n2];
        bxn2 = p.itAx[jj + g.Ny * kc(k + 2)];
        byn2 = p.itAy[ii + g.Nx * kc(k + 2)];
        axN = p.iAx[jj + g.Ny * kc(k + 1)];
        ayN = p.iAy[ii + g.Nx * kc(k + 1)];
      }
      const bool arrn2 = !LUT || esc_at(cn2), arrn1 = !LUT || esc_at(ck1);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        Wn2[c] = *at(Win, (I)c * N + n2);
        bn2[c] = arrn2 ? *at(b, (I)c * N + n2) : S(0);
        lhN[c] = need_lh ? *at(lHin, (I)c * N + n1) : A(0);
        acN[c] = (need_lh && arrn1) ? *at(a, (I)c * N + n1) : S(0);
        if (cond && need_c && arrn1) {
          ieN[c] = (!kap || in_region(k + 1)) ? *at(ieps, (I)c * N + n1) : S(0);
          sgN[c] = *at(sgsrc, (I)c * N + n1);
        } else {
          ieN[c] = S(0); sgN[c] = S(0);
        }
      }
    }
    // ---- stage P(k)
    C Pk[3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
      Pk[c] = valid ? (C)bval(c, ck, bxk, byk, bk[c]) * (C)Wk[c] : C(0);
    if (inrow) {
#pragma unroll
      for (int c = 0; c < 3; ++c) sP[c][row][slot] = Pk[c];
    }
    const C tLz = (C)m.tLz[k], Lz = (C)m.Lz[k];
    __syncthreads();

    // ---- lH'(k) and Q(k) (row waves: all 64 columns; halo wave: column i0-1)
    C hn[3] = {0, 0, 0};
    if (makes_q) {
      C Pk1[3];
#pragma unroll
      for (int c = 0; c < 3; ++c)
        Pk1[c] = (valid && kp1) ? (C)bval(c, ck1, bxk1, byk1, bk1[c]) * (C)Wk1[c] : C(0);
      // h_terms order: x: +Py, -Py(z+1), -Pz, +Pz(y+1); y: -Px, +Px(z+1), +Pz, -Pz(x+1);
      // z: +Px, -Px(y+1), -Py, +Py(x+1)
      C s0 = 0;
      s0 = s0 + Pk[1]; s0 = s0 - Pk1[1]; s0 = s0 - Pk[2]; s0 = s0 + sP[2][row + 1][slot];
      C s1 = 0;
      s1 = s1 - Pk[0]; s1 = s1 + Pk1[0]; s1 = s1 + Pk[2]; s1 = s1 - sP[2][row][slot + 1];
      C s2 = 0;
      s2 = s2 + Pk[0]; s2 = s2 - sP[0][row + 1][slot]; s2 = s2 - Pk[1];
      s2 = s2 + sP[1][row][slot + 1];
      const C h0 = (C)lh[0], h1 = (C)lh[1], h2 = (C)lh[2];
      hn[0] = keep_row(g.pmcz, 0, i, j, k, g) ? h0 + dt * (tLx * s0) : h0;
      hn[1] = keep_row(g.pmcz, 1, i, j, k, g) ? h1 + dt * (tLy * s1) : h1;
      hn[2] = keep_row(g.pmcz, 2, i, j, k, g) ? h2 + dt * (tLz * s2) : h2;
#pragma unroll
      for (int c = 0; c < 3; ++c)
        sQ[c][row][slot] = valid ? (C)aval(c, ck, axk, ayk, ac[c]) * hn[c] : C(0);
    }
    __syncthreads();

    // ---- W'(k) of the owned cells on body planes
    C Qx = 0, Qy = 0;
    if (!halo) { Qx = sQ[0][row][slot]; Qy = sQ[1][row][slot]; }
    if (own && body) {
      const C Qz = sQ[2][row][slot];
      // e_terms order: x: +Qy(z-1), -Qy, -Qz(y-1), +Qz; y: -Qx(z-1), +Qx, +Qz(x-1), -Qz;
      // z: +Qx(y-1), -Qx, -Qy(x-1), +Qy
      C e0 = 0;
      e0 = e0 + Qpy; e0 = e0 - Qy; e0 = e0 - sQ[2][row - 1][slot]; e0 = e0 + Qz;
      C e1 = 0;
      e1 = e1 - Qpx; e1 = e1 + Qx; e1 = e1 + sQ[2][row][slot - 1]; e1 = e1 - Qz;
      C e2 = 0;
      e2 = e2 + sQ[0][row - 1][slot]; e2 = e2 - Qx; e2 = e2 - sQ[1][row][slot - 1]; e2 = e2 + Qy;
      const C m0 = keep_row(g.pecz, 0, i, j, k, g) ? Lx * e0 : C(0);
      const C m1 = keep_row(g.pecz, 1, i, j, k, g) ? Ly * e1 : C(0);
      const C m2 = keep_row(g.pecz, 2, i, j, k, g) ? Lz * e2 : C(0);
      const C wv[3] = {(C)Wk[0], (C)Wk[1], (C)Wk[2]};
      C le[3] = {wv[0] - dt * m0, wv[1] - dt * m1, wv[2] - dt * m2};
      if (capcol) le[2] += (C)gcap[(capbase + x.cap_sample) * g.Nz + k];
      C wn[3];
      const bool inreg = regcol && k >= x.r0[2] && k < x.r1[2];
      const int64_t tr = inreg ? (int64_t)(i - x.r0[0]) +
                                     (int64_t)rx * ((j - x.r0[1]) + (int64_t)ry * (k - x.r0[2]))
                               : 0;
      C ljv[3] = {0, 0, 0};
      if (cond) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const C iec = (C)ieval(c, ck, ie[c]), sgc = (C)sgval(c, ck, sg[c]);
          ljv[c] = (c == 2 && online) ? C(0) : -dt * iec * wv[c];
          wn[c] = kap ? le[c] + sgc * wv[c] : le[c] + sgc * ljv[c];
        }
      } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) wn[c] = le[c];
      }
      if (inreg) {
        if (rW) {  // region W read by this step and its lJ, for the one-pass gradient
          const int64_t base = rslot * 3 * nreg + tr;
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            rW[base + c * nreg] = Wk[c];
            rJ[base + c * nreg] = (double)ljv[c];
          }
        } else if (lJr && cond) {
#pragma unroll
          for (int c = 0; c < 3; ++c) lJr[c * nreg + tr] = (double)ljv[c];
        }
      }
      const I nk = col + sz * (I)k;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        *atw(Wout, (I)c * N + nk) = (A)wn[c];
        *atw(lHout, (I)c * N + nk) = (A)hn[c];
      }
    }
    Qpx = Qx;  // Q(k) of this cell seeds the z-1 terms of plane k+1
    Qpy = Qy;
    // advance the register pipeline. Plane k+1 stages sP after this plane's
    // second barrier (every read of this plane's sP is before it) and writes
    // sQ after its own first barrier (every read of this plane's sQ is before it).
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      Wk[c] = Wk1[c]; bk[c] = bk1[c];
      Wk1[c] = Wn2[c]; bk1[c] = bn2[c];
      lh[c] = lhN[c]; ac[c] = acN[c]; ie[c] = ieN[c]; sg[c] = sgN[c];
    }
    if (LUT) {
      ck = ck1; ck1 = cn2;
      bxk = bxk1; bxk1 = bxn2; byk = byk1; byk1 = byn2;
      axk = axN; ayk = ayN;
    }
  }
}

template <typename S, typename A, int TY, bool LUT = false, typename C = double>
__global__ void __launch_bounds__(64 * (TY + 3)) k_adj_fused(AdjStep<S, A> p) {
  adj_fused_tile<S, A, TY, LUT, C>(p, xcd_tile(blockIdx.x, p.t.ntx * p.t.nty * p.t.ntz));
}

// One grid of a reverse batch. Reverse step r of the batch is step Nt-1-r of
// the grid (grids with fewer steps start later); even r reads W/lH and writes
// W2/lH2, odd r the reverse.
template <typename S, typename A>
struct AdjBatchDesc {
  AdjStep<S, A> p;  // W/lH in p.Win/p.lHin, W2/lH2 in p.Wout/p.lHout
  int Nt;
  int tile0;
};

template <typename S, typename A, int TY, typename C = double>
__global__ void __launch_bounds__(64 * (TY + 3))
k_adj_fused_batch(const AdjBatchDesc<S, A> *__restrict__ d, int ninst, int ntiles,
                  const int *__restrict__ rptr) {
  const int r = *rptr;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  int lo = 0, hi = ninst - 1;  // last descriptor with tile0 <= tile
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].tile0 <= tile) lo = mid;
    else hi = mid - 1;
  }
  const int n = d[lo].Nt - 1 - r;
  if (n < 0) return;  // whole workgroup, before any barrier
  AdjStep<S, A> p = d[lo].p;
  if (r & 1) {
    const A *wi = p.Win, *li = p.lHin;
    p.Win = p.Wout; p.lHin = p.lHout;
    p.Wout = const_cast<A *>(wi); p.lHout = const_cast<A *>(li);
  }
  p.x.cap_sample = n;
  p.rslot = n;
  adj_fused_tile<S, A, TY, false, C>(p, tile - d[lo].tile0);
}

}  // namespace wf
