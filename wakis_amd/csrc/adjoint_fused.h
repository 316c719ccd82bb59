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

// a copy of v in a vector register, made anew where it is used: values derived
// from it inside a loop are not hoisted out of it (and not held as lane masks
// across it)
__device__ __forceinline__ uint32_t vreg_copy(uint32_t v) {
  uint32_t r;
  asm volatile("v_mov_b32 %0, %1" : "=v"(r) : "v"(v));
  return r;
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
  // LUT: per material 12 values in three 16-byte groups, [ieps x y z, -],
  // [imu x y z, -], [kappa (or sigma) x y z, -]; kappa is the per-cell fold
  // (-dt ieps) sigma done once per material and component
  extern __shared__ __align__(16) unsigned char s_adj_tab[];
  S *sMat = reinterpret_cast<S *>(s_adj_tab);
  const bool kap = kappa != nullptr;
  if (LUT) {
    for (int u = threadIdx.x; u < p.nmat * 12; u += 64 * (NW + 1)) {
      const int cd = u / 12, r = u - 12 * cd, grp = r >> 2, c = r & 3;
      const S *mt = p.mat + 9 * cd;
      S v = S(0);
      if (c < 3) {
        if (grp == 0) v = mt[3 + c];
        else if (grp == 1) v = mt[c];
        else v = kap ? (S)((-dtd * (double)mt[3 + c]) * (double)mt[6 + c]) : mt[6 + c];
      }
      sMat[u] = v;
    }
  }
  using S4 = S __attribute__((ext_vector_type(4)));
  const auto mgroup = [&](int cd, int grp) -> S4 {
    return *reinterpret_cast<const S4 *>(sMat + 12 * cd + 4 * grp);
  };

  const C tLx = valid ? (C)m.tLx[i] : C(0), tLy = valid ? (C)m.tLy[j] : C(0);
  const C Lx = valid ? (C)m.Lx[i] : C(0), Ly = valid ? (C)m.Ly[j] : C(0);
  // the zero-face rows (the face_mask of the general path: PMC-masked lH rows
  // keep lH, PEC-masked W rows drop the curl term), split into the in-plane
  // part (per thread, one bit each: bits 0..2 the lH rows x, y, z, bits 3..5
  // the W rows) and the z-face part (per plane)
  const uint32_t kbits = (uint32_t)(ypart(g.pmcz, j, g.Ny) != 0) |
                         ((uint32_t)(xpart(g.pmcz, i, g.Nx) != 0) << 1) |
                         ((uint32_t)mz(g.pmcz, i, j, g) << 2) |
                         ((uint32_t)(ypart(g.pecz, j, g.Ny) != 0) << 3) |
                         ((uint32_t)(xpart(g.pecz, i, g.Nx) != 0) << 4) |
                         ((uint32_t)mz(g.pecz, i, j, g) << 5);
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
  const S *sgsrc = kap ? kappa : sigma;
  const auto in_region = [&](int p) { return regcol && p >= x.r0[2] && p < x.r1[2]; };

  // plane p's offset of this cell, clamped into the grid (masked at use)
  const auto off = [&](int p) { return col + sz * (I)min(p, g.Nz - 1); };
  const int kstart = (k0 > 0) ? k0 - 1 : k0;
  // LUT: material codes of planes k, k+1 (k+2 in the prefetch) and the metric
  // those planes need (32-bit offsets into the 2-D tables); escaped cells (code
  // 255) read the arrays where the value is used, in branches that waves
  // without such cells skip. Without codes the arrays are prefetched.
  const I ii = valid ? (I)i : 0, jj = valid ? (I)j : 0, nx = (I)g.Nx, ny = (I)g.Ny;
  const auto kc = [&](int p) { return (I)min(p, g.Nz - 1); };
  const auto esc_at = [&](int cd) { return LUT && p.esc && cd == 255; };
  const auto tab = [&](const S *t, I idx) { return *at(t, idx); };
  int ck = 0, ck1 = 0;
  S bxk = 0, bxk1 = 0, byk = 0, byk1 = 0, axk = 0, ayk = 0, bz = 0, az = 0;
  if (LUT) {
    ck = p.code[off(kstart)];
    ck1 = p.code[off(kstart + 1)];
    bz = tab(p.itAz, ii + nx * jj);
    az = tab(p.iAz, ii + nx * jj);
    bxk = tab(p.itAx, jj + ny * kc(kstart)); bxk1 = tab(p.itAx, jj + ny * kc(kstart + 1));
    byk = tab(p.itAy, ii + nx * kc(kstart)); byk1 = tab(p.itAy, ii + nx * kc(kstart + 1));
    axk = tab(p.iAx, jj + ny * kc(kstart));
    ayk = tab(p.iAy, ii + nx * kc(kstart));
  }
  A Wk[3], Wk1[3], lh[3];
  S bk1[3], ac[3], ie[3], sg[3];
  C Pk[3];  // P of plane k (computed once, as P(k+1) of the plane before)
  {
    const I n0 = off(kstart), n1 = off(kstart + 1);
    S bk[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      Wk[c] = *at(Win, (I)c * N + n0);
      Wk1[c] = *at(Win, (I)c * N + n1);
      lh[c] = need_lh ? *at(lHin, (I)c * N + n0) : A(0);
      if (LUT) {
        bk[c] = bk1[c] = ac[c] = ie[c] = sg[c] = S(0);
      } else {
        bk[c] = *at(b, (I)c * N + n0);
        bk1[c] = *at(b, (I)c * N + n1);
        ac[c] = need_lh ? *at(a, (I)c * N + n0) : S(0);
        ie[c] = (need_c && (!kap || in_region(kstart))) ? *at(ieps, (I)c * N + n0) : S(0);
        sg[c] = need_c ? *at(sgsrc, (I)c * N + n0) : S(0);
      }
    }
    if (esc_at(ck)) {
#pragma unroll
      for (int c = 0; c < 3; ++c) bk[c] = *at(b, (I)c * N + n0);
    }
    __syncthreads();  // zeroed LDS, material table
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const S bv = (!LUT || esc_at(ck))
                       ? bk[c]
                       : (S)(mgroup(ck, 0)[c] * (c == 0 ? bxk : (c == 1 ? byk : bz)));
      Pk[c] = valid ? (C)bv * (C)Wk[c] : C(0);
    }
  }
  // coefficients of plane k from its code (escaped cells and the array path:
  // the arrays)
  const bool onl = valid && (i == x.ixs) && (j == x.iys);  // beam line: kappa_z = 0
  C Qpx = 0, Qpy = 0;  // Q(k-1) of this cell, x and y components

  for (int k = kstart; k < kend; ++k) {
    const bool body = (k >= k0);
    const bool kp1 = (k + 1 < g.Nz);
    const I nk = off(k), nk1 = off(k + 1);
    // ---- this plane's loads first (escaped cells' arrays, the z metric), so
    // that waiting for them never waits for the prefetch issued below
    if (esc_at(ck1)) {
#pragma unroll
      for (int c = 0; c < 3; ++c) bk1[c] = *at(b, (I)c * N + nk1);
    }
    if (esc_at(ck)) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        ac[c] = *at(a, (I)c * N + nk);
        ie[c] = (cond && need_c && (!kap || in_region(k))) ? *at(ieps, (I)c * N + nk) : S(0);
        sg[c] = (cond && need_c) ? *at(sgsrc, (I)c * N + nk) : S(0);
      }
    }
    const C tLz = (C)m.tLz[k], Lz = (C)m.Lz[k];
    // ---- stage P(k); P(k+1) of this cell (the z+1 terms, P(k) of the next plane)
    if (inrow) {
#pragma unroll
      for (int c = 0; c < 3; ++c) sP[c][row][slot] = Pk[c];
    }
    C Pk1[3];
    {
      S bv[3];
      const S4 ie1 = LUT ? mgroup(esc_at(ck1) ? 0 : ck1, 0) : S4{};
#pragma unroll
      for (int c = 0; c < 3; ++c)
        bv[c] = (!LUT || esc_at(ck1))
                    ? bk1[c]
                    : (S)(ie1[c] * (c == 0 ? bxk1 : (c == 1 ? byk1 : bz)));
#pragma unroll
      for (int c = 0; c < 3; ++c) Pk1[c] = (valid && kp1) ? (C)bv[c] * (C)Wk1[c] : C(0);
    }
    // ---- prefetch: W (and b without codes) of plane k+2; lH (and a, ieps,
    // sigma/kappa without codes) of plane k+1; codes and metric
    A Wn2[3], lhN[3];
    S bn2[3], acN[3], ieN[3], sgN[3];
    int cn2 = 0;
    S bxn2 = 0, byn2 = 0, axN = 0, ayN = 0;
    {
      const I n2 = off(k + 2);
      if (LUT) {
        cn2 = p.code[n2];
        bxn2 = tab(p.itAx, jj + ny * kc(k + 2));
        byn2 = tab(p.itAy, ii + nx * kc(k + 2));
        axN = tab(p.iAx, jj + ny * kc(k + 1));
        ayN = tab(p.iAy, ii + nx * kc(k + 1));
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        Wn2[c] = *at(Win, (I)c * N + n2);
        lhN[c] = need_lh ? *at(lHin, (I)c * N + nk1) : A(0);
        if (LUT) {
          bn2[c] = acN[c] = ieN[c] = sgN[c] = S(0);
        } else {
          bn2[c] = *at(b, (I)c * N + n2);
          acN[c] = need_lh ? *at(a, (I)c * N + nk1) : S(0);
          if (cond && need_c) {
            ieN[c] = (!kap || in_region(k + 1)) ? *at(ieps, (I)c * N + nk1) : S(0);
            sgN[c] = *at(sgsrc, (I)c * N + nk1);
          } else {
            ieN[c] = S(0); sgN[c] = S(0);
          }
        }
      }
    }
    const uint32_t kb = vreg_copy(kbits);
    const int zh = zface(g.pmcz, k, g.Nz), zp = zface(g.pecz, k, g.Nz);
    __syncthreads();

    // ---- lH'(k) and Q(k) (row waves: all 64 columns; halo wave: column i0-1)
    C hn[3] = {0, 0, 0};
    if (makes_q) {
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
      hn[0] = zsel(zh, kb & 1u) ? h0 + dt * (tLx * s0) : h0;
      hn[1] = zsel(zh, (kb >> 1) & 1u) ? h1 + dt * (tLy * s1) : h1;
      hn[2] = ((kb >> 2) & 1u) ? h2 + dt * (tLz * s2) : h2;
      S av[3];
      const S4 imu = LUT ? mgroup(esc_at(ck) ? 0 : ck, 1) : S4{};
#pragma unroll
      for (int c = 0; c < 3; ++c)
        av[c] = (!LUT || esc_at(ck))
                    ? ac[c]
                    : (S)((c == 0 ? axk : (c == 1 ? ayk : az)) * imu[c]);
#pragma unroll
      for (int c = 0; c < 3; ++c) sQ[c][row][slot] = valid ? (C)av[c] * hn[c] : C(0);
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
      const C m0 = zsel(zp, (kb >> 3) & 1u) ? Lx * e0 : C(0);
      const C m1 = zsel(zp, (kb >> 4) & 1u) ? Ly * e1 : C(0);
      const C m2 = ((kb >> 5) & 1u) ? Lz * e2 : C(0);
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
        S iev[3], sgv[3];
        if (!LUT || esc_at(ck)) {
#pragma unroll
          for (int c = 0; c < 3; ++c) { iev[c] = ie[c]; sgv[c] = sg[c]; }
        } else {
          const S4 ie4 = mgroup(ck, 0), sg4 = mgroup(ck, 2);
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            iev[c] = ie4[c];
            sgv[c] = (kap && c == 2 && onl) ? S(0) : sg4[c];
          }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const C iec = (C)iev[c], sgc = (C)sgv[c];
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
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if constexpr (sizeof(A) == 8) {  // FP64 outputs streamed past the caches
          __builtin_nontemporal_store((A)wn[c], atw(Wout, (I)c * N + nk));
          __builtin_nontemporal_store((A)hn[c], atw(lHout, (I)c * N + nk));
        } else {  // (4-byte streaming stores measured slower: 91 against 77 us)
          *atw(Wout, (I)c * N + nk) = (A)wn[c];
          *atw(lHout, (I)c * N + nk) = (A)hn[c];
        }
      }
    }
    Qpx = Qx;  // Q(k) of this cell seeds the z-1 terms of plane k+1
    Qpy = Qy;
    // advance the register pipeline. Plane k+1 stages sP after this plane's
    // second barrier (every read of this plane's sP is before it) and writes
    // sQ after its own first barrier (every read of this plane's sQ is before it).
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      Pk[c] = Pk1[c];
      Wk[c] = Wk1[c];
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

// waves per SIMD each row count is built for. FP32 arithmetic (79 VGPRs):
// two workgroups per CU with 7 or 9 rows (10 or 12 waves each), three with 5
// (8 waves), one with 13 (16 waves); FP64 arithmetic: one workgroup per CU
// (more would spill registers), and likewise with FP64 coefficients.
constexpr int adj_waves_per_simd(int TY, bool c64) {
  return c64 ? (TY + 3 + 3) / 4 : (TY == 5 ? 6 : TY == 7 ? 5 : TY == 9 ? 6 : 4);
}

template <typename S, typename A, int TY, bool LUT = false, typename C = double>
__global__ void __launch_bounds__(64 * (TY + 3), adj_waves_per_simd(TY, sizeof(C) == 8 || sizeof(S) == 8))
k_adj_fused(AdjStep<S, A> p) {
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
