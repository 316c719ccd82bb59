// Two leapfrog steps per launch (temporal blocking) for the wakefield step:
// material codes, masks folded into the coefficients (zero_masked), the beam
// line fed from the beam table, with or without conduction (COND: J = sigma E
// everywhere, recomputed from E as in k_fused4 and stored by the last launch
// of a call), rows of Nx <= 256 cells: one wave per row with FP32 (V = 4
// cells per lane), two waves per row with FP64 (V = 2).
//
// A workgroup owns TY rows j0 .. j0+TY-1 and a chunk of z-planes and marches
// up in z like k_fused4, but carries two time levels, the second one plane
// behind the first: at plane k it computes step n+1 (level 1: H1(k), E1(k))
// and step n+2 at plane k-1 (level 2: H2(k-1), E2(k-1)), and stores only
// level 2. Level 2 on the own rows needs level-1 values on rows j0-1 .. j0+TY
// (E1) and j0-2 .. j0+TY (H1), so the workgroup has TY + 4 rows of waves,
// row w = j0-2+w (the last one only stages E0 of row j0+TY+1); halo values are
// recomputed, never exchanged. Every value is formed by the same expression
// as in k_fused4 (h_row, e_sum, the coefficient recomputation from the
// codes), so both levels are bit-identical to two single steps.
//
// Per plane, with one barrier after each phase:
//   P1 H1(k) on waves 0..TY+2; stage H1(k)
//   P2 E1(k) on waves 1..TY+2; H2(k-1) on waves 1..TY+1 (reads E1(k-1)
//      staged in the previous P3); stage H2(k-1)
//   P3 stage E0(k+1); issue the loads of E0(k+2), of the codes of plane k+1
//      and (waves 0, 1) of its iAy / itAy row; E2(k-1) on waves 2..TY+1,
//      stores; stage E1(k) and the iAy / itAy row; issue the loads of H0(k+1)
// Loads are in flight during the E2 update before their use, no register is
// copied from a load before that use, and the per-plane scalars (Lz, tLz, the
// beam, iAx, itAx) are staged in LDS once per launch, so the waves rarely wait
// for memory. LDS rows are q-major, so a single-word read takes consecutive
// words in consecutive lanes (no bank conflicts). E and H are read once and
// written once per two steps.
#pragma once
#include "step_fused4.h"

namespace wf {

// A read of read-only data through the constant address space: with a uniform
// index it compiles to a scalar load (SGPR result, scalar cache).
template <typename T>
__device__ __forceinline__ T ldc(const T *p, int64_t i) {
  return ((const __attribute__((address_space(4))) T *)p)[i];
}

template <typename S, bool COND, bool DESC, int TY>
__global__ void __launch_bounds__(64 * (TY + 4) * (int)(sizeof(S) / 4))
k_fused4_t2(GridDev g, MetricDev<S> m, const S *__restrict__ Ein, const S *__restrict__ Hin,
            S *__restrict__ Eout, S *__restrict__ Hout, S *__restrict__ J, Coefs<S> cf, S dt,
            FusedAux x, const S *__restrict__ beam1, const S *__restrict__ beam2,
            S *__restrict__ cap) {
  using I = uint32_t;  // byte offsets: every 3N array within 4 GiB
  constexpr int V = 16 / sizeof(S);
  constexpr int WPR = sizeof(S) / 4;  // waves per row: 1 FP32, 2 FP64
  constexpr int LR = 64 * WPR;        // lanes per row
  constexpr int W = LR * V;           // cells per row (256)
  // LDS rows are q-major: the cell V*ln+q of row lane ln at slot q*QS + ln + 1,
  // so that one value per lane is LR consecutive words (no bank conflicts for
  // the single-word reads of the x neighbours and of cells one at a time).
  // Slot 0 of block V-1 is cell -1 and slot QS-1 of block 0 is cell W: both
  // stay 0.
  constexpr int QS = LR + 2;
  constexpr int RW = V * QS;
  constexpr int NM = 9;
  constexpr int RE0 = TY + 4, RH1 = TY + 3, RE1 = TY + 2, RH2 = TY + 1;
  constexpr int NT = 64 * (TY + 4) * WPR;
  __shared__ S sE0_[3 * RE0 * RW];  // E0(k), rows j0-2 .. j0+TY+1
  __shared__ S sH1_[3 * RH1 * RW];  // H1(k), rows j0-2 .. j0+TY
  __shared__ S sE1_[3 * RE1 * RW];  // E1(k-1) until P3 stages E1(k), rows j0-1 .. j0+TY
  __shared__ S sH2_[3 * RH2 * RW];  // H2(k-1), rows j0-1 .. j0+TY-1
  // iAy and itAy of planes k-1, k, k+1 (slot plane % 3), lane-major, written
  // by waves 0 and 1
  __shared__ __attribute__((aligned(16))) S sCy_[3 * 2 * W];
  extern __shared__ unsigned char s_dyn[];
  S *sMat = reinterpret_cast<S *>(s_dyn);
  const auto sE0 = [&](int c, int r) { return sE0_ + (c * RE0 + r) * RW; };
  const auto sH1 = [&](int c, int r) { return sH1_ + (c * RH1 + r) * RW; };
  const auto sE1 = [&](int c, int r) { return sE1_ + (c * RE1 + r) * RW; };
  const auto sH2 = [&](int c, int r) { return sH2_ + (c * RH2 + r) * RW; };
  const auto sCy = [&](int kk, int which) { return sCy_ + (((kk + 3) % 3) * 2 + which) * W; };
  typedef S vec __attribute__((ext_vector_type(V)));
  const auto ldsv = [&](const S *p, S (&o)[V]) {
    const vec v = *reinterpret_cast<const vec *>(p);
#pragma unroll
    for (int q = 0; q < V; ++q) o[q] = v[q];
  };

  for (int t = threadIdx.x; t < cf.nmat * NM; t += NT) sMat[t] = cf.mat[t];
  // the x-halo slots (cells -1 and W) are zero for the whole launch
  for (int t = threadIdx.x; t < 3 * (RE0 + RH1 + RE1 + RH2) * 2; t += NT) {
    int r = t >> 1;  // component row over the four buffers
    const int slot = (t & 1) ? QS - 1 : (V - 1) * QS;
    S *base;
    if (r < 3 * RE0) {
      base = sE0_;
    } else if ((r -= 3 * RE0) < 3 * RH1) {
      base = sH1_;
    } else if ((r -= 3 * RH1) < 3 * RE1) {
      base = sE1_;
    } else {
      r -= 3 * RE1;
      base = sH2_;
    }
    base[r * RW + slot] = S(0);
  }

  const int tile = xcd_tile(blockIdx.x, x.nty * x.ntz);
  const int tz = tile / x.nty, ty = tile - tz * x.nty;
  const int j0 = ty * TY, k0 = tz * x.KZ;
  const int kend = min(k0 + x.KZ, g.Nz);
  // the row of a wave is uniform: row flags, row metrics and branches are scalar
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int w = wave / WPR;                              // row of the workgroup
  const int ln = (wave - w * WPR) * 64 + (threadIdx.x & 63);  // lane in the row
  const int j = j0 - 2 + w;
  const bool rowok = (j >= 0) && (j < g.Ny);
  const int xi = V * ln;
  // LDS slot of cell xi + q, -1 <= q <= V (q is a constant after unrolling)
  const auto atx = [&](int q) {
    return q < 0 ? (V - 1) * QS + ln : q >= V ? ln + 2 : q * QS + ln + 1;
  };
  const bool lact = rowok && (xi < g.Nx);
  const I N = (I)g.N, sz = (I)g.sz;
  const I col = lact ? (I)xi + (I)g.sy * (I)j : (I)0;
  const bool doH1 = w <= TY + 2, doE1 = w >= 1 && w <= TY + 2;
  const bool doH2 = w >= 1 && w <= TY + 1, own = w >= 2 && w <= TY + 1;

  S Lx[V], tLx[V], iAz[V], itAz[V];
#pragma unroll
  for (int q = 0; q < V; ++q) {
    Lx[q] = lact ? m.Lx[xi + q] : S(0);
    tLx[q] = lact ? m.tLx[xi + q] : S(0);
    iAz[q] = S(0);
    itAz[q] = S(0);
  }
  if (lact) {
    const I o = (I)xi + (I)g.Nx * (I)j;
    ldv<V>(cf.iAz, o, iAz);
    ldv<V>(cf.itAz, o, itAz);
  }
  const S Ly = rowok ? ldc(m.Ly, j) : S(0), tLy = rowok ? ldc(m.tLy, j) : S(0);
  const int qbeam = (lact && j == x.iys && x.ixs >= xi && x.ixs < xi + V) ? x.ixs - xi : -1;
  // Ez window capture: own rows only, one sample per level
  int capq0 = V, capq1 = 0;
  int64_t capb1 = 0, capb2 = 0;
  if (cap && own && lact && j >= x.cap_iy0 && j < x.cap_iy0 + x.cap_ny) {
    const int lo = max(xi, x.cap_ix0), hi = min(xi + V, x.cap_ix0 + x.cap_nx);
    if (lo < hi) {
      capq0 = lo - xi; capq1 = hi - xi;
      capb1 = ((int64_t)(j - x.cap_iy0) * x.cap_samples + x.cap_sample) * g.Nz;
      capb2 = capb1 + g.Nz;
    }
  }
  const bool cap1 = x.cap_sample < x.cap_samples, cap2 = x.cap_sample + 1 < x.cap_samples;
  const int64_t capl = (int64_t)x.cap_ny * x.cap_samples * g.Nz;  // stride of one x line

  const int kstart = max(k0 - 2, 0);
  const int kE1 = max(k0 - 1, 0);  // first plane with E1 (and H2)
  // per-plane values of planes kstart .. kend, staged once (no scalar memory
  // load in the z-march: its out-of-order return would make every LDS wait
  // also wait for it): Lz, tLz, the beam of both levels, and iAx, itAx of the
  // workgroup's rows
  const int NP = kend - kstart + 1;
  S *sPl = sMat + cf.nmat * NM;     // [4][NP]
  S *sRow = sPl + 4 * NP;           // [2][TY+4][NP]
  for (int t = threadIdx.x; t < 4 * NP; t += NT) {
    const int which = t / NP, kk = kstart + t - which * NP;
    S v = S(0);
    if (kk < g.Nz) {
      if (which == 0) v = m.Lz[kk];
      else if (which == 1) v = m.tLz[kk];
      else if (which == 2) v = beam1 ? beam1[kk] : S(0);
      else v = beam2 ? beam2[kk] : S(0);
    }
    sPl[t] = v;
  }
  for (int t = threadIdx.x; t < 2 * (TY + 4) * NP; t += NT) {
    const int which = t / ((TY + 4) * NP), rem = t - which * (TY + 4) * NP;
    const int r = rem / NP, kk = kstart + rem - r * NP, jj = j0 - 2 + r;
    S v = S(0);
    if (jj >= 0 && jj < g.Ny && kk < g.Nz)
      v = (which ? cf.itAx : cf.iAx)[(int64_t)jj + (int64_t)g.Ny * kk];
    sRow[t] = v;
  }
  const auto plv = [&](int which, int kk) { return sPl[which * NP + kk - kstart]; };
  const auto rowv = [&](int which, int kk) {
    return sRow[(which * (TY + 4) + w) * NP + kk - kstart];
  };
  S Ek[3][V], Ek1[3][V];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int q = 0; q < V; ++q) { Ek[c][q] = S(0); Ek1[c][q] = S(0); }
  if (lact) {
    const I n = col + sz * (I)kstart;
    ldv<V>(Ein, n, Ek[0]); ldv<V>(Ein, N + n, Ek[1]); ldv<V>(Ein, 2 * N + n, Ek[2]);
    if (kstart + 1 < g.Nz) {
      ldv<V>(Ein, n + sz, Ek1[0]); ldv<V>(Ein, N + n + sz, Ek1[1]);
      ldv<V>(Ein, 2 * N + n + sz, Ek1[2]);
    }
  }
  // carried between planes: H1(k-1), H2(k-2) (x, y) and the codes of plane
  // k-1 (level 2's coefficients; its metric values come from the tables)
  S H1p[3][V], H2p[2][V];
#pragma unroll
  for (int q = 0; q < V; ++q) {
    H1p[0][q] = H1p[1][q] = H1p[2][q] = S(0);
    H2p[0][q] = H2p[1][q] = S(0);
  }
  uint32_t codep = 0;
  // the plane in hand, loaded one plane ahead: H0 and the codes; waves 0 and 1
  // also carry its iAy / itAy row on the way to sCy
  struct Plane {
    S h[3][V], cy[V];
    uint32_t code;
  };
  // iAy / itAy of plane kk (issued early: few registers)
  const auto load_plane_a = [&](int kk, Plane &P) {
#pragma unroll
    for (int q = 0; q < V; ++q) P.cy[q] = S(0);
    if (kk < g.Nz && w < 2 && xi < g.Nx)
      ldv<V>(w == 0 ? cf.iAy : cf.itAy, (I)xi + (I)g.Nx * (I)kk, P.cy);
  };
  // the codes of plane kk
  const auto load_code = [&](int kk, Plane &P) {
    P.code = (kk < g.Nz && lact && doH1) ? ldcode<V>(cf.code, col + sz * (I)kk) : 0u;
  };
  // H0 of plane kk (issued last: its registers live until H1)
  const auto load_plane_b = [&](int kk, Plane &P) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < V; ++q) P.h[c][q] = S(0);
    if (kk < g.Nz && lact && doH1) {
      const I n = col + sz * (I)kk;
      ldv<V>(Hin, n, P.h[0]); ldv<V>(Hin, N + n, P.h[1]); ldv<V>(Hin, 2 * N + n, P.h[2]);
    }
  };
  const auto put_ring = [&](int kk, const Plane &P) {
    if (w < 2 && kk <= g.Nz) {
      vec v;
#pragma unroll
      for (int q = 0; q < V; ++q) v[q] = P.cy[q];
      *reinterpret_cast<vec *>(sCy(kk, w) + xi) = v;
    }
  };
  Plane P;
  load_plane_a(kstart, P);
  load_code(kstart, P);
  load_plane_b(kstart, P);
  put_ring(kstart, P);
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int q = 0; q < V; ++q) sE0(c, w)[atx(q)] = lact ? Ek[c][q] : S(0);
  __syncthreads();

  for (int k = kstart; k <= kend; ++k) {
    const bool pl1 = k < g.Nz;                // level-1 plane exists
    const int p = k - 1;                      // level-2 plane
    const bool l2h = p >= kE1 && p < kend;    // H2(p) needed
    const bool l2e = p >= k0 && p < kend;     // E2(p) stored
    const uint32_t code = P.code;
    const S iax = rowv(0, k), itax = rowv(1, k), Lz = plv(0, k), tLz = plv(1, k);
    const int pp = max(p, kstart);  // level 2 reads only for p >= kE1 >= kstart
    const S iaxp = rowv(0, pp), itaxp = rowv(1, pp), Lzp = plv(0, pp), tLzp = plv(1, pp);

    // ---- P1: H1(k)
    S H1[3][V];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < V; ++q) H1[c][q] = S(0);
    if (lact && pl1 && doH1) {
      S c0y[V];
      ldsv(sCy(k, 0) + xi, c0y);
      S EX[V + 1], EY[V + 1], EZ[V + 1];
#pragma unroll
      for (int q = 0; q < V + 1; ++q) {
        EX[q] = sE0(0, w)[atx(q)]; EY[q] = sE0(1, w)[atx(q)];
        EZ[q] = sE0(2, w)[atx(q)];
      }
#pragma unroll
      for (int q = 0; q < V; ++q) {
        const S *t = sMat + NM * (int)((code >> (8 * q)) & 0xffu);
        const S a0 = (S)(iax * t[0]), a1 = (S)(c0y[q] * t[1]), a2 = (S)(iAz[q] * t[2]);
        H1[0][q] = h_row<S, S, DESC>(P.h[0][q], a0, dt, Ly, Lz, EY[q], Ek1[1][q], EZ[q],
                                     sE0(2, w + 1)[atx(q)], +1, -1);
        H1[1][q] = h_row<S, S, DESC>(P.h[1][q], a1, dt, Lx[q], Lz, EX[q], Ek1[0][q], EZ[q],
                                     EZ[q + 1], -1, +1);
        H1[2][q] = h_row<S, S, DESC>(P.h[2][q], a2, dt, Lx[q], Ly, EX[q],
                                     sE0(0, w + 1)[atx(q)], EY[q], EY[q + 1], +1, -1);
      }
    }
    if (doH1) {
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int q = 0; q < V; ++q) sH1(c, w)[atx(q)] = H1[c][q];
    }
    __syncthreads();

    // ---- P2: E1(k), then H2(k-1)
    S E1[3][V];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < V; ++q) E1[c][q] = S(0);
    if (lact && pl1 && doE1 && k >= kE1) {
      S c1y[V];
      ldsv(sCy(k, 1) + xi, c1y);
      S HX[V + 1], HY[V + 1], HZ[V + 1];  // [0] = -x neighbour
#pragma unroll
      for (int q = 0; q < V + 1; ++q) {
        HX[q] = sH1(0, w)[atx(q - 1)];
        HY[q] = sH1(1, w)[atx(q - 1)];
        HZ[q] = sH1(2, w)[atx(q - 1)];
      }
#pragma unroll
      for (int q = 0; q < V; ++q) {
        const S *t = sMat + NM * (int)((code >> (8 * q)) & 0xffu);
        const S b0 = (S)(t[3] * itax), b1 = (S)(t[4] * c1y[q]), b2 = (S)(t[5] * itAz[q]);
        S d0 = e_sum<S>(b0, tLy, tLz, H1p[1][q], HY[q + 1], sH1(2, w - 1)[atx(q)],
                        HZ[q + 1], +1, -1);
        S d1 = e_sum<S>(b1, tLx[q], tLz, H1p[0][q], HX[q + 1], HZ[q], HZ[q + 1], -1, +1);
        S d2 = e_sum<S>(b2, tLx[q], tLy, sH1(0, w - 1)[atx(q)], HX[q + 1], HY[q],
                        HY[q + 1], +1, -1);
        const S e0 = sE0(0, w)[atx(q)], e1 = sE0(1, w)[atx(q)];
        const S e2 = sE0(2, w)[atx(q)];  // E0(k) as staged
        if (COND) {  // J = sigma E0 (the beam's current on its line)
          d0 = d0 - t[3] * (S)(t[6] * e0);
          d1 = d1 - t[4] * (S)(t[7] * e1);
          d2 = d2 - t[5] * ((q == qbeam) ? plv(2, k) : (S)(t[8] * e2));
        } else if (q == qbeam) {
          d2 = d2 - t[5] * plv(2, k);
        }
        E1[0][q] = (S)(e0 + dt * d0);
        E1[1][q] = (S)(e1 + dt * d1);
        E1[2][q] = (S)(e2 + dt * d2);
        if (cap1 && k >= k0 && k < kend && q >= capq0 && q < capq1)
          cap[capb1 + (int64_t)(q + xi - x.cap_ix0) * capl + k] = E1[2][q];
      }
    }
    S H2[3][V];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < V; ++q) H2[c][q] = S(0);
    if (lact && doH2 && l2h) {
      S c0yp[V];
      ldsv(sCy(p, 0) + xi, c0yp);
      // E1(p): own row from sE1 row w-1 (x neighbour in slot q = V), row j+1 from row w
      S EX[V + 1], EY[V + 1], EZ[V + 1];
#pragma unroll
      for (int q = 0; q < V + 1; ++q) {
        EX[q] = sE1(0, w - 1)[atx(q)]; EY[q] = sE1(1, w - 1)[atx(q)];
        EZ[q] = sE1(2, w - 1)[atx(q)];
      }
#pragma unroll
      for (int q = 0; q < V; ++q) {
        const S *t = sMat + NM * (int)((codep >> (8 * q)) & 0xffu);
        const S a0 = (S)(iaxp * t[0]), a1 = (S)(c0yp[q] * t[1]), a2 = (S)(iAz[q] * t[2]);
        H2[0][q] = h_row<S, S, DESC>(H1p[0][q], a0, dt, Ly, Lzp, EY[q], E1[1][q], EZ[q],
                                     sE1(2, w)[atx(q)], +1, -1);
        H2[1][q] = h_row<S, S, DESC>(H1p[1][q], a1, dt, Lx[q], Lzp, EX[q], E1[0][q], EZ[q],
                                     EZ[q + 1], -1, +1);
        H2[2][q] = h_row<S, S, DESC>(H1p[2][q], a2, dt, Lx[q], Ly, EX[q],
                                     sE1(0, w)[atx(q)], EY[q], EY[q + 1], +1, -1);
      }
    }
    if (doH2) {
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int q = 0; q < V; ++q) sH2(c, w - 1)[atx(q)] = H2[c][q];
    }
    // H2(k-1) of the own rows is final: stored now; its x, y components are
    // kept for E2(k) of the next plane
    S H2c[2][V];
#pragma unroll
    for (int q = 0; q < V; ++q) { H2c[0][q] = H2[0][q]; H2c[1][q] = H2[1][q]; }
    if (lact && own && l2e) {
      const I np = col + sz * (I)p;
      stv_nt<V>(Hout, np, H2[0]); stv_nt<V>(Hout, N + np, H2[1]);
      stv_nt<V>(Hout, 2 * N + np, H2[2]);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < V; ++q) H1p[c][q] = H1[c][q];
    __syncthreads();

    // ---- P3: stage E0(k+1) (sE0 is free once E1(k) is done); loads of plane
    // k+1 and of E0(k+2); E2(k-1) on the own rows, stores; stage E1(k)
    const bool more = k < kend;
    const S bv2 = plv(3, pp);  // the beam at plane p, level 2
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < V; ++q) sE0(c, w)[atx(q)] = Ek1[c][q];
    load_plane_a(more ? k + 1 : g.Nz + 1, P);
    load_code(more ? k + 1 : g.Nz, P);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < V; ++q) Ek1[c][q] = S(0);
    if (lact && more && k + 2 < g.Nz) {
      const I n2 = col + sz * (I)(k + 2);
      ldv<V>(Ein, n2, Ek1[0]); ldv<V>(Ein, N + n2, Ek1[1]); ldv<V>(Ein, 2 * N + n2, Ek1[2]);
    }
    if (lact && own && l2e) {
      S c1yp[V];
      ldsv(sCy(p, 1) + xi, c1yp);
      S HX[V + 1], HY[V + 1], HZ[V + 1];
#pragma unroll
      for (int q = 0; q < V + 1; ++q) {
        HX[q] = sH2(0, w - 1)[atx(q - 1)];
        HY[q] = sH2(1, w - 1)[atx(q - 1)];
        HZ[q] = sH2(2, w - 1)[atx(q - 1)];
      }
      const I np = col + sz * (I)p;
      S E2[3][V], Jn[3][V];
#pragma unroll
      for (int q = 0; q < V; ++q) {
        const S *t = sMat + NM * (int)((codep >> (8 * q)) & 0xffu);
        const S b0 = (S)(t[3] * itaxp), b1 = (S)(t[4] * c1yp[q]), b2 = (S)(t[5] * itAz[q]);
        S d0 = e_sum<S>(b0, tLy, tLzp, H2p[1][q], HY[q + 1], sH2(2, w - 2)[atx(q)],
                        HZ[q + 1], +1, -1);
        S d1 = e_sum<S>(b1, tLx[q], tLzp, H2p[0][q], HX[q + 1], HZ[q], HZ[q + 1], -1, +1);
        S d2 = e_sum<S>(b2, tLx[q], tLy, sH2(0, w - 2)[atx(q)], HX[q + 1], HY[q],
                        HY[q + 1], +1, -1);
        const S e0 = sE1(0, w - 1)[atx(q)], e1 = sE1(1, w - 1)[atx(q)];
        const S e2 = sE1(2, w - 1)[atx(q)];  // E1(k-1)
        if (COND) {
          d0 = d0 - t[3] * (S)(t[6] * e0);
          d1 = d1 - t[4] * (S)(t[7] * e1);
          d2 = d2 - t[5] * ((q == qbeam) ? bv2 : (S)(t[8] * e2));
        } else if (q == qbeam) {
          d2 = d2 - t[5] * bv2;
        }
        E2[0][q] = (S)(e0 + dt * d0);
        E2[1][q] = (S)(e1 + dt * d1);
        E2[2][q] = (S)(e2 + dt * d2);
        if (!COND && q == qbeam) *atw(J, 2 * N + np + (I)q) = bv2;
        if (COND && !x.skip_j) {  // J = sigma E2 by the call's last launch
#pragma unroll
          for (int c = 0; c < 3; ++c) Jn[c][q] = (S)(t[6 + c] * E2[c][q]);
        }
        if (cap2 && q >= capq0 && q < capq1)
          cap[capb2 + (int64_t)(q + xi - x.cap_ix0) * capl + p] = E2[2][q];
      }
      stv_nt<V>(Eout, np, E2[0]); stv_nt<V>(Eout, N + np, E2[1]);
      stv_nt<V>(Eout, 2 * N + np, E2[2]);
      if (COND && !x.skip_j) {
        stv_nt<V>(J, np, Jn[0]); stv_nt<V>(J, N + np, Jn[1]); stv_nt<V>(J, 2 * N + np, Jn[2]);
      }
    }
#pragma unroll
    for (int q = 0; q < V; ++q) { H2p[0][q] = H2c[0][q]; H2p[1][q] = H2c[1][q]; }
    if (doE1) {
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int q = 0; q < V; ++q) sE1(c, w - 1)[atx(q)] = E1[c][q];
    }
    put_ring(k + 1, P);
    load_plane_b(more ? k + 1 : g.Nz, P);
    // level 2 of the next plane uses this plane's coefficients
    codep = code;
    __syncthreads();
  }
}

}  // namespace wf
