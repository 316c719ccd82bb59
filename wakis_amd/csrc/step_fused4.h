// Fused single-pass leapfrog step, x-vectorised (fast path for Nx % 4 == 0).
//
// Same per-row arithmetic as k_fused (step_fused.h), organised for CDNA issue
// rate: each lane owns 4 consecutive x cells (16-byte loads/stores, one
// 32-bit load for four material codes), each wave owns one row segment of
// 256 cells, and a workgroup holds TY own rows plus one extra wave that
// recomputes H on the row below the tile (j0-1) and stages the row above
// (j0+TY). x neighbours come from the lane's own registers or from an aligned
// LDS row (slot 3 = i0-1, slots 4..259 = cells, slot 260 = i0+256); the
// workgroup marches up its chunk of z planes with E two planes and the other
// per-plane inputs one plane ahead. PEC/PMC column masks and absent
// neighbours are applied once, when a value is staged (as 0).
//
// Full-row tiles (FR): for rows whose length is not close to a multiple of the
// 64-lane segment (e.g. 200 cells: 22% of the lanes idle), a workgroup takes
// whole rows instead. Thread t then owns the V cells at x = V * (t % Q) of
// tile row t / Q (Q = Nx / V lanes per row; tile row 0 is the halo row), the
// tile has x.rows own rows, there are no x halos, and the LDS rows (Nx + 2V
// slots) live in dynamic LDS ahead of the material table. Only the thread ->
// (row, cell) map and the LDS strides change; the arithmetic is the same.
#pragma once
#include "fit_common.h"
#include "step_fused.h"

#include <type_traits>

namespace wf {

template <typename S>
struct alignas(16) Q4 {
  S v[4];
};

template <typename S, typename I>
__device__ __forceinline__ const S *at(const S *p, I idx) {
  if constexpr (sizeof(I) == 4)
    return reinterpret_cast<const S *>(reinterpret_cast<const char *>(p) +
                                       (uint32_t)(idx * (uint32_t)sizeof(S)));
  else
    return p + idx;
}
template <typename S, typename I>
__device__ __forceinline__ S *atw(S *p, I idx) {
  if constexpr (sizeof(I) == 4)
    return reinterpret_cast<S *>(reinterpret_cast<char *>(p) + (uint32_t)(idx * (uint32_t)sizeof(S)));
  else
    return p + idx;
}
// V consecutive values starting at a multiple-of-V index: one 16-byte access
// when V*sizeof(S) == 16, else V element accesses
template <int V, typename S, typename I>
__device__ __forceinline__ void ldv(const S *__restrict__ p, I idx, S (&o)[V]) {
  const S *q = at(p, idx);
  if constexpr (V * sizeof(S) == 16 && sizeof(S) == 4) {
    const float4 t = *reinterpret_cast<const float4 *>(q);
    o[0] = t.x; o[1] = t.y; o[2] = t.z; o[3] = t.w;
  } else if constexpr (V * sizeof(S) == 16) {
    const double2 t = *reinterpret_cast<const double2 *>(q);
    o[0] = t.x; o[1] = t.y;
  } else {
#pragma unroll
    for (int v = 0; v < V; ++v) o[v] = q[v];
  }
}
template <int V, typename S, typename I>
__device__ __forceinline__ void stv(S *__restrict__ p, I idx, const S (&v)[V]) {
  S *q = atw(p, idx);
  if constexpr (V * sizeof(S) == 16 && sizeof(S) == 4) {
    *reinterpret_cast<float4 *>(q) = make_float4(v[0], v[1], v[2], v[3]);
  } else if constexpr (V * sizeof(S) == 16) {
    *reinterpret_cast<double2 *>(q) = make_double2(v[0], v[1]);
  } else {
#pragma unroll
    for (int u = 0; u < V; ++u) q[u] = v[u];
  }
}
// stv with non-temporal stores (streamed past the caches): the step's outputs E,
// H, J, which the launch never reads back; leaving L2 to the input rows that
// neighbouring tiles share took the 256x256x512 step from 0.365 to 0.355 ms
// (FP32) and from 0.801 to 0.770 ms (FP64)
template <int V, typename S, typename I>
__device__ __forceinline__ void stv_nt(S *__restrict__ p, I idx, const S (&v)[V]) {
  S *q = atw(p, idx);
  if constexpr (V * sizeof(S) == 16 && sizeof(S) == 4) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const f4 t = {v[0], v[1], v[2], v[3]};
    __builtin_nontemporal_store(t, reinterpret_cast<f4 *>(q));
  } else if constexpr (V * sizeof(S) == 16) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    const d2 t = {v[0], v[1]};
    __builtin_nontemporal_store(t, reinterpret_cast<d2 *>(q));
  } else {
#pragma unroll
    for (int u = 0; u < V; ++u) __builtin_nontemporal_store(v[u], q + u);
  }
}
// V consecutive uint8 material codes packed little-endian into one word
template <int V, typename I>
__device__ __forceinline__ uint32_t ldcode(const uint8_t *__restrict__ p, I idx) {
  if constexpr (V == 4) return *reinterpret_cast<const uint32_t *>(p + idx);
  else if constexpr (V == 2) return *reinterpret_cast<const uint16_t *>(p + idx);
  else return p[idx];
}

// MK = false: the caller guarantees that E is zero on every PEC-masked edge and
// H on every PMC-masked face (and that no source writes there), and passes
// coefficients with the row masks folded in. The staged values then need no
// boundary masks: with zero mask words every mask test below folds to "true"
// at compile time, which frees the registers the per-lane masks take.
// XCD-aware block -> tile map: the hardware deals consecutive block ids to the
// 8 XCDs in turn, so block b runs on XCD b % 8; give each XCD a contiguous run
// of tiles so that neighbouring tiles (which share halo rows) meet in one L2.
__device__ __forceinline__ int xcd_tile(int blin, int nb) {
  const int xcd = blin & 7, slot = blin >> 3;
  const int q8 = nb >> 3, r8 = nb & 7;
  return xcd * q8 + (xcd < r8 ? xcd : r8) + slot;
}

// One workgroup's z-march over tile `tile` of the grid (the whole step of that
// tile); shared by the single-grid and the batched kernels.
// WIDE: 64-bit element offsets (arrays past 4 GiB); otherwise 32-bit byte
// offsets from the array base (every 3N array within 4 GiB).
template <typename S, int JM, bool COND, bool LUT, bool DESC, int TY, bool MK, int V, bool FR,
          bool ESC, bool WIDE = false>
__device__ __forceinline__ void fused4_tile(const GridDev &g, const MetricDev<S> &m,
                                            const S *__restrict__ Ein, const S *__restrict__ Hin,
                                            S *__restrict__ Eout, S *__restrict__ Hout,
                                            S *__restrict__ J, const Coefs<S> &cf, S dt,
                                            const FusedAux &x, const S *__restrict__ beam_row,
                                            S *__restrict__ cap, int tile) {
  using I = typename std::conditional<WIDE, uint64_t, uint32_t>::type;
  // V cells per lane: 16 / sizeof(S) for 16-byte accesses, or 1 for narrow grids
  constexpr int OFF = V;                       // cell xi at LDS slot xi + OFF
  constexpr int NM = 9;
  // segment cells, own rows, LDS row length and lanes per row (FR: run time)
  const int W = FR ? g.Nx : 64 * V;
  const int TYr = FR ? x.rows : TY;
  const int RW = W + 2 * V;
  const int Q = FR ? x.qrow : 64;
  // one buffer each: a plane's E staging happens after the barrier that ends
  // every read of the previous plane's E, and its H writes after the barrier
  // that ends every read of the previous plane's H.
  // sE rows: 0 = j0-1, 1..TYr own, TYr+1 = j0+TYr; sH rows: 0 = j0-1, 1..TYr
  // own (PMC-masked)
  extern __shared__ unsigned char s_dyn[];
  S *sEp, *sHp, *sMat;
  if constexpr (FR) {
    sEp = reinterpret_cast<S *>(s_dyn);
    sHp = sEp + 3 * (TYr + 2) * RW;
    sMat = sHp + 3 * (TYr + 1) * RW;
  } else {
    __shared__ S sE_[3 * (TY + 2) * (64 * V + 2 * V)];
    __shared__ S sH_[3 * (TY + 1) * (64 * V + 2 * V)];
    sEp = sE_;
    sHp = sH_;
    sMat = reinterpret_cast<S *>(s_dyn);
  }
  const auto sE = [&](int c, int r) { return sEp + (c * (TYr + 2) + r) * RW; };
  const auto sH = [&](int c, int r) { return sHp + (c * (TYr + 1) + r) * RW; };
  if (LUT) {
    const int nthr = FR ? (int)blockDim.x : 64 * (TY + 1);
    for (int t = threadIdx.x; t < cf.nmat * NM; t += nthr) sMat[t] = cf.mat[t];
  }

  const int tzl = tile / (x.ntx * x.nty);
  const int rem = tile - tzl * x.ntx * x.nty;
  const int tyl = rem / x.ntx;
  const int tz = tzl + x.tz0, tyb = tyl + x.ty0, txb = rem - tyl * x.ntx + x.tx0;
  const int i0 = txb * W, j0 = tyb * TYr, k0 = tz * x.KZ;
  const int kend = min(k0 + x.KZ, g.Nz);

  const int lane = FR ? (int)threadIdx.x % Q : (int)(threadIdx.x & 63);
  const int w = FR ? (int)threadIdx.x / Q : (int)(threadIdx.x >> 6);
  const int LL = Q - 1;                          // last lane of a row segment
  const bool wok = !FR || w <= TYr;              // FR: threads past the tile rows idle
  const int j = (w == 0) ? j0 - 1 : j0 + w - 1;  // row of this wave
  const bool rowok = wok && (j >= 0) && (j < g.Ny);
  const int xi = V * lane;
  const int ic = i0 + xi;
  const bool lact = rowok && (ic < g.Nx);          // Nx % V == 0: all V cells or none
  const bool own = (w >= 1) && lact;
  const I N = (I)g.N, sz = (I)g.sz;
  const I col = lact ? (I)ic + (I)g.sy * (I)j : (I)0;
  const int pz = MK ? g.pecz : 0, hz = MK ? g.pmcz : 0;
  const int jt = j0 + TYr;
  const bool top = (w == 0) && (jt < g.Ny) && (ic < g.Nx);
  const I tcol = top ? (I)ic + (I)g.sy * (I)jt : (I)0;
  const bool hasL = rowok && (lane == 0) && (i0 > 0);
  const bool hasR = rowok && (lane == LL) && (i0 + W < g.Nx);
  const bool topR = (w == 0) && (lane == LL) && (jt < g.Ny) && (i0 + W < g.Nx);
  const bool topL = (w == 0) && (lane == 0) && (jt < g.Ny) && (i0 > 0);

  const bool mXj = rowok && ypart(pz, j, g.Ny);
  const bool mXt = ypart(pz, jt, g.Ny);
  bool mYi[V], mZi[V], mZt[V], mYt[V];
#pragma unroll
  for (int q = 0; q < V; ++q) {
    mYi[q] = lact && xpart(pz, ic + q, g.Nx);
    mZi[q] = lact && mz(pz, ic + q, j, g);
    mZt[q] = top && mz(pz, ic + q, jt, g);
    mYt[q] = top && xpart(pz, ic + q, g.Nx);
  }
  const bool mYL = xpart(pz, i0 - 1, g.Nx), mZL = rowok && (i0 > 0) && mz(pz, i0 - 1, j, g);
  const bool mYR = xpart(pz, i0 + W, g.Nx), mZR = rowok && (i0 + W < g.Nx) && mz(pz, i0 + W, j, g);
  const bool mZtR = topR && mz(pz, i0 + W, jt, g);

  S Lx[V], tLx[V];
#pragma unroll
  for (int q = 0; q < V; ++q) {
    Lx[q] = lact ? m.Lx[ic + q] : S(0);
    tLx[q] = lact ? m.tLx[ic + q] : S(0);
  }
  const S Ly = rowok ? m.Ly[j] : S(0), tLy = rowok ? m.tLy[j] : S(0);
  const S LxL = (i0 > 0) ? m.Lx[i0 - 1] : S(0);
  const int qbeam = (own && j == x.iys && x.ixs >= ic && x.ixs < ic + V) ? x.ixs - ic : -1;
  int capq0 = V, capq1 = 0;
  int64_t capbase = 0;
  if (cap && own && x.cap_sample < x.cap_samples && j >= x.cap_iy0 && j < x.cap_iy0 + x.cap_ny) {
    const int lo = max(ic, x.cap_ix0), hi = min(ic + V, x.cap_ix0 + x.cap_nx);
    if (lo < hi) {
      capq0 = lo - ic; capq1 = hi - ic;
      capbase = ((int64_t)(j - x.cap_iy0) * x.cap_samples + x.cap_sample) * g.Nz;
    }
  }
  S iAz[V], itAz[V];
  S iAzL = 0;
#pragma unroll
  for (int q = 0; q < V; ++q) { iAz[q] = S(0); itAz[q] = S(0); }
  if (LUT && lact) {
    const I o = (I)ic + (I)g.Nx * (I)j;
    ldv<V>(cf.iAz, o, iAz);
    ldv<V>(cf.itAz, o, itAz);
  }
  if (LUT && hasL) iAzL = cf.iAz[(int64_t)(i0 - 1) + (int64_t)g.Nx * j];

  const int kstart = (k0 > 0) ? k0 - 1 : k0;
  S Ek[3][V], Ek1[3][V];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int q = 0; q < V; ++q) { Ek[c][q] = S(0); Ek1[c][q] = S(0); }
  if (lact) {
    const I n = col + sz * (I)kstart;
    ldv<V>(Ein, n, Ek[0]); ldv<V>(Ein, N + n, Ek[1]); ldv<V>(Ein, 2 * N + n, Ek[2]);
    if (kstart + 1 < g.Nz) {
      ldv<V>(Ein, n + sz, Ek1[0]); ldv<V>(Ein, N + n + sz, Ek1[1]); ldv<V>(Ein, 2 * N + n + sz, Ek1[2]);
    }
  }
  if (LUT) __syncthreads();
  S Hpx[V], Hpy[V];  // PMC-masked H_new at k-1
#pragma unroll
  for (int q = 0; q < V; ++q) { Hpx[q] = S(0); Hpy[q] = S(0); }

  for (int k = kstart; k < kend; ++k) {
    const bool body = (k >= k0);
    const I nk = col + sz * (I)k;
    const bool kp1 = (k + 1 < g.Nz), kp2 = (k + 2 < g.Nz);
    // ---- loads of this plane (issued before the staging barrier) and E(k+2)
    S Ho[3][V], En2[3][V];
    S c0y[V], c1y[V], iax = 0, itax = 0;
    uint32_t code = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < V; ++q) { Ho[c][q] = S(0); En2[c][q] = S(0); }
#pragma unroll
    for (int q = 0; q < V; ++q) { c0y[q] = S(0); c1y[q] = S(0); }
    S ao[3][V], bo[3][V], ie[3][V], sg[3][V];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < V; ++q) { ao[c][q] = S(0); bo[c][q] = S(0); ie[c][q] = S(0); sg[c][q] = S(0); }
    if (lact && (w >= 1 || body)) {
      ldv<V>(Hin, nk, Ho[0]); ldv<V>(Hin, N + nk, Ho[1]); ldv<V>(Hin, 2 * N + nk, Ho[2]);
      if (LUT) {
        code = ldcode<V>(cf.code, nk);
        iax = cf.iAx[(int64_t)j + (int64_t)g.Ny * k];
        ldv<V>(cf.iAy, (I)ic + (I)g.Nx * (I)k, c0y);
        if (w >= 1 && body) {
          itax = cf.itAx[(int64_t)j + (int64_t)g.Ny * k];
          ldv<V>(cf.itAy, (I)ic + (I)g.Nx * (I)k, c1y);
        }
        if (ESC) {  // code 255: this cell's coefficients come from the arrays
#pragma unroll
          for (int q = 0; q < V; ++q) {
            if (((code >> (8 * q)) & 0xffu) != 0xffu) continue;
            const I nq = nk + (I)q;
#pragma unroll
            for (int c = 0; c < 3; ++c) ao[c][q] = *at(cf.a, (I)c * N + nq);
            if (w >= 1 && body) {
#pragma unroll
              for (int c = 0; c < 3; ++c) {
                bo[c][q] = *at(cf.b, (I)c * N + nq);
                if (COND || JM == 1) ie[c][q] = *at(cf.ieps, (I)c * N + nq);
                if (COND) sg[c][q] = *at(cf.sigma, (I)c * N + nq);
              }
            }
          }
        }
      } else {
        ldv<V>(cf.a, nk, ao[0]); ldv<V>(cf.a, N + nk, ao[1]); ldv<V>(cf.a, 2 * N + nk, ao[2]);
        if (w >= 1 && body) {
          ldv<V>(cf.b, nk, bo[0]); ldv<V>(cf.b, N + nk, bo[1]); ldv<V>(cf.b, 2 * N + nk, bo[2]);
          if (COND) {
            ldv<V>(cf.ieps, nk, ie[0]); ldv<V>(cf.ieps, N + nk, ie[1]); ldv<V>(cf.ieps, 2 * N + nk, ie[2]);
            ldv<V>(cf.sigma, nk, sg[0]); ldv<V>(cf.sigma, N + nk, sg[1]); ldv<V>(cf.sigma, 2 * N + nk, sg[2]);
          } else if (JM == 1 && qbeam >= 0) {
            ldv<V>(cf.ieps, 2 * N + nk, ie[2]);
          }
        }
      }
    }
    if (lact && kp2) {
      const I n2 = nk + 2 * sz;
      ldv<V>(Ein, n2, En2[0]); ldv<V>(Ein, N + n2, En2[1]); ldv<V>(Ein, 2 * N + n2, En2[2]);
    }
    S Tp[3][V];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < V; ++q) Tp[c][q] = S(0);
    if (top) {
      const I nt = tcol + sz * (I)k;
      ldv<V>(Ein, nt, Tp[0]); ldv<V>(Ein, N + nt, Tp[1]); ldv<V>(Ein, 2 * N + nt, Tp[2]);
    }
    S hxL[3] = {0, 0, 0}, hxR[3] = {0, 0, 0}, htL = 0, htR[3] = {0, 0, 0};
    S LH0 = 0, LH1 = 0, LA0 = 0, LA1 = 0, LE1 = 0;
    int Lcode = 0;
    if (hasL) {
      const I n = col - 1 + sz * (I)k;
      hxL[0] = *at(Ein, n); hxL[1] = *at(Ein, N + n); hxL[2] = *at(Ein, 2 * N + n);
      if (w >= 1 && body) {
        LH0 = *at(Hin, N + n); LH1 = *at(Hin, 2 * N + n);
        if (LUT) {
          Lcode = cf.code[n];
          LA0 = cf.iAy[(int64_t)(i0 - 1) + (int64_t)g.Nx * k];
          if (ESC && Lcode == 0xff) { LA0 = *at(cf.a, N + n); LA1 = *at(cf.a, 2 * N + n); }
        } else {
          LA0 = *at(cf.a, N + n); LA1 = *at(cf.a, 2 * N + n);
        }
        LE1 = kp1 ? *at(Ein, n + sz) : S(0);
      }
    }
    if (hasR) {
      const I n = col + V + sz * (I)k;
      hxR[0] = *at(Ein, n); hxR[1] = *at(Ein, N + n); hxR[2] = *at(Ein, 2 * N + n);
    }
    if (topL) htL = *at(Ein, tcol - 1 + sz * (I)k);
    if (topR) {
      const I n = tcol + V + sz * (I)k;
      htR[0] = *at(Ein, n); htR[1] = *at(Ein, N + n); htR[2] = *at(Ein, 2 * N + n);
    }

    // ---- stage masked E(k)
    const int zk = zface(pz, k, g.Nz), zk1 = kp1 ? zface(pz, k + 1, g.Nz) : -1;
    const bool mEx = zsel(zk, mXj), mEx1 = kp1 && zsel(zk1, mXj);
    bool mEy[V], mEy1[V];
#pragma unroll
    for (int q = 0; q < V; ++q) {
      mEy[q] = lact && zsel(zk, mYi[q]);
      mEy1[q] = lact && kp1 && zsel(zk1, mYi[q]);
    }
    if (wok) {
#pragma unroll
      for (int q = 0; q < V; ++q) {
        sE(0, w)[OFF + xi + q] = mEx ? Ek[0][q] : S(0);
        sE(1, w)[OFF + xi + q] = mEy[q] ? Ek[1][q] : S(0);
        sE(2, w)[OFF + xi + q] = mZi[q] ? Ek[2][q] : S(0);
      }
      if (lane == 0) {
        sE(0, w)[OFF - 1] = (hasL && mEx) ? hxL[0] : S(0);
        sE(1, w)[OFF - 1] = (hasL && zsel(zk, mYL)) ? hxL[1] : S(0);
        sE(2, w)[OFF - 1] = mZL ? hxL[2] : S(0);
      }
      if (lane == LL) {
        sE(0, w)[OFF + W] = (hasR && mEx) ? hxR[0] : S(0);
        sE(1, w)[OFF + W] = (hasR && zsel(zk, mYR)) ? hxR[1] : S(0);
        sE(2, w)[OFF + W] = mZR ? hxR[2] : S(0);
      }
    }
    if (w == 0) {
      const bool mEt = (jt < g.Ny) && zsel(zk, mXt);
#pragma unroll
      for (int q = 0; q < V; ++q) {
        sE(0, TYr + 1)[OFF + xi + q] = (top && mEt) ? Tp[0][q] : S(0);
        sE(1, TYr + 1)[OFF + xi + q] = (top && zsel(zk, mYt[q])) ? Tp[1][q] : S(0);
        sE(2, TYr + 1)[OFF + xi + q] = mZt[q] ? Tp[2][q] : S(0);
      }
      if (lane == 0) sE(0, TYr + 1)[OFF - 1] = (topL && mEt) ? htL : S(0);
      if (lane == LL) {
        sE(0, TYr + 1)[OFF + W] = (topR && mEt) ? htR[0] : S(0);
        sE(1, TYr + 1)[OFF + W] = (topR && zsel(zk, mYR)) ? htR[1] : S(0);
        sE(2, TYr + 1)[OFF + W] = mZtR ? htR[2] : S(0);
      }
    }
    const S Lz = m.Lz[k];
    __syncthreads();

    // ---- H_new(k) of this row (own rows always, the halo row on body planes)
    S Hn[3][V];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < V; ++q) Hn[c][q] = S(0);
    if (lact && (w >= 1 || body)) {
      if (LUT) {
#pragma unroll
        for (int q = 0; q < V; ++q) {
          if (ESC && ((code >> (8 * q)) & 0xffu) == 0xffu) continue;  // loaded above
          const S *t = sMat + NM * (int)((code >> (8 * q)) & 0xffu);
          ao[0][q] = (S)(iax * t[0]);
          ao[1][q] = (S)(c0y[q] * t[1]);
          ao[2][q] = (S)(iAz[q] * t[2]);
          if (w >= 1 && body) {
            bo[0][q] = mEx ? (S)(t[3] * itax) : S(0);
            bo[1][q] = mEy[q] ? (S)(t[4] * c1y[q]) : S(0);
            bo[2][q] = mZi[q] ? (S)(t[5] * itAz[q]) : S(0);
#pragma unroll
            for (int c = 0; c < 3; ++c) { ie[c][q] = t[3 + c]; sg[c][q] = t[6 + c]; }
          }
        }
      }
      if (hz) {
        const int zh = zface(hz, k, g.Nz);
#pragma unroll
        for (int q = 0; q < V; ++q) {
          if (!zsel(zh, ypart(hz, j, g.Ny))) ao[0][q] = S(0);
          if (!zsel(zh, xpart(hz, ic + q, g.Nx))) ao[1][q] = S(0);
          if (!mz(hz, ic + q, j, g)) ao[2][q] = S(0);
        }
      }
      S EX[V + 1], EY[V + 1], EZ[V + 1];
#pragma unroll
      for (int q = 0; q < V + 1; ++q) {
        EX[q] = sE(0, w)[OFF + xi + q]; EY[q] = sE(1, w)[OFF + xi + q]; EZ[q] = sE(2, w)[OFF + xi + q];
      }
#pragma unroll
      for (int q = 0; q < V; ++q) {
        const S ex1 = mEx1 ? Ek1[0][q] : S(0), ey1 = mEy1[q] ? Ek1[1][q] : S(0);
        Hn[0][q] = h_row<S, S, DESC>(Ho[0][q], ao[0][q], dt, Ly, Lz, EY[q], ey1, EZ[q],
                                     sE(2, w + 1)[OFF + xi + q], +1, -1);
        Hn[1][q] = h_row<S, S, DESC>(Ho[1][q], ao[1][q], dt, Lx[q], Lz, EX[q], ex1, EZ[q],
                                     EZ[q + 1], -1, +1);
        Hn[2][q] = h_row<S, S, DESC>(Ho[2][q], ao[2][q], dt, Lx[q], Ly, EX[q],
                                     sE(0, w + 1)[OFF + xi + q], EY[q], EY[q + 1], +1, -1);
      }
    }
    {
      S hm[3][V];
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int q = 0; q < V; ++q) hm[c][q] = Hn[c][q];
      if (hz && lact) {
        const int zh = zface(hz, k, g.Nz);
#pragma unroll
        for (int q = 0; q < V; ++q) {
          if (!zsel(zh, ypart(hz, j, g.Ny))) hm[0][q] = S(0);
          if (!zsel(zh, xpart(hz, ic + q, g.Nx))) hm[1][q] = S(0);
          if (!mz(hz, ic + q, j, g)) hm[2][q] = S(0);
        }
      }
      if (wok) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int q = 0; q < V; ++q) sH(c, w)[OFF + xi + q] = hm[c][q];
      }
      if (wok && lane == 0) {
        S hy = 0, hzv = 0;
        if (hasL && w >= 1 && body) {
          S lay, laz;
          if (LUT && !(ESC && Lcode == 0xff)) {
            const S *t = sMat + NM * Lcode;
            lay = (S)(LA0 * t[1]);
            laz = (S)(iAzL * t[2]);
          } else {
            lay = LA0; laz = LA1;
          }
          if (hz) {
            const int zh = zface(hz, k, g.Nz);
            if (!zsel(zh, xpart(hz, i0 - 1, g.Nx))) lay = S(0);
            if (!mz(hz, i0 - 1, j, g)) laz = S(0);
          }
          const S exl = sE(0, w)[OFF - 1], eyl = sE(1, w)[OFF - 1], ezl = sE(2, w)[OFF - 1];
          const S ex1 = mEx1 ? LE1 : S(0);
          hy = h_row<S, S, DESC>(LH0, lay, dt, LxL, Lz, exl, ex1, ezl, sE(2, w)[OFF], -1, +1);
          hzv = h_row<S, S, DESC>(LH1, laz, dt, LxL, Ly, exl, sE(0, w + 1)[OFF - 1], eyl,
                                  sE(1, w)[OFF], +1, -1);
          if (hz) {
            const int zh = zface(hz, k, g.Nz);
            if (!zsel(zh, xpart(hz, i0 - 1, g.Nx))) hy = S(0);
            if (!mz(hz, i0 - 1, j, g)) hzv = S(0);
          }
        }
        sH(0, w)[OFF - 1] = S(0);
        sH(1, w)[OFF - 1] = hy;
        sH(2, w)[OFF - 1] = hzv;
      }
    }
    __syncthreads();

    // ---- E_new(k) for own rows
    if (own && body) {
      const S tLz = m.tLz[k];
      S HX[V + 1], HY[V + 1], HZ[V + 1];  // [0] = -x neighbour
#pragma unroll
      for (int q = 0; q < V + 1; ++q) {
        HX[q] = sH(0, w)[OFF + xi + q - 1];
        HY[q] = sH(1, w)[OFF + xi + q - 1];
        HZ[q] = sH(2, w)[OFF + xi + q - 1];
      }
      S En[3][V];
#pragma unroll
      for (int q = 0; q < V; ++q) {
        S d0 = e_sum<S>(bo[0][q], tLy, tLz, Hpy[q], HY[q + 1], sH(2, w - 1)[OFF + xi + q],
                        HZ[q + 1], +1, -1);
        S d1 = e_sum<S>(bo[1][q], tLx[q], tLz, Hpx[q], HX[q + 1], HZ[q], HZ[q + 1], -1, +1);
        S d2 = e_sum<S>(bo[2][q], tLx[q], tLy, sH(0, w - 1)[OFF + xi + q], HX[q + 1], HY[q],
                        HY[q + 1], +1, -1);
        const bool online = (q == qbeam);
        S jline = 0;
        if (online && (JM == 1 || COND)) jline = beam_row ? beam_row[k] : *at(J, 2 * N + nk + q);
        if (COND) {
          d0 = d0 - ie[0][q] * (S)(sg[0][q] * Ek[0][q]);
          d1 = d1 - ie[1][q] * (S)(sg[1][q] * Ek[1][q]);
          d2 = d2 - ie[2][q] * (online ? jline : (S)(sg[2][q] * Ek[2][q]));
        } else if (JM == 1 && online) {
          d2 = d2 - ie[2][q] * jline;
        }
        En[0][q] = (S)(Ek[0][q] + dt * d0);
        En[1][q] = (S)(Ek[1][q] + dt * d1);
        En[2][q] = (S)(Ek[2][q] + dt * d2);
        if (!COND && online && beam_row) *atw(J, 2 * N + nk + q) = jline;
        if (q >= capq0 && q < capq1)
          cap[capbase + (int64_t)(q + ic - x.cap_ix0) * x.cap_ny * x.cap_samples * g.Nz + k] =
              En[2][q];
      }
      stv_nt<V>(Eout, nk, En[0]); stv_nt<V>(Eout, N + nk, En[1]);
      stv_nt<V>(Eout, 2 * N + nk, En[2]);
      stv_nt<V>(Hout, nk, Hn[0]); stv_nt<V>(Hout, N + nk, Hn[1]);
      stv_nt<V>(Hout, 2 * N + nk, Hn[2]);
      if (COND && !x.skip_j) {
        S Jn[3][V];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int q = 0; q < V; ++q) Jn[c][q] = (S)(sg[c][q] * En[c][q]);
        stv_nt<V>(J, nk, Jn[0]); stv_nt<V>(J, N + nk, Jn[1]); stv_nt<V>(J, 2 * N + nk, Jn[2]);
      } else if (COND && qbeam >= 0 && !beam_row) {
        // the line's J is read back by the next step when no beam table feeds it
        *atw(J, 2 * N + nk + qbeam) = (S)(sg[2][qbeam] * En[2][qbeam]);
      }
#pragma unroll
      for (int q = 0; q < V; ++q) { Hpx[q] = HX[q + 1]; Hpy[q] = HY[q + 1]; }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < V; ++q) { Ek[c][q] = Ek1[c][q]; Ek1[c][q] = En2[c][q]; }
    if (w >= 1 && wok && !body) {  // warm-up plane: H_new(k0-1) seeds the z-1 terms
#pragma unroll
      for (int q = 0; q < V; ++q) {
        Hpx[q] = sH(0, w)[OFF + xi + q];
        Hpy[q] = sH(1, w)[OFF + xi + q];
      }
    }
  }
}

// FR: TY only bounds the workgroup (64 * (TY + 1) threads); x.rows is the tile
// ESC (with LUT): material code 255 marks cells whose coefficients are read
// from the arrays a, b, ieps, sigma (per-cell parameters, e.g. a block of
// cells being optimised); the table then has at most 255 rows
template <typename S, int JM, bool COND, bool LUT, bool DESC, int TY, bool MK, int V, bool FR,
          bool ESC = false>
__global__ void __launch_bounds__(64 * (TY + 1))
k_fused4(GridDev g, MetricDev<S> m, const S *__restrict__ Ein, const S *__restrict__ Hin,
         S *__restrict__ Eout, S *__restrict__ Hout, S *__restrict__ J, Coefs<S> cf, S dt,
         FusedAux x, const S *__restrict__ beam_row, S *__restrict__ cap) {
  const int tile = xcd_tile(blockIdx.x, x.ntx * x.nty * x.ntz);
  fused4_tile<S, JM, COND, LUT, DESC, TY, MK, V, FR, ESC>(g, m, Ein, Hin, Eout, Hout, J, cf,
                                                          dt, x, beam_row, cap, tile);
}

// The same step for grids whose 3N arrays pass 4 GiB: 64-bit offsets (two
// VGPRs per address), row segments, no escape code.
template <typename S, int JM, bool COND, bool LUT, bool DESC, int TY, bool MK, int V>
__global__ void __launch_bounds__(64 * (TY + 1))
k_fused4_wide(GridDev g, MetricDev<S> m, const S *__restrict__ Ein, const S *__restrict__ Hin,
              S *__restrict__ Eout, S *__restrict__ Hout, S *__restrict__ J, Coefs<S> cf, S dt,
              FusedAux x, const S *__restrict__ beam_row, S *__restrict__ cap) {
  const int tile = xcd_tile(blockIdx.x, x.ntx * x.nty * x.ntz);
  fused4_tile<S, JM, COND, LUT, DESC, TY, MK, V, false, false, true>(
      g, m, Ein, Hin, Eout, Hout, J, cf, dt, x, beam_row, cap, tile);
}

// One grid of a batch: everything a step of it needs except the step number.
template <typename S>
struct BatchDesc {
  GridDev g;
  MetricDev<S> m;
  S *E, *H, *E2, *H2, *J;      // even steps read E/H and write E2/H2, odd steps the reverse
  Coefs<S> cf;
  S dt;
  FusedAux x;                  // cap_sample is set per step
  const S *beam_table;         // row of step s: beam_table + (beam_row0 + s) * Nz, or null
  int64_t beam_row0;
  S *cap;
  int64_t cap_sample0;         // capture sample written by step s: cap_sample0 + s
  int nsteps;                  // the grid takes part in steps 0 .. nsteps-1
  int tile0;                   // its first tile in the batch's tile numbering
  // optional region trajectory (reverse mode): after step s, slot traj_slot0+s+1
  S *tE, *tH;                  // [slots][3 * box cells], or null
  int tlo[3], thi[3];
  int64_t traj_slot0;
  int64_t thr0;                // first thread of this grid in the batched gather
};

// One step of every grid of a batch in one launch. The step number comes from
// device memory (k_batch_next advances it), so a captured graph of these
// launches replays the right steps.
// FR: whole-row tiles, each grid with its own x.rows / x.qrow within the
// workgroup of 64 * (TY + 1) threads
template <typename S, int JM, bool COND, bool LUT, bool DESC, int TY, bool MK, int V, bool FR>
__global__ void __launch_bounds__(64 * (TY + 1))
k_fused4_batch(const BatchDesc<S> *__restrict__ d, int ninst, int ntiles,
               const int *__restrict__ step_ptr) {
  const int step = *step_ptr;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  int lo = 0, hi = ninst - 1;  // last descriptor with tile0 <= tile
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].tile0 <= tile) lo = mid;
    else hi = mid - 1;
  }
  const BatchDesc<S> &b = d[lo];
  if (step >= b.nsteps) return;  // whole workgroup, before any barrier
  const bool odd = step & 1;
  FusedAux x = b.x;
  x.cap_sample = b.cap_sample0 + step;
  x.skip_j = (step + 1 < b.nsteps) ? 1 : 0;
  const S *row = b.beam_table ? b.beam_table + (b.beam_row0 + step) * b.g.Nz : nullptr;
  fused4_tile<S, JM, COND, LUT, DESC, TY, MK, V, FR, false>(
      b.g, b.m, odd ? b.E2 : b.E, odd ? b.H2 : b.H, odd ? b.E : b.E2, odd ? b.H : b.H2, b.J,
      b.cf, b.dt, x, row, b.cap, tile - b.tile0);
}

static __global__ void k_batch_next(int *step) { step[0] = step[0] + 1; }
static __global__ void k_batch_set(int *step, int v) { step[0] = v; }

}  // namespace wf
