// Fused single-pass leapfrog step (fast path).
//
// One workgroup owns a TX x TY column tile and a chunk of KZ z-planes and
// marches up in z. Per plane k it stages E_old(k) of the tile plus a one-cell
// halo ring in LDS, computes H_new(k) for its own cells and for the low halo
// row/column (recomputed instead of exchanged), then E_new(k) from H_new(k),
// H_new(k-1) (kept in registers) and the LDS neighbours. E and H are read once
// and written once per step (ping-pong buffers, so halo reads never see a
// neighbour's new values). Per row the arithmetic is exactly the general
// path's: m = fl(a*L), fl(fl(dt*m)*E) summed from 0 in CSR column order,
// H - sum; fl(b*tL)*H summed ascending, E + fl(dt*(S - fl(ieps*J))).
//
// Neighbours outside the box are dropped. The reference's flat index also
// couples i=Nx-1 to the next row and j=Ny-1 to the next plane ("wrap"
// terms); this kernel is only used after the host has proven every such term
// contributes exactly zero (zero weight, or a column frozen at zero).
#pragma once
#include "fit_common.h"

namespace wf {

struct FusedAux {
  int ixs, iys;
  int cap_ix0, cap_iy0, cap_nx, cap_ny;
  int64_t cap_samples, cap_sample;
  int KZ, ntx, nty, ntz;
};

// z-face override of the face mask for components x and y at plane k:
// -1 when plane k is not on a z face.
__device__ __forceinline__ int zface(int zbits, int k, int Nz) {
  if (k == Nz - 1) return ((zbits >> 5) & 1) ? 0 : 1;
  if (k == 0) return ((zbits >> 2) & 1) ? 0 : 1;
  return -1;
}
// In-plane part of the face mask (faces of the listed axes only).
__device__ __forceinline__ int xpart(int zbits, int i, int Nx) {
  if (i == Nx - 1) return ((zbits >> 3) & 1) ? 0 : 1;
  if (i == 0) return ((zbits >> 0) & 1) ? 0 : 1;
  return 1;
}
__device__ __forceinline__ int ypart(int zbits, int j, int Ny) {
  if (j == Ny - 1) return ((zbits >> 4) & 1) ? 0 : 1;
  if (j == 0) return ((zbits >> 1) & 1) ? 0 : 1;
  return 1;
}
// full masks: x needs (j,k), y needs (i,k), z needs (i,j)
__device__ __forceinline__ bool mx(int zb, int j, int k, const GridDev &g) {
  const int z = zface(zb, k, g.Nz);
  return (z >= 0) ? z : ypart(zb, j, g.Ny);
}
__device__ __forceinline__ bool my(int zb, int i, int k, const GridDev &g) {
  const int z = zface(zb, k, g.Nz);
  return (z >= 0) ? z : xpart(zb, i, g.Nx);
}
__device__ __forceinline__ bool mz(int zb, int i, int j, const GridDev &g) {
  if (j == g.Ny - 1) return !((zb >> 4) & 1);
  if (j == 0) return !((zb >> 1) & 1);
  return xpart(zb, i, g.Nx);
}

// One M1 row: H - sum(fl(fl(dt*fl(a*L_p))*E_p)) over the 4 stencil terms in
// ascending column order (t0..t3), or descending when desc. Terms 0,1 share
// edge length L0, terms 2,3 share L1. Returns the new H.
template <typename S, typename Cm>
__device__ __forceinline__ S h_row(S h, S ar, Cm dt, S L0, S L1, S e0, S e1, S e2, S e3,
                                   bool u0, bool u1, bool u2, bool u3, int s0, int s1, int s2,
                                   int s3, bool desc) {
  const Cm w0 = dt * (Cm)(S)(ar * L0);
  const Cm w1 = dt * (Cm)(S)(ar * L1);
  const Cm p0 = w0 * (Cm)e0, p1 = w0 * (Cm)e1, p2 = w1 * (Cm)e2, p3 = w1 * (Cm)e3;
  Cm acc = Cm(0);
  if (!desc) {
    if (u0) acc = s0 > 0 ? acc + p0 : acc - p0;
    if (u1) acc = s1 > 0 ? acc + p1 : acc - p1;
    if (u2) acc = s2 > 0 ? acc + p2 : acc - p2;
    if (u3) acc = s3 > 0 ? acc + p3 : acc - p3;
  } else {
    if (u3) acc = s3 > 0 ? acc + p3 : acc - p3;
    if (u2) acc = s2 > 0 ? acc + p2 : acc - p2;
    if (u1) acc = s1 > 0 ? acc + p1 : acc - p1;
    if (u0) acc = s0 > 0 ? acc + p0 : acc - p0;
  }
  return (S)((Cm)h - acc);
}

// One M2 row sum (storage precision, ascending order).
template <typename S>
__device__ __forceinline__ S e_sum(S br, S tL0, S tL1, S h0, S h1, S h2, S h3, bool u0,
                                   bool u1, bool u2, bool u3, int s0, int s1, int s2, int s3) {
  const S m0 = br * tL0, m1 = br * tL1;
  const S p0 = m0 * h0, p1 = m0 * h1, p2 = m1 * h2, p3 = m1 * h3;
  S acc = S(0);
  if (u0) acc = s0 > 0 ? acc + p0 : acc - p0;
  if (u1) acc = s1 > 0 ? acc + p1 : acc - p1;
  if (u2) acc = s2 > 0 ? acc + p2 : acc - p2;
  if (u3) acc = s3 > 0 ? acc + p3 : acc - p3;
  return acc;
}

template <typename S, typename Cm, int JM, bool COND, int TX, int TY>
__global__ void __launch_bounds__(TX *TY)
k_fused(GridDev g, MetricDev<S> m, const S *__restrict__ Ein, const S *__restrict__ Hin,
        S *__restrict__ Eout, S *__restrict__ Hout, S *__restrict__ J, const S *__restrict__ a,
        const S *__restrict__ b, const S *__restrict__ ieps, const S *__restrict__ sigma, Cm dt,
        FusedAux x, const S *__restrict__ beam_row, S *__restrict__ cap) {
  constexpr int NT = TX * TY;
  constexpr int NRING = 2 * (TX + 2) + 2 * TY;
  __shared__ S sE[3][TY + 2][TX + 2];
  __shared__ S sH[3][TY + 1][TX + 1];

  // XCD-aware tile order: consecutive dispatch slots of one XCD take
  // consecutive tiles, so x/y halo neighbours share that XCD's L2.
  const int nb = x.ntx * x.nty * x.ntz;
  const int blin = blockIdx.x;
  const int xcd = blin & 7, slot = blin >> 3;
  const int q = nb >> 3, r = nb & 7;
  const int tile = xcd * q + (xcd < r ? xcd : r) + slot;
  const int tz = tile / (x.ntx * x.nty);
  const int rem = tile - tz * x.ntx * x.nty;
  const int tyb = rem / x.ntx, txb = rem - tyb * x.ntx;
  const int i0 = txb * TX, j0 = tyb * TY, k0 = tz * x.KZ;
  const int kend = min(k0 + x.KZ, g.Nz);

  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;
  const int i = i0 + tx, j = j0 + ty;
  const bool act = (i < g.Nx) && (j < g.Ny);
  const int64_t N = g.N, sy = g.sy, sz = g.sz;
  const int64_t col = (int64_t)i + sy * j;  // in-plane offset of the own column

  // halo ring cell of this thread (threads < NRING): position in sE
  int rjj = -1, rii = -1;
  if (tid < NRING) {
    if (tid < TX + 2) { rjj = 0; rii = tid; }
    else if (tid < 2 * (TX + 2)) { rjj = TY + 1; rii = tid - (TX + 2); }
    else if (tid < 2 * (TX + 2) + TY) { rjj = tid - 2 * (TX + 2) + 1; rii = 0; }
    else { rjj = tid - 2 * (TX + 2) - TY + 1; rii = TX + 1; }
  }
  const int ri = i0 - 1 + rii, rj = j0 - 1 + rjj;
  const bool ring_in = (tid < NRING) && ri >= 0 && ri < g.Nx && rj >= 0 && rj < g.Ny;
  const int64_t rcol = ring_in ? (int64_t)ri + sy * rj : 0;

  // halo H cells: bottom row (ty==0) and left column (tx==0)
  const int jb = j0 - 1, ib = i0 - 1;
  const bool hb = (ty == 0) && jb >= 0 && i < g.Nx;
  const bool hl = (tx == 0) && ib >= 0 && j < g.Ny;
  const int64_t bcol = (int64_t)i + sy * jb, lcol = (int64_t)ib + sy * j;

  // per-column constants
  const int pz = g.pecz, hz = g.pmcz;
  const bool desc = g.m1desc != 0;
  S Lx = 0, Ly = 0, Lz = 0, tLx = 0, tLy = 0;
  if (i < g.Nx) { Lx = m.Lx[i]; tLx = m.tLx[i]; }
  if (j < g.Ny) { Ly = m.Ly[j]; tLy = m.tLy[j]; }

  const S Lx_l = (ib >= 0) ? m.Lx[ib] : S(0);   // left halo column
  const S Ly_b = (jb >= 0) ? m.Ly[jb] : S(0);   // bottom halo row
  const bool in_x1 = (i + 1 < g.Nx), in_y1 = (j + 1 < g.Ny);
  const bool in_xm = (i >= 1), in_ym = (j >= 1);
  const bool online = act && (i == x.ixs) && (j == x.iys);
  const int li = i - x.cap_ix0, lj = j - x.cap_iy0;
  const bool incap = cap && act && li >= 0 && li < x.cap_nx && lj >= 0 && lj < x.cap_ny &&
                     x.cap_sample < x.cap_samples;
  const int64_t capbase = incap ? (((int64_t)li * x.cap_ny + lj) * x.cap_samples +
                                   x.cap_sample) * g.Nz : 0;

  const int kstart = (k0 > 0) ? k0 - 1 : k0;
  S Ek[3] = {0, 0, 0}, Ek1[3] = {0, 0, 0};
  S ring[3] = {0, 0, 0};
  if (act) {
    const int64_t n = col + sz * kstart;
    Ek[0] = Ein[n]; Ek[1] = Ein[N + n]; Ek[2] = Ein[2 * N + n];
    if (kstart + 1 < g.Nz) {
      Ek1[0] = Ein[n + sz]; Ek1[1] = Ein[N + n + sz]; Ek1[2] = Ein[2 * N + n + sz];
    }
  }
  if (ring_in) {
    const int64_t n = rcol + sz * kstart;
    ring[0] = Ein[n]; ring[1] = Ein[N + n]; ring[2] = Ein[2 * N + n];
  }
  S Hpx = 0, Hpy = 0;  // H_new at plane k-1, own column

  for (int k = kstart; k < kend; ++k) {
    const bool body = (k >= k0);  // k0-1 is the warm-up plane (H only)
    const int64_t nk = col + sz * k;
    // stage E_old(k)
    if (act) {
      sE[0][ty + 1][tx + 1] = Ek[0];
      sE[1][ty + 1][tx + 1] = Ek[1];
      sE[2][ty + 1][tx + 1] = Ek[2];
    }
    if (tid < NRING) {
      sE[0][rjj][rii] = ring[0];
      sE[1][rjj][rii] = ring[1];
      sE[2][rjj][rii] = ring[2];
    }
    // loads for this plane (own) and prefetch of the next plane
    S Ho[3] = {0, 0, 0}, ao[3] = {0, 0, 0}, bo[3] = {0, 0, 0}, En2[3] = {0, 0, 0};
    S iep[3] = {0, 0, 0}, sg[3] = {0, 0, 0};
    const bool kp1 = (k + 1 < g.Nz), kp2 = (k + 2 < g.Nz);
    if (act) {
      Ho[0] = Hin[nk]; Ho[1] = Hin[N + nk]; Ho[2] = Hin[2 * N + nk];
      ao[0] = a[nk]; ao[1] = a[N + nk]; ao[2] = a[2 * N + nk];
      if (body) {
        bo[0] = b[nk]; bo[1] = b[N + nk]; bo[2] = b[2 * N + nk];
        if (COND) {
          iep[0] = ieps[nk]; iep[1] = ieps[N + nk]; iep[2] = ieps[2 * N + nk];
          sg[0] = sigma[nk]; sg[1] = sigma[N + nk]; sg[2] = sigma[2 * N + nk];
        } else if (JM == 1 && online) {
          iep[2] = ieps[2 * N + nk];
        }
      }
      if (kp2) {
        const int64_t n2 = nk + 2 * sz;
        En2[0] = Ein[n2]; En2[1] = Ein[N + n2]; En2[2] = Ein[2 * N + n2];
      }
    }
    S ringn[3] = {0, 0, 0};
    if (ring_in && kp1) {
      const int64_t n1 = rcol + sz * (k + 1);
      ringn[0] = Ein[n1]; ringn[1] = Ein[N + n1]; ringn[2] = Ein[2 * N + n1];
    }
    // halo-H inputs
    S bHx = 0, bHz = 0, bax = 0, baz = 0, bEy1 = 0;
    S lHy = 0, lHz = 0, lay = 0, laz = 0, lEx1 = 0;
    if (body && hb) {
      const int64_t n = bcol + sz * k;
      bHx = Hin[n]; bHz = Hin[2 * N + n]; bax = a[n]; baz = a[2 * N + n];
      if (kp1) bEy1 = Ein[N + n + sz];
    }
    if (body && hl) {
      const int64_t n = lcol + sz * k;
      lHy = Hin[N + n]; lHz = Hin[2 * N + n]; lay = a[N + n]; laz = a[2 * N + n];
      if (kp1) lEx1 = Ein[n + sz];
    }
    if (k < g.Nz) Lz = m.Lz[k];
    __syncthreads();

    // ---- H_new(k), own column
    S Hn[3] = {0, 0, 0};
    if (act) {
      const int zk = zface(pz, k, g.Nz), zk1 = kp1 ? zface(pz, k + 1, g.Nz) : -1;
      const bool myk = (zk >= 0) ? zk : xpart(pz, i, g.Nx);
      const bool myk1 = (zk1 >= 0) ? zk1 : xpart(pz, i, g.Nx);
      const bool mxk = (zk >= 0) ? zk : ypart(pz, j, g.Ny);
      const bool mxk1 = (zk1 >= 0) ? zk1 : ypart(pz, j, g.Ny);
      const bool myk_i1 = in_x1 && ((zk >= 0) ? zk : xpart(pz, i + 1, g.Nx));
      const bool mxk_j1 = in_y1 && ((zk >= 0) ? zk : ypart(pz, j + 1, g.Ny));
      const S exk = sE[0][ty + 1][tx + 1], eyk = sE[1][ty + 1][tx + 1], ezk = sE[2][ty + 1][tx + 1];
      // H_x: Ey(k)+, Ey(k+1)-, Ez(k)-, Ez(j+1)+
      Hn[0] = h_row<S, Cm>(Ho[0], ao[0], dt, Ly, Lz, eyk, Ek1[1], ezk, sE[2][ty + 2][tx + 1],
                           myk, kp1 && myk1, mz(pz, i, j, g), in_y1 && mz(pz, i, j + 1, g),
                           +1, -1, -1, +1, desc);
      // H_y: Ex(k)-, Ex(k+1)+, Ez(k)+, Ez(i+1)-
      Hn[1] = h_row<S, Cm>(Ho[1], ao[1], dt, Lx, Lz, exk, Ek1[0], ezk, sE[2][ty + 1][tx + 2],
                           mxk, kp1 && mxk1, mz(pz, i, j, g), in_x1 && mz(pz, i + 1, j, g),
                           -1, +1, +1, -1, desc);
      // H_z: Ex(k)+, Ex(j+1)-, Ey(k)-, Ey(i+1)+
      Hn[2] = h_row<S, Cm>(Ho[2], ao[2], dt, Lx, Ly, exk, sE[0][ty + 2][tx + 1], eyk,
                           sE[1][ty + 1][tx + 2], mxk, mxk_j1, myk, myk_i1, +1, -1, -1, +1,
                           desc);
      sH[0][ty + 1][tx + 1] = Hn[0];
      sH[1][ty + 1][tx + 1] = Hn[1];
      sH[2][ty + 1][tx + 1] = Hn[2];
    }
    // ---- H_new(k) on the low halo row (Hx, Hz) and column (Hy, Hz)
    if (body && hb) {
      const int zk = zface(pz, k, g.Nz), zk1 = kp1 ? zface(pz, k + 1, g.Nz) : -1;
      const bool myk = (zk >= 0) ? zk : xpart(pz, i, g.Nx);
      const bool myk1 = (zk1 >= 0) ? zk1 : xpart(pz, i, g.Nx);
      const bool mxk = (zk >= 0) ? zk : ypart(pz, jb, g.Ny);
      const bool mxk_j1 = (zk >= 0) ? zk : ypart(pz, j0, g.Ny);
      const bool myk_i1 = in_x1 && ((zk >= 0) ? zk : xpart(pz, i + 1, g.Nx));
      const S exb = sE[0][0][tx + 1], eyb = sE[1][0][tx + 1], ezb = sE[2][0][tx + 1];
      sH[0][0][tx + 1] = h_row<S, Cm>(bHx, bax, dt, Ly_b, Lz, eyb, bEy1, ezb, sE[2][1][tx + 1],
                                      myk, kp1 && myk1, mz(pz, i, jb, g), mz(pz, i, j0, g),
                                      +1, -1, -1, +1, desc);
      sH[2][0][tx + 1] = h_row<S, Cm>(bHz, baz, dt, Lx, Ly_b, exb, sE[0][1][tx + 1], eyb,
                                      sE[1][0][tx + 2], mxk, mxk_j1, myk, myk_i1, +1, -1, -1,
                                      +1, desc);
    }
    if (body && hl) {
      const int zk = zface(pz, k, g.Nz), zk1 = kp1 ? zface(pz, k + 1, g.Nz) : -1;
      const bool mxk = (zk >= 0) ? zk : ypart(pz, j, g.Ny);
      const bool mxk1 = (zk1 >= 0) ? zk1 : ypart(pz, j, g.Ny);
      const bool myk = (zk >= 0) ? zk : xpart(pz, ib, g.Nx);
      const bool myk_i1 = (zk >= 0) ? zk : xpart(pz, i0, g.Nx);
      const bool mxk_j1 = in_y1 && ((zk >= 0) ? zk : ypart(pz, j + 1, g.Ny));
      const S exl = sE[0][ty + 1][0], eyl = sE[1][ty + 1][0], ezl = sE[2][ty + 1][0];
      sH[1][ty + 1][0] = h_row<S, Cm>(lHy, lay, dt, Lx_l, Lz, exl, lEx1, ezl, sE[2][ty + 1][1],
                                      mxk, kp1 && mxk1, mz(pz, ib, j, g), mz(pz, i0, j, g),
                                      -1, +1, +1, -1, desc);
      sH[2][ty + 1][0] = h_row<S, Cm>(lHz, laz, dt, Lx_l, Ly, exl, sE[0][ty + 2][0], eyl,
                                      sE[1][ty + 1][1], mxk, mxk_j1, myk, myk_i1, +1, -1, -1,
                                      +1, desc);
    }
    __syncthreads();

    // ---- E_new(k), own column
    if (body && act) {
      const bool km = (k >= 1);
      // PMC column masks on H (all true when there is no PMC face)
      bool hxk = true, hxkm = true, hyk = true, hykm = true, hzk = true, hzjm = true,
           hzim = true, hxjm = true, hyim = true;
      if (hz) {
        const int zk = zface(hz, k, g.Nz), zkm = km ? zface(hz, k - 1, g.Nz) : -1;
        hxk = (zk >= 0) ? zk : ypart(hz, j, g.Ny);
        hxkm = (zkm >= 0) ? zkm : ypart(hz, j, g.Ny);
        hyk = (zk >= 0) ? zk : xpart(hz, i, g.Nx);
        hykm = (zkm >= 0) ? zkm : xpart(hz, i, g.Nx);
        hzk = mz(hz, i, j, g);
        hzjm = in_ym && mz(hz, i, j - 1, g);
        hzim = in_xm && mz(hz, i - 1, j, g);
        hxjm = in_ym && ((zk >= 0) ? zk : ypart(hz, j - 1, g.Ny));
        hyim = in_xm && ((zk >= 0) ? zk : xpart(hz, i - 1, g.Nx));
      }
      const S hxc = sH[0][ty + 1][tx + 1], hyc = sH[1][ty + 1][tx + 1], hzc = sH[2][ty + 1][tx + 1];
      const S tLz = m.tLz[k];
      // E_x: Hy(k-1)+, Hy(k)-, Hz(j-1)-, Hz(k)+
      const S sx = e_sum<S>(bo[0], tLy, tLz, Hpy, hyc, sH[2][ty][tx + 1], hzc, km && hykm, hyk,
                            in_ym && hzjm, hzk, +1, -1, -1, +1);
      // E_y: Hx(k-1)-, Hx(k)+, Hz(i-1)+, Hz(k)-
      const S sy_ = e_sum<S>(bo[1], tLx, tLz, Hpx, hxc, sH[2][ty + 1][tx], hzc, km && hxkm, hxk,
                             in_xm && hzim, hzk, -1, +1, +1, -1);
      // E_z: Hx(j-1)+, Hx(k)-, Hy(i-1)-, Hy(k)+
      const S sz_ = e_sum<S>(bo[2], tLx, tLy, sH[0][ty][tx + 1], hxc, sH[1][ty + 1][tx], hyc,
                             in_ym && hxjm, hxk, in_xm && hyim, hyk, +1, -1, -1, +1);
      S d[3] = {sx, sy_, sz_};
      S jline = 0;
      if (online && (JM == 1 || COND)) jline = beam_row ? beam_row[k] : J[2 * N + nk];
      if (COND) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const S jv = (c == 2 && online) ? jline : (S)(sg[c] * Ek[c]);
          d[c] = d[c] - iep[c] * jv;
        }
      } else if (JM == 1 && online) {
        d[2] = d[2] - iep[2] * jline;
      }
      S En[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) En[c] = (S)((Cm)Ek[c] + dt * (Cm)d[c]);
      Eout[nk] = En[0]; Eout[N + nk] = En[1]; Eout[2 * N + nk] = En[2];
      Hout[nk] = Hn[0]; Hout[N + nk] = Hn[1]; Hout[2 * N + nk] = Hn[2];
      if (COND) {
        if (J) { J[nk] = sg[0] * En[0]; J[N + nk] = sg[1] * En[1]; J[2 * N + nk] = sg[2] * En[2]; }
      } else if (online && beam_row) {
        J[2 * N + nk] = jline;
      }
      if (incap) cap[capbase + k] = En[2];
    }
    Hpx = Hn[0];
    Hpy = Hn[1];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      Ek[c] = Ek1[c];
      Ek1[c] = En2[c];
      ring[c] = ringn[c];
    }
    __syncthreads();
  }
}

}  // namespace wf
