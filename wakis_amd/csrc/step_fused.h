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
  // k_fused4, conductive steps: 1 = store J = sigma E only on the beam line
  // without a beam table (the next step of the same call recomputes sigma E
  // from E; J is stored in full by the call's last step)
  int skip_j;
  // k_fused4 with full-row tiles: own rows per tile and lanes per row (Nx / V)
  int rows, qrow;
  // k_fused4: first tile along x, y, z of the launched tile range (ntx, nty,
  // ntz tiles from there); 0 for the whole grid
  int tx0, ty0, tz0;
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

// One M1 row without masks: H - sum fl(fl(dt*fl(a*L_p))*E_p) over the four
// stencil terms in CSR column order (ascending, or descending when DESC).
// Terms 0,1 share edge length L0, terms 2,3 share L1. Masked or absent
// columns come in as E_p = 0 (their product is an exact zero).
template <typename S, typename Cm, bool DESC>
__device__ __forceinline__ S h_row(S h, S ar, Cm dt, S L0, S L1, S e0, S e1, S e2, S e3,
                                   int s0, int s2) {
  // s0: sign of term 0 (term 1 has the opposite sign); s2 likewise for terms 2,3
  const Cm w0 = dt * (Cm)(S)(ar * L0);
  const Cm w1 = dt * (Cm)(S)(ar * L1);
  const Cm p0 = w0 * (Cm)e0, p1 = w0 * (Cm)e1, p2 = w1 * (Cm)e2, p3 = w1 * (Cm)e3;
  Cm acc = Cm(0);
  if (!DESC) {
    acc = s0 > 0 ? acc + p0 : acc - p0;
    acc = s0 > 0 ? acc - p1 : acc + p1;
    acc = s2 > 0 ? acc + p2 : acc - p2;
    acc = s2 > 0 ? acc - p3 : acc + p3;
  } else {
    acc = s2 > 0 ? acc - p3 : acc + p3;
    acc = s2 > 0 ? acc + p2 : acc - p2;
    acc = s0 > 0 ? acc - p1 : acc + p1;
    acc = s0 > 0 ? acc + p0 : acc - p0;
  }
  return (S)((Cm)h - acc);
}

// One M2 row sum (storage precision, ascending order), masked columns as 0.
template <typename S>
__device__ __forceinline__ S e_sum(S br, S tL0, S tL1, S h0, S h1, S h2, S h3, int s0, int s2) {
  const S m0 = br * tL0, m1 = br * tL1;
  const S p0 = m0 * h0, p1 = m0 * h1, p2 = m1 * h2, p3 = m1 * h3;
  S acc = S(0);
  acc = s0 > 0 ? acc + p0 : acc - p0;
  acc = s0 > 0 ? acc - p1 : acc + p1;
  acc = s2 > 0 ? acc + p2 : acc - p2;
  acc = s2 > 0 ? acc - p3 : acc + p3;
  return acc;
}

// Coefficient source: either the full 3N arrays a/b (and ieps/sigma), or a
// per-cell uint8 material code with a table of per-component material values
// (imu_xyz, ieps_xyz, sigma_xyz) staged in LDS plus 2-D metric tables
// iA_c/itA_c over the two transverse axes of component c. The kernel then
// recomputes a = fl(iA*imu) and b = fl(ieps*itA) exactly as the host fold.
template <typename S>
struct Coefs {
  const S *a, *b, *ieps, *sigma;
  const uint8_t *code;
  const S *mat;   // [nmat][9]
  int nmat;
  const S *iAx, *iAy, *iAz, *itAx, *itAy, *itAz;  // (j,k), (i,k), (i,j) tables
};

// Global access with a 32-bit byte offset (global_load ... v_off, s[base]) when
// the whole array fits in 4 GiB, otherwise 64-bit indexing.
template <typename S, typename I>
__device__ __forceinline__ S ldg(const S *__restrict__ p, I idx) {
  if constexpr (sizeof(I) == 4)
    return *reinterpret_cast<const S *>(reinterpret_cast<const char *>(p) +
                                        (uint32_t)(idx * (uint32_t)sizeof(S)));
  else
    return p[idx];
}
template <typename S, typename I>
__device__ __forceinline__ void stg(S *__restrict__ p, I idx, S v) {
  if constexpr (sizeof(I) == 4)
    *reinterpret_cast<S *>(reinterpret_cast<char *>(p) + (uint32_t)(idx * (uint32_t)sizeof(S))) = v;
  else
    p[idx] = v;
}

// mask value of a component-x/y face rule at plane k given the in-plane part
__device__ __forceinline__ bool zsel(int zk, bool inplane) {
  return (zk >= 0) ? (zk != 0) : inplane;
}

template <typename S, typename Cm, typename I, int JM, bool COND, bool LUT, bool DESC, int TX,
          int TY>
__global__ void __launch_bounds__(TX *TY)
k_fused(GridDev g, MetricDev<S> m, const S *__restrict__ Ein, const S *__restrict__ Hin,
        S *__restrict__ Eout, S *__restrict__ Hout, S *__restrict__ J, Coefs<S> cf, Cm dt,
        FusedAux x, const S *__restrict__ beam_row, S *__restrict__ cap) {
  constexpr int NT = TX * TY;
  constexpr int NRING = 2 * (TX + 2) + 2 * TY;
  constexpr int NM = 9;
  // E staged with PEC column masks applied; H staged with PMC column masks
  __shared__ S sE[1][3][TY + 2][TX + 2];
  __shared__ S sH[1][3][TY + 1][TX + 1];
  extern __shared__ unsigned char s_dyn[];
  S *sMat = reinterpret_cast<S *>(s_dyn);
  if (LUT) {
    for (int t = threadIdx.x; t < cf.nmat * NM; t += NT) sMat[t] = cf.mat[t];
  }

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
  const I N = (I)g.N, sy = (I)g.sy, sz = (I)g.sz;
  const I col = act ? (I)i + sy * (I)j : (I)0;
  const int pz = g.pecz, hz = g.pmcz;

  // halo ring cell of this thread (threads < NRING): position in sE
  int rjj = 0, rii = 0;
  if (tid < NRING) {
    if (tid < TX + 2) { rjj = 0; rii = tid; }
    else if (tid < 2 * (TX + 2)) { rjj = TY + 1; rii = tid - (TX + 2); }
    else if (tid < 2 * (TX + 2) + TY) { rjj = tid - 2 * (TX + 2) + 1; rii = 0; }
    else { rjj = tid - 2 * (TX + 2) - TY + 1; rii = TX + 1; }
  }
  const int ri = i0 - 1 + rii, rj = j0 - 1 + rjj;
  const bool ring_in = (tid < NRING) && ri >= 0 && ri < g.Nx && rj >= 0 && rj < g.Ny;
  const I rcol = ring_in ? (I)ri + sy * (I)rj : (I)0;
  // in-plane mask parts of the ring cell: x-comp (faces y), y-comp (faces x), z-comp
  const bool rmx = ring_in && ypart(pz, rj, g.Ny), rmy = ring_in && xpart(pz, ri, g.Nx);
  const bool rmz = ring_in && mz(pz, ri, rj, g);

  // halo H cells: bottom row (ty==0) and left column (tx==0)
  const int jb = j0 - 1, ib = i0 - 1;
  const bool hb = (ty == 0) && jb >= 0 && i < g.Nx;
  const bool hl = (tx == 0) && ib >= 0 && j < g.Ny;
  const I bcol = hb ? (I)i + sy * (I)jb : (I)0;
  const I lcol = hl ? (I)ib + sy * (I)j : (I)0;

  S Lx = 0, Ly = 0, tLx = 0, tLy = 0;
  if (i < g.Nx) { Lx = m.Lx[i]; tLx = m.tLx[i]; }
  if (j < g.Ny) { Ly = m.Ly[j]; tLy = m.tLy[j]; }
  const S Lx_l = hl ? m.Lx[ib] : S(0);
  const S Ly_b = hb ? m.Ly[jb] : S(0);
  const bool online = act && (i == x.ixs) && (j == x.iys);
  const int li = i - x.cap_ix0, lj = j - x.cap_iy0;
  const bool incap = cap && act && li >= 0 && li < x.cap_nx && lj >= 0 && lj < x.cap_ny &&
                     x.cap_sample < x.cap_samples;
  const int64_t capbase = incap ? (((int64_t)li * x.cap_ny + lj) * x.cap_samples +
                                   x.cap_sample) * g.Nz : 0;
  // in-plane mask parts of the own cell (PEC: E columns and b rows)
  const bool mI_y = act && xpart(pz, i, g.Nx), mJ_x = act && ypart(pz, j, g.Ny);
  const bool mZ = act && mz(pz, i, j, g);
  // halo-row/column cells' own E_y(k+1) / E_x(k+1) masks use the same in-plane parts
  const bool bmI_y = hb && xpart(pz, i, g.Nx), lmJ_x = hl && ypart(pz, j, g.Ny);
  // code mode: k-independent metric entries
  S iAz_ij = 0, itAz_ij = 0, iAz_b = 0, iAz_l = 0;
  if (LUT) {
    if (act) { iAz_ij = cf.iAz[(int64_t)i + (int64_t)g.Nx * j]; itAz_ij = cf.itAz[(int64_t)i + (int64_t)g.Nx * j]; }
    if (hb) iAz_b = cf.iAz[(int64_t)i + (int64_t)g.Nx * jb];
    if (hl) iAz_l = cf.iAz[(int64_t)ib + (int64_t)g.Nx * j];
  }

  // per-plane inputs (one plane ahead)
  struct In {
    S Ho[3], c0[3], c1[3], iep[3], sg[3], ring[3];
    S bH[2], bA, bE1, lH[2], lA, lE1, bA2, lA2;
    int code, bcode, lcode;
  };
  auto load = [&](int kk, In &P) {
    const bool body = kk >= k0;
    const I nk = col + sz * (I)kk;
    if (act) {
      P.Ho[0] = ldg(Hin, nk); P.Ho[1] = ldg(Hin, N + nk); P.Ho[2] = ldg(Hin, 2 * N + nk);
      if (LUT) {
        P.code = (int)cf.code[nk];
        P.c0[0] = cf.iAx[(int64_t)j + (int64_t)g.Ny * kk];
        P.c0[1] = cf.iAy[(int64_t)i + (int64_t)g.Nx * kk];
        if (body) {
          P.c1[0] = cf.itAx[(int64_t)j + (int64_t)g.Ny * kk];
          P.c1[1] = cf.itAy[(int64_t)i + (int64_t)g.Nx * kk];
        }
      } else {
        P.c0[0] = ldg(cf.a, nk); P.c0[1] = ldg(cf.a, N + nk); P.c0[2] = ldg(cf.a, 2 * N + nk);
        if (body) {
          P.c1[0] = ldg(cf.b, nk); P.c1[1] = ldg(cf.b, N + nk); P.c1[2] = ldg(cf.b, 2 * N + nk);
          if (COND) {
            P.iep[0] = ldg(cf.ieps, nk); P.iep[1] = ldg(cf.ieps, N + nk);
            P.iep[2] = ldg(cf.ieps, 2 * N + nk);
            P.sg[0] = ldg(cf.sigma, nk); P.sg[1] = ldg(cf.sigma, N + nk);
            P.sg[2] = ldg(cf.sigma, 2 * N + nk);
          } else if (JM == 1 && online) {
            P.iep[2] = ldg(cf.ieps, 2 * N + nk);
          }
        }
      }
    }
    if (ring_in) {
      const I n = rcol + sz * (I)kk;
      P.ring[0] = ldg(Ein, n); P.ring[1] = ldg(Ein, N + n); P.ring[2] = ldg(Ein, 2 * N + n);
    } else {
      P.ring[0] = P.ring[1] = P.ring[2] = S(0);
    }
    const bool kp1 = kk + 1 < g.Nz;
    if (body && hb) {
      const I n = bcol + sz * (I)kk;
      P.bH[0] = ldg(Hin, n); P.bH[1] = ldg(Hin, 2 * N + n);
      if (LUT) {
        P.bcode = (int)cf.code[n];
        P.bA = cf.iAx[(int64_t)jb + (int64_t)g.Ny * kk];
      } else {
        P.bA = ldg(cf.a, n); P.bA2 = ldg(cf.a, 2 * N + n);
      }
      P.bE1 = kp1 ? ldg(Ein, N + n + sz) : S(0);
    }
    if (body && hl) {
      const I n = lcol + sz * (I)kk;
      P.lH[0] = ldg(Hin, N + n); P.lH[1] = ldg(Hin, 2 * N + n);
      if (LUT) {
        P.lcode = (int)cf.code[n];
        P.lA = cf.iAy[(int64_t)ib + (int64_t)g.Nx * kk];
      } else {
        P.lA = ldg(cf.a, N + n); P.lA2 = ldg(cf.a, 2 * N + n);
      }
      P.lE1 = kp1 ? ldg(Ein, n + sz) : S(0);
    }
  };

  const int kstart = (k0 > 0) ? k0 - 1 : k0;
  S Ek[3] = {0, 0, 0}, Ek1[3] = {0, 0, 0};
  if (act) {
    const I n = col + sz * (I)kstart;
    Ek[0] = ldg(Ein, n); Ek[1] = ldg(Ein, N + n); Ek[2] = ldg(Ein, 2 * N + n);
    if (kstart + 1 < g.Nz) {
      Ek1[0] = ldg(Ein, n + sz); Ek1[1] = ldg(Ein, N + n + sz); Ek1[2] = ldg(Ein, 2 * N + n + sz);
    }
  }
  In P0, P1;
  load(kstart, P0);
  if (LUT) __syncthreads();  // material table staged
  S Hpx = 0, Hpy = 0;        // H_new at plane k-1 (PMC-masked), own column

  auto plane = [&](const int k, const int buf, const In &cur, In &nxt) {
    const bool body = (k >= k0);  // k0-1 is the warm-up plane (H only)
    const I nk = col + sz * (I)k;
    const bool kp1 = (k + 1 < g.Nz), kp2 = (k + 2 < g.Nz);
    if (k + 1 < kend) load(k + 1, nxt);
    S En2[3] = {0, 0, 0};
    if (act && kp2) {
      const I n2 = nk + 2 * sz;
      En2[0] = ldg(Ein, n2); En2[1] = ldg(Ein, N + n2); En2[2] = ldg(Ein, 2 * N + n2);
    }
    // PEC column masks of E at planes k and k+1 (z faces override in-plane parts)
    const int zk = zface(pz, k, g.Nz), zk1 = kp1 ? zface(pz, k + 1, g.Nz) : -1;
    const bool mEx = zsel(zk, mJ_x), mEy = zsel(zk, mI_y);
    const bool mEx1 = kp1 && zsel(zk1, mJ_x), mEy1 = kp1 && zsel(zk1, mI_y);
    sE[buf][0][ty + 1][tx + 1] = mEx ? Ek[0] : S(0);
    sE[buf][1][ty + 1][tx + 1] = mEy ? Ek[1] : S(0);
    sE[buf][2][ty + 1][tx + 1] = mZ ? Ek[2] : S(0);
    if (tid < NRING) {
      sE[buf][0][rjj][rii] = (ring_in && zsel(zk, rmx)) ? cur.ring[0] : S(0);
      sE[buf][1][rjj][rii] = (ring_in && zsel(zk, rmy)) ? cur.ring[1] : S(0);
      sE[buf][2][rjj][rii] = rmz ? cur.ring[2] : S(0);
    }
    const S Lz = m.Lz[k];
    __syncthreads();

    S Hn[3] = {0, 0, 0};
    S bo[3] = {0, 0, 0}, iep[3] = {0, 0, 0}, sg[3] = {0, 0, 0};
    if (act) {
      S ao[3];
      if (LUT) {
        const S *t = sMat + NM * cur.code;
        ao[0] = (S)(cur.c0[0] * t[0]);
        ao[1] = (S)(cur.c0[1] * t[1]);
        ao[2] = (S)(iAz_ij * t[2]);
        if (hz) {  // PMC row mask
          const int zh = zface(hz, k, g.Nz);
          if (!zsel(zh, ypart(hz, j, g.Ny))) ao[0] = S(0);
          if (!zsel(zh, xpart(hz, i, g.Nx))) ao[1] = S(0);
          if (!mz(hz, i, j, g)) ao[2] = S(0);
        }
        if (body) {  // PEC row mask
          bo[0] = mEx ? (S)(t[3] * cur.c1[0]) : S(0);
          bo[1] = mEy ? (S)(t[4] * cur.c1[1]) : S(0);
          bo[2] = mZ ? (S)(t[5] * itAz_ij) : S(0);
          if (COND) {
            iep[0] = t[3]; iep[1] = t[4]; iep[2] = t[5];
            sg[0] = t[6]; sg[1] = t[7]; sg[2] = t[8];
          } else if (JM == 1) {
            iep[2] = t[5];
          }
        }
      } else {
        ao[0] = cur.c0[0]; ao[1] = cur.c0[1]; ao[2] = cur.c0[2];
        bo[0] = cur.c1[0]; bo[1] = cur.c1[1]; bo[2] = cur.c1[2];
        if (COND) {
          iep[0] = cur.iep[0]; iep[1] = cur.iep[1]; iep[2] = cur.iep[2];
          sg[0] = cur.sg[0]; sg[1] = cur.sg[1]; sg[2] = cur.sg[2];
        } else if (JM == 1) {
          iep[2] = cur.iep[2];
        }
      }
      const S exk = sE[buf][0][ty + 1][tx + 1], eyk = sE[buf][1][ty + 1][tx + 1];
      const S ezk = sE[buf][2][ty + 1][tx + 1];
      const S ex1 = mEx1 ? Ek1[0] : S(0), ey1 = mEy1 ? Ek1[1] : S(0);
      // H_x: Ey(k)+, Ey(k+1)-, Ez(k)-, Ez(j+1)+
      Hn[0] = h_row<S, Cm, DESC>(cur.Ho[0], ao[0], dt, Ly, Lz, eyk, ey1, ezk,
                                 sE[buf][2][ty + 2][tx + 1], +1, -1);
      // H_y: Ex(k)-, Ex(k+1)+, Ez(k)+, Ez(i+1)-
      Hn[1] = h_row<S, Cm, DESC>(cur.Ho[1], ao[1], dt, Lx, Lz, exk, ex1, ezk,
                                 sE[buf][2][ty + 1][tx + 2], -1, +1);
      // H_z: Ex(k)+, Ex(j+1)-, Ey(k)-, Ey(i+1)+
      Hn[2] = h_row<S, Cm, DESC>(cur.Ho[2], ao[2], dt, Lx, Ly, exk, sE[buf][0][ty + 2][tx + 1],
                                 eyk, sE[buf][1][ty + 1][tx + 2], +1, -1);
    }
    // staged H (PMC column masks; absent cells are zero)
    {
      S h0 = Hn[0], h1 = Hn[1], h2 = Hn[2];
      if (hz && act) {
        const int zh = zface(hz, k, g.Nz);
        if (!zsel(zh, ypart(hz, j, g.Ny))) h0 = S(0);
        if (!zsel(zh, xpart(hz, i, g.Nx))) h1 = S(0);
        if (!mz(hz, i, j, g)) h2 = S(0);
      }
      sH[buf][0][ty + 1][tx + 1] = h0;
      sH[buf][1][ty + 1][tx + 1] = h1;
      sH[buf][2][ty + 1][tx + 1] = h2;
    }
    // ---- H_new(k) on the low halo row (Hx, Hz) and column (Hy, Hz); zero when absent
    if (ty == 0) {
      S hx = 0, hzv = 0;
      if (body && hb) {
        S bax, baz;
        if (LUT) {
          const S *t = sMat + NM * cur.bcode;
          bax = (S)(cur.bA * t[0]);
          baz = (S)(iAz_b * t[2]);
          if (hz) {
            const int zh = zface(hz, k, g.Nz);
            if (!zsel(zh, ypart(hz, jb, g.Ny))) bax = S(0);
            if (!mz(hz, i, jb, g)) baz = S(0);
          }
        } else {
          bax = cur.bA; baz = cur.bA2;
        }
        const S ey1 = (kp1 && zsel(zk1, bmI_y)) ? cur.bE1 : S(0);
        hx = h_row<S, Cm, DESC>(cur.bH[0], bax, dt, Ly_b, Lz, sE[buf][1][0][tx + 1], ey1,
                                sE[buf][2][0][tx + 1], sE[buf][2][1][tx + 1], +1, -1);
        hzv = h_row<S, Cm, DESC>(cur.bH[1], baz, dt, Lx, Ly_b, sE[buf][0][0][tx + 1],
                                 sE[buf][0][1][tx + 1], sE[buf][1][0][tx + 1],
                                 sE[buf][1][0][tx + 2], +1, -1);
        if (hz) {
          const int zh = zface(hz, k, g.Nz);
          if (!zsel(zh, ypart(hz, jb, g.Ny))) hx = S(0);
          if (!mz(hz, i, jb, g)) hzv = S(0);
        }
      }
      sH[buf][0][0][tx + 1] = hx;
      sH[buf][2][0][tx + 1] = hzv;
    }
    if (tx == 0) {
      S hy = 0, hzv = 0;
      if (body && hl) {
        S lay, laz;
        if (LUT) {
          const S *t = sMat + NM * cur.lcode;
          lay = (S)(cur.lA * t[1]);
          laz = (S)(iAz_l * t[2]);
          if (hz) {
            const int zh = zface(hz, k, g.Nz);
            if (!zsel(zh, xpart(hz, ib, g.Nx))) lay = S(0);
            if (!mz(hz, ib, j, g)) laz = S(0);
          }
        } else {
          lay = cur.lA; laz = cur.lA2;
        }
        const S ex1 = (kp1 && zsel(zk1, lmJ_x)) ? cur.lE1 : S(0);
        hy = h_row<S, Cm, DESC>(cur.lH[0], lay, dt, Lx_l, Lz, sE[buf][0][ty + 1][0], ex1,
                                sE[buf][2][ty + 1][0], sE[buf][2][ty + 1][1], -1, +1);
        hzv = h_row<S, Cm, DESC>(cur.lH[1], laz, dt, Lx_l, Ly, sE[buf][0][ty + 1][0],
                                 sE[buf][0][ty + 2][0], sE[buf][1][ty + 1][0],
                                 sE[buf][1][ty + 1][1], +1, -1);
        if (hz) {
          const int zh = zface(hz, k, g.Nz);
          if (!zsel(zh, xpart(hz, ib, g.Nx))) hy = S(0);
          if (!mz(hz, ib, j, g)) hzv = S(0);
        }
      }
      sH[buf][1][ty + 1][0] = hy;
      sH[buf][2][ty + 1][0] = hzv;
    }
    __syncthreads();

    // ---- E_new(k), own column
    const S hxm = sH[buf][0][ty + 1][tx + 1], hym = sH[buf][1][ty + 1][tx + 1];
    if (body && act) {
      const S hzm = sH[buf][2][ty + 1][tx + 1];
      const S tLz = m.tLz[k];
      // E_x: Hy(k-1)+, Hy(k)-, Hz(j-1)-, Hz(k)+
      const S sx = e_sum<S>(bo[0], tLy, tLz, Hpy, hym, sH[buf][2][ty][tx + 1], hzm, +1, -1);
      // E_y: Hx(k-1)-, Hx(k)+, Hz(i-1)+, Hz(k)-
      const S sy_ = e_sum<S>(bo[1], tLx, tLz, Hpx, hxm, sH[buf][2][ty + 1][tx], hzm, -1, +1);
      // E_z: Hx(j-1)+, Hx(k)-, Hy(i-1)-, Hy(k)+
      const S sz_ = e_sum<S>(bo[2], tLx, tLy, sH[buf][0][ty][tx + 1], hxm,
                             sH[buf][1][ty + 1][tx], hym, +1, -1);
      S d[3] = {sx, sy_, sz_};
      S jline = 0;
      if (online && (JM == 1 || COND)) jline = beam_row ? beam_row[k] : ldg(J, 2 * N + nk);
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
      stg(Eout, nk, En[0]); stg(Eout, N + nk, En[1]); stg(Eout, 2 * N + nk, En[2]);
      stg(Hout, nk, Hn[0]); stg(Hout, N + nk, Hn[1]); stg(Hout, 2 * N + nk, Hn[2]);
      if (COND) {
        stg(J, nk, (S)(sg[0] * En[0]));
        stg(J, N + nk, (S)(sg[1] * En[1]));
        stg(J, 2 * N + nk, (S)(sg[2] * En[2]));
      } else if (online && beam_row) {
        stg(J, 2 * N + nk, jline);
      }
      if (incap) cap[capbase + k] = En[2];
    }
    Hpx = hxm;
    Hpy = hym;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      Ek[c] = Ek1[c];
      Ek1[c] = En2[c];
    }
  };

  // unrolled by two so the plane inputs ping-pong between P0 and P1
  for (int k = kstart; k < kend; k += 2) {
    plane(k, 0, P0, P1);
    if (k + 1 < kend) plane(k + 1, 0, P1, P0);
  }
}

}  // namespace wf
