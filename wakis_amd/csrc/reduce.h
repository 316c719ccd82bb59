// Deterministic FP64 reductions: the FIT field energy of the leapfrog and the
// loss-factor integral. Each thread accumulates with Neumaier compensation over
// a fixed grid-stride set of elements; partials are combined in a fixed tree
// order inside a block, then one block combines the per-block partials, so a
// result does not depend on scheduling.
#pragma once
#include "fit_common.h"
#include "step_general.h"

namespace wf {

// running sum with its Neumaier compensation term
struct Acc {
  double s, c;
  __device__ __forceinline__ void add(double x) {
    const double t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
  }
  __device__ __forceinline__ void merge(const Acc &o) {
    add(o.s);
    c += o.c;
  }
};

constexpr int RED_THREADS = 256;
constexpr int RED_BLOCKS = 1024;  // fixed grid: the partials buffer holds RED_BLOCKS*4 doubles

// block tree combine of NV accumulators per thread; thread 0 ends with the block total
template <int NV>
__device__ void block_combine(Acc (&a)[NV]) {
  __shared__ double ss[NV][RED_THREADS], sc[NV][RED_THREADS];
  const int t = threadIdx.x;
#pragma unroll
  for (int v = 0; v < NV; ++v) { ss[v][t] = a[v].s; sc[v][t] = a[v].c; }
  __syncthreads();
  for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
    if (t < w) {
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        Acc x{ss[v][t], sc[v][t]};
        x.merge(Acc{ss[v][t + w], sc[v][t + w]});
        ss[v][t] = x.s; sc[v][t] = x.c;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int v = 0; v < NV; ++v) { a[v].s = ss[v][0]; a[v].c = sc[v][0]; }
}

// Electric part: 1/2 * L_c / (ieps * itA_c) * E^2 over edges with ieps > 0 and
// itA_c > 0 (edges of PEC cells and dual-less edges carry no energy); magnetic
// part: 1/2 * tL_c / (imu * iA_c) * Hprev * H. The weights are formed in FP64
// from the storage-precision coefficients the step uses.
template <typename S>
__global__ void __launch_bounds__(RED_THREADS)
k_energy(GridDev g, MetricDev<S> m, const S *__restrict__ E, const S *__restrict__ Hp,
         const S *__restrict__ H, const S *__restrict__ ieps, const S *__restrict__ imu,
         const S *__restrict__ metric2d, double *__restrict__ part) {
  const int64_t yz = (int64_t)g.Ny * g.Nz, xz = (int64_t)g.Nx * g.Nz, xy = (int64_t)g.Nx * g.Ny;
  const S *iA[3] = {metric2d, metric2d + yz, metric2d + yz + xz};
  const S *itA[3] = {metric2d + yz + xz + xy, metric2d + 2 * yz + xz + xy,
                     metric2d + 2 * yz + 2 * xz + xy};
  Acc a[2] = {{0.0, 0.0}, {0.0, 0.0}};
  for (int64_t n = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; n < g.N;
       n += (int64_t)gridDim.x * RED_THREADS) {
    const Cell q = cell_of(n, g);
    const int64_t t2[3] = {(int64_t)q.j + (int64_t)g.Ny * q.k, (int64_t)q.i + (int64_t)g.Nx * q.k,
                           (int64_t)q.i + (int64_t)g.Nx * q.j};
    const double L[3] = {(double)m.Lx[q.i], (double)m.Ly[q.j], (double)m.Lz[q.k]};
    const double tL[3] = {(double)m.tLx[q.i], (double)m.tLy[q.j], (double)m.tLz[q.k]};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t r = c * g.N + n;
      const double ie = (double)ieps[r], ta = (double)itA[c][t2[c]];
      if (ie > 0.0 && ta > 0.0) {
        const double e = (double)E[r];
        a[0].add(0.5 * (L[c] / (ie * ta)) * (e * e));
      }
      const double im = (double)imu[r], ia = (double)iA[c][t2[c]];
      if (tL[c] != 0.0) a[1].add(0.5 * (tL[c] / (im * ia)) * ((double)Hp[r] * (double)H[r]));
    }
  }
  block_combine<2>(a);
  if (threadIdx.x == 0) {
    part[4 * blockIdx.x + 0] = a[0].s; part[4 * blockIdx.x + 1] = a[0].c;
    part[4 * blockIdx.x + 2] = a[1].s; part[4 * blockIdx.x + 3] = a[1].c;
  }
}

// one block: combine nb partial pairs (in block order) into out[0], out[1]
static __global__ void __launch_bounds__(RED_THREADS)
k_combine2(const double *__restrict__ part, int nb, double *__restrict__ out) {
  Acc a[2] = {{0.0, 0.0}, {0.0, 0.0}};
  for (int b = threadIdx.x; b < nb; b += RED_THREADS) {
    a[0].merge(Acc{part[4 * b + 0], part[4 * b + 1]});
    a[1].merge(Acc{part[4 * b + 2], part[4 * b + 3]});
  }
  block_combine<2>(a);
  if (threadIdx.x == 0) {
    out[0] = a[0].s + a[0].c;
    out[1] = a[1].s + a[1].c;
  }
}

// one block: out = scale * sum_n x[n] * y[n]
static __global__ void __launch_bounds__(RED_THREADS)
k_dot(const double *__restrict__ x, const double *__restrict__ y, int64_t n, double scale,
      double *__restrict__ out) {
  Acc a[1] = {{0.0, 0.0}};
  for (int64_t i = threadIdx.x; i < n; i += RED_THREADS) a[0].add(x[i] * y[i]);
  block_combine<1>(a);
  if (threadIdx.x == 0) out[0] = scale * (a[0].s + a[0].c);
}

// non-finite values (Inf or NaN: all exponent bits set) of x[0, n): per-block
// counts into part[blockIdx.x], in a fixed tree order
__device__ __forceinline__ int nonfinite(float v) {
  return (__float_as_uint(v) & 0x7f800000u) == 0x7f800000u;
}
__device__ __forceinline__ int nonfinite(double v) {
  return (__double_as_longlong(v) & 0x7ff0000000000000ll) == 0x7ff0000000000000ll;
}

template <typename S>
__global__ void __launch_bounds__(RED_THREADS)
k_count_nonfinite(const S *__restrict__ x, int64_t n, long long *__restrict__ part) {
  __shared__ long long sh[RED_THREADS];
  long long c = 0;
  for (int64_t i = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * RED_THREADS)
    c += nonfinite(x[i]);
  sh[threadIdx.x] = c;
  __syncthreads();
  for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// one block: out[0] = sum of nb per-block counts
static __global__ void __launch_bounds__(RED_THREADS)
k_sum_counts(const long long *__restrict__ part, int nb, long long *__restrict__ out) {
  __shared__ long long sh[RED_THREADS];
  long long c = 0;
  for (int b = threadIdx.x; b < nb; b += RED_THREADS) c += part[b];
  sh[threadIdx.x] = c;
  __syncthreads();
  for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

}  // namespace wf
