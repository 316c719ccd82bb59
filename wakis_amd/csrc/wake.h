// Wake potential and impedance kernels (FP64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wf {

// WakeSolver.calc_long_WP_3d (wakeSolver.py:485-537): one thread per s sample.
// The thread walks the captured lines in order carrying the running value
// (WP is zeroed once, then WP += sum_k Ez*dz and WP /= q*1e12 per line),
// accumulating over k in increasing order: the exact operation sequence of
// the reference's loop for this s.
//   forward (CM false, wakeSolver.py:520-533): row k of the line,
//     ts = (z[k] + s)/v - zref_v - t0 + ti, weight dz[k], zref_v = zmin/v;
//   counter-moving (CM true, wakeSolver.py:503-517): row nz-1-k,
//     ts = (z[nz-1-k] - s)/(-v) - zref_v - t0 + ti, weight (-dz[k]),
//     zref_v = zmax/(-v).
// it = trunc(ts/dt) is kept when it < nt; a negative it indexes from the end
// of the reference's Ezt, which has ncols columns of which the first nt hold
// samples (the others are zero).
template <typename S, bool CM>
__global__ void __launch_bounds__(256)
k_wake_wp(const S *__restrict__ cap, int nlines, int64_t cap_samples, int Nz, int k0, int nz,
          const double *__restrict__ z, const double *__restrict__ dz,
          const double *__restrict__ s, int64_t ns, double v, double zref_v, double t0,
          double ti, double dt, int64_t nt, int64_t ncols, double qe12,
          double *__restrict__ WP3d) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= ns) return;
  const double sn = s[n];
  const double nv = -v;
  double acc = 0.0;
  for (int l = 0; l < nlines; ++l) {
    const S *line = cap + (int64_t)l * cap_samples * Nz + k0;
    for (int k = 0; k < nz; ++k) {
      const int kr = CM ? nz - 1 - k : k;  // row of Ezt
      double ts = CM ? (z[kr] - sn) / nv : (z[kr] + sn) / v;
      ts = ts - zref_v;
      ts = ts - t0;
      ts = ts + ti;
      const int64_t it = (int64_t)(ts / dt);
      if (it < nt) {
        const int64_t col = it < 0 ? it + ncols : it;  // python negative index
        double e = 0.0;
        if (col >= 0 && col < nt) e = (double)line[col * Nz + kr];
        acc = acc + e * (CM ? -dz[k] : dz[k]);
      }
    }
    acc = acc / qe12;
    WP3d[(int64_t)l * ns + n] = acc;
  }
}

// numpy's float64 add-reduce of a contiguous block of at most 8192 values:
// fewer than 8 -> sequential; up to 128 -> 8 interleaved accumulators combined
// ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)) then the tail; otherwise split at
// n2 = n/2 - (n/2 % 8) and add the two halves. Iterative (explicit stack).
__device__ inline double np_leaf(const double *a, int64_t n) {
  if (n < 8) {
    double r = 0.0;
    for (int64_t i = 0; i < n; ++i) r = r + a[i];
    return r;
  }
  double r0 = a[0], r1 = a[1], r2 = a[2], r3 = a[3], r4 = a[4], r5 = a[5], r6 = a[6], r7 = a[7];
  int64_t i = 8;
  for (; i < n - (n % 8); i += 8) {
    r0 = r0 + a[i + 0]; r1 = r1 + a[i + 1]; r2 = r2 + a[i + 2]; r3 = r3 + a[i + 3];
    r4 = r4 + a[i + 4]; r5 = r5 + a[i + 5]; r6 = r6 + a[i + 6]; r7 = r7 + a[i + 7];
  }
  double res = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
  for (; i < n; ++i) res = res + a[i];
  return res;
}

__device__ inline double np_pairwise(const double *a, int64_t n) {
  int64_t lo[16], len[16];
  int st[16];
  double left[16];
  int sp = 0;
  lo[0] = 0; len[0] = n; st[0] = 0;
  double ret = 0.0;
  for (;;) {
    // invariant: node sp is unvisited (st[sp] == 0)
    bool finished = false;
    if (len[sp] <= 128) {
      ret = np_leaf(a + lo[sp], len[sp]);
      finished = true;
    } else {
      int64_t n2 = len[sp] / 2;
      n2 -= n2 % 8;
      st[sp] = 1;
      lo[sp + 1] = lo[sp]; len[sp + 1] = n2; st[sp + 1] = 0;
      ++sp;
    }
    // pop finished nodes; a parent whose left child finished pushes its right child
    while (finished) {
      if (sp == 0) return ret;
      --sp;
      if (st[sp] == 1) {
        left[sp] = ret;
        st[sp] = 2;
        int64_t n2 = len[sp] / 2;
        n2 -= n2 % 8;
        lo[sp + 1] = lo[sp] + n2; len[sp + 1] = len[sp] - n2; st[sp + 1] = 0;
        ++sp;
        finished = false;
      } else {
        ret = left[sp] + ret;
      }
    }
  }
}

__device__ inline double np_sum(const double *a, int64_t n) {
  double res = 0.0;
  for (int64_t lo = 0; lo < n; lo += 8192) {
    const int64_t m = (n - lo) < 8192 ? (n - lo) : 8192;
    res = res + np_pairwise(a + lo, m);
  }
  return res;
}

// WakeSolver.calc_trans_WP (wakeSolver.py:583-612) on a 3x3 window (i-major).
static __global__ void __launch_bounds__(256)
k_wake_trans(const double *__restrict__ WP3d, int64_t ns, double ds, double dxsum,
             double dysum, double *__restrict__ WPx, double *__restrict__ WPy) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= ns) return;
  const double i21 = np_sum(WP3d + 7 * ns, n) * ds;
  const double i01 = np_sum(WP3d + 1 * ns, n) * ds;
  const double i12 = np_sum(WP3d + 5 * ns, n) * ds;
  const double i10 = np_sum(WP3d + 3 * ns, n) * ds;
  WPx[n] = (-(i21 - i01)) / dxsum;
  WPy[n] = (-(i12 - i10)) / dysum;
}

// Neumaier compensated accumulation.
struct Comp {
  double s, c;
  __device__ void add(double x) {
    const double t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
  }
};

// One workgroup per (signal, bin): X[f] = sum_n x[n] exp(-2 pi i f n / N).
static __global__ void __launch_bounds__(256)
k_dft(const double *__restrict__ x, int64_t len, int64_t N, int64_t nbins,
      double *__restrict__ out) {
  const int64_t f = blockIdx.x;
  const int sig = blockIdx.y;
  const double *xs = x + (int64_t)sig * len;
  Comp re{0.0, 0.0}, im{0.0, 0.0};
  for (int64_t n = threadIdx.x; n < len; n += blockDim.x) {
    const int64_t ph = (f * n) % N;  // exact phase reduction
    double sn, cs;
    sincospi(-2.0 * (double)ph / (double)N, &sn, &cs);
    re.add(xs[n] * cs);
    im.add(xs[n] * sn);
  }
  __shared__ double sre[256], sim[256];
  sre[threadIdx.x] = re.s + re.c;
  sim[threadIdx.x] = im.s + im.c;
  __syncthreads();
  if (threadIdx.x == 0) {
    Comp r{0.0, 0.0}, i{0.0, 0.0};
    for (int t = 0; t < (int)blockDim.x; ++t) {
      r.add(sre[t]);
      i.add(sim[t]);
    }
    out[((int64_t)sig * nbins + f) * 2 + 0] = r.s + r.c;
    out[((int64_t)sig * nbins + f) * 2 + 1] = i.s + i.c;
  }
}

}  // namespace wf
