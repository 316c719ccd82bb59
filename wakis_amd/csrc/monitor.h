// Running single-bin DFTs of the electric field (FieldMonitor.update,
// field_monitors.py:59-102): acc[c][f][n] += E_c[n] * phase[f] for every cell
// n and component c, with phase[f] = exp(-2 pi i f t) supplied by the caller
// (the reference evaluates it with numpy in complex128). For a real E the
// complex product is (E*re, E*im), which is what numpy's complex multiply of
// (E + 0i)(re + i im) rounds to.
#pragma once
#include "fit_common.h"

namespace wf {

template <typename S>
__global__ void __launch_bounds__(256)
k_monitor(const S *__restrict__ E, int64_t N, const double *__restrict__ phase, int nf,
          double *__restrict__ acc) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * N) return;
  const int c = (int)(t / N);
  const int64_t n = t - (int64_t)c * N;
  const double e = (double)E[t];
  for (int f = 0; f < nf; ++f) {
    double *a = acc + 2 * (((int64_t)c * nf + f) * N + n);
    const double re = phase[2 * f], im = phase[2 * f + 1];
    a[0] = a[0] + e * re;
    a[1] = a[1] + e * im;
  }
}

}  // namespace wf
