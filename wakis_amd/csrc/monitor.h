// Running single-bin DFTs of the electric field (FieldMonitor.update,
// field_monitors.py:59-102): acc[c][f][n] += E_c[n] * phase[f] for every cell
// n and component c, with phase[f] = exp(-2 pi i f t) supplied by the caller,
// as numpy evaluates it: complex64 (P = float) when t is an FP32 scalar,
// complex128 (P = double) otherwise. The product E * phase is numpy's complex
// multiply of (E + 0i)(re + i im) = (E re - 0 im, E im + 0 re) in the promoted
// type (FP32 for FP32 E with a complex64 phase, FP64 otherwise), then added to
// the complex128 accumulator.
#pragma once
#include "fit_common.h"

#include <type_traits>

namespace wf {

template <typename S, typename P>
__global__ void __launch_bounds__(256)
k_monitor(const S *__restrict__ E, int64_t N, const P *__restrict__ phase, int nf,
          double *__restrict__ acc) {
  using R = typename std::conditional<sizeof(S) == 4 && sizeof(P) == 4, float, double>::type;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * N) return;
  const int c = (int)(t / N);
  const int64_t n = t - (int64_t)c * N;
  const R e = (R)E[t];
  for (int f = 0; f < nf; ++f) {
    double *a = acc + 2 * (((int64_t)c * nf + f) * N + n);
    const R re = (R)phase[2 * f], im = (R)phase[2 * f + 1];
    const R pr = e * re - R(0) * im;
    const R pi = e * im + R(0) * re;
    a[0] = a[0] + (double)pr;
    a[1] = a[1] + (double)pi;
  }
}

}  // namespace wf
