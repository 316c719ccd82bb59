// Internal declarations shared by the translation units of libwakis_fit.so.
// Not part of the C ABI (include/wakis_fit.h); every symbol here is hidden.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/wakis_fit.h"
#include "fit_common.h"
#include "step_general.h"
#include "adjoint.h"
#include "monitor.h"

#define WF_HIDDEN __attribute__((visibility("hidden")))

// Sets the thread's last-error text (wf_last_error) and returns `code`.
WF_HIDDEN int fail(int code, const char *fmt, ...);

#define WF_HIP(call)                                                               \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(WF_EHIP, "%s failed: %s", #call, hipGetErrorString(e_));         \
  } while (0)

#define WF_LAUNCH_CHECK()                                                          \
  do {                                                                             \
    hipError_t e_ = hipGetLastError();                                             \
    if (e_ != hipSuccess) return fail(WF_EHIP, "kernel launch: %s", hipGetErrorString(e_)); \
  } while (0)

inline unsigned nblocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

struct wf_handle {
  wf::GridDev g;
  int dtype;
  int device;
  int ncu;  // compute units of the device
  struct Occ {
    const void *kern;
    int threads;
    size_t lds;
    int per_cu;
  };
  std::vector<Occ> occ;  // workgroups per CU of the fused kernels launched so far
  void *metric_buf;  // device: 6 vectors in the storage type
  double *red_buf;   // device: per-block partials of the energy reduction
  wf::MetricDev<double> md;
  wf::MetricDev<float> mf;
};

// trajectory slot `slot` of a->traj_E / traj_H from E, H (after a step)
template <typename S>
WF_HIDDEN void traj_store(const wf_handle *h, const wf_step_args *a, const S *E, const S *H, int64_t slot,
                hipStream_t st) {
  const wf::TrajBox b = wf::traj_box(a->traj_lo, a->traj_hi);
  const int64_t tot = 3 * (b.nE + b.nH);
  hipLaunchKernelGGL((wf::k_traj_gather<S>), dim3(nblocks(tot, 256)), dim3(256), 0, st, h->g,
                     E, H, b, (S *)a->traj_E + slot * 3 * b.nE, (S *)a->traj_H + slot * 3 * b.nH);
}

// One monitor accumulation: E (FP32 or FP64) times a row of nf phases (FP32
// or FP64 (re, im) pairs).
inline void monitor_launch(const void *E, bool e32, int64_t N, const void *ph, bool p32, int nf,
                           double *acc, hipStream_t st) {
  const dim3 grid(nblocks(3 * N, 256));
  if (e32 && p32)
    hipLaunchKernelGGL((wf::k_monitor<float, float>), grid, dim3(256), 0, st, (const float *)E,
                       N, (const float *)ph, nf, acc);
  else if (e32)
    hipLaunchKernelGGL((wf::k_monitor<float, double>), grid, dim3(256), 0, st,
                       (const float *)E, N, (const double *)ph, nf, acc);
  else if (p32)
    hipLaunchKernelGGL((wf::k_monitor<double, float>), grid, dim3(256), 0, st,
                       (const double *)E, N, (const float *)ph, nf, acc);
  else
    hipLaunchKernelGGL((wf::k_monitor<double, double>), grid, dim3(256), 0, st,
                       (const double *)E, N, (const double *)ph, nf, acc);
}

// FieldMonitor accumulation of the E that step s of the call produced
// (wf_step_args.mon_*).
template <typename S>
WF_HIDDEN void monitor_store(const wf_handle *h, const wf_step_args *a, const S *E, int64_t s,
                             hipStream_t st) {
  const size_t pb = a->mon_phase_f32 ? sizeof(float) : sizeof(double);
  const char *ph = (const char *)a->mon_phase + (size_t)((a->mon_row0 + s) * 2 * a->mon_nf) * pb;
  monitor_launch(E, sizeof(S) == 4, h->g.N, ph, a->mon_phase_f32 != 0, a->mon_nf, a->mon_acc,
                 st);
}

// Workgroups of `kern` that one CU runs at once (cached per handle).
WF_HIDDEN int fused_occupancy(wf_handle *h, const void *kern, int threads, size_t lds);

// z-planes per workgroup for nxy tiles per plane and `conc` resident workgroups.
WF_HIDDEN int auto_kz(int64_t nxy, int Nz, int64_t conc);

// Full-row tiles: dynamic LDS bytes of a tile of `rows` own rows (E and H rows
// of Nx + 2V slots, then the material table).
inline size_t full_rows_lds(int Nx, int V, int rows, size_t sz, size_t table) {
  return (size_t)(6 * rows + 9) * (size_t)(Nx + 2 * V) * sz + table;
}
constexpr int FR_BOUND = 7;            // full-row kernels: up to 64 * 8 threads
constexpr size_t FR_LDS_MAX = 65536;   // and 64 KiB of LDS per workgroup

constexpr int BATCH_GRAPH_STEPS = 16;  // batch steps captured in one HIP graph

// Argument checks shared by wf_step and wf_batch_create.
WF_HIDDEN const char *check_traj(const wf_handle *h, const void *tE, const void *tH,
                                 const int32_t *lo, const int32_t *hi);
WF_HIDDEN int check_step_args(const wf_handle *h, const wf_step_args *a, const char *who);

// Device address (current device) of the static scratch of the
// no-allocation entry points; resolved once per device.
WF_HIDDEN int noalloc_scratch(long long **out);

// The fused single-pass step (path 1) for FP32 and FP64 storage.
WF_HIDDEN int step_fused_f32(wf_handle *h, wf_step_args *a, hipStream_t st);
// two steps per launch (lib_fused_t2.hip): 1 if this call qualifies
WF_HIDDEN int fused_t2_applies(const wf_handle *h, const wf_step_args *a);
WF_HIDDEN int step_fused_t2_f32(wf_handle *h, wf_step_args *a, hipStream_t st);
WF_HIDDEN int step_fused_t2_f64(wf_handle *h, wf_step_args *a, hipStream_t st);
WF_HIDDEN int step_fused_f64(wf_handle *h, wf_step_args *a, hipStream_t st);
