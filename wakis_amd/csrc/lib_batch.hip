// Batches: one launch per step over many grids (wf_batch_*).
#include "lib_internal.h"
#include "step_fused4.h"

struct wf_batch {
  int dtype = 0, ninst = 0, ntiles = 0, maxsteps = 0;
  int64_t traj_threads = 0;   // batched trajectory gather threads (0: none)
  int rows = 3, vw = 1, cond = 0, jm = 0, desc = 0;
  int lut = 1;                // material codes (1) or coefficient arrays (0)
  int fr = 0;                 // whole-row tiles (16-byte lanes), rows per grid
  size_t lds = 0;
  void *desc_dev = nullptr;   // BatchDesc<S>[ninst]
  int *step_dev = nullptr;    // step number read by the batch kernel
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_steps = 0;
};

namespace {

constexpr int BATCH_ROWS = 3;

template <typename S>
int batch_build(wf_handle *const *hs, const wf_step_args *as, int n, int vw, int kz,
                wf_batch *b) {
  std::vector<wf::BatchDesc<S>> d(n);
  int64_t tiles = 0;
  size_t lds = 0;
  const int W = 64 * vw;
  for (int i = 0; i < n; ++i) {
    const wf_handle *h = hs[i];
    const wf_step_args *a = &as[i];
    wf::BatchDesc<S> &e = d[i];
    e.g = h->g;
    if constexpr (sizeof(S) == 8) e.m = h->md;
    else e.m = h->mf;
    e.E = (S *)a->E; e.H = (S *)a->H; e.E2 = (S *)a->E2; e.H2 = (S *)a->H2; e.J = (S *)a->J;
    wf::Coefs<S> &cf = e.cf;
    cf.a = (const S *)a->a; cf.b = (const S *)a->b;
    cf.ieps = (const S *)a->ieps; cf.sigma = (const S *)a->sigma;
    cf.code = (const uint8_t *)a->code; cf.mat = (const S *)a->mat; cf.nmat = a->nmat;
    if (b->lut) {
      const S *t = (const S *)a->metric2d;
      const int64_t yz = (int64_t)h->g.Ny * h->g.Nz, xz = (int64_t)h->g.Nx * h->g.Nz,
                    xy = (int64_t)h->g.Nx * h->g.Ny;
      cf.iAx = t; cf.iAy = t + yz; cf.iAz = t + yz + xz;
      cf.itAx = t + yz + xz + xy; cf.itAy = cf.itAx + yz; cf.itAz = cf.itAy + xz;
    }
    e.dt = (S)a->dt;
    wf::FusedAux &x = e.x;
    x.ixs = (a->jmode == WF_J_LINE || a->beam_table || a->conductive) ? a->ixs : -1;
    x.iys = (x.ixs >= 0) ? a->iys : -1;
    x.cap_ix0 = a->cap_ix0; x.cap_iy0 = a->cap_iy0; x.cap_nx = a->cap_nx; x.cap_ny = a->cap_ny;
    x.cap_samples = a->cap_samples;
    x.cap_sample = 0;
    x.tx0 = x.ty0 = x.tz0 = 0;
    if (b->fr) {
      // whole rows: as many as fill the batch workgroup (and 64 KiB of LDS)
      const size_t table = b->lut ? (size_t)a->nmat * 9 * sizeof(S) : 0;
      x.qrow = h->g.Nx / vw;
      int T = 64 * (BATCH_ROWS + 1) / x.qrow - 1;
      if (T > h->g.Ny) T = h->g.Ny;
      while (T >= 1 && full_rows_lds(h->g.Nx, vw, T, sizeof(S), table) > FR_LDS_MAX) --T;
      if (T < 1)
        return fail(WF_EINVAL, "wf_batch_create[%d]: rows of %d cells do not fit a whole-row "
                    "tile", i, h->g.Nx);
      x.rows = T;
      x.ntx = 1;
      x.nty = (h->g.Ny + T - 1) / T;
      lds = std::max(lds, full_rows_lds(h->g.Nx, vw, T, sizeof(S), table));
    } else {
      x.rows = BATCH_ROWS;
      x.qrow = 64;
      x.ntx = (h->g.Nx + W - 1) / W;
      x.nty = (h->g.Ny + BATCH_ROWS - 1) / BATCH_ROWS;
    }
    x.KZ = kz > 0 ? kz : (h->g.Nz < 32 ? h->g.Nz : 32);
    x.ntz = (h->g.Nz + x.KZ - 1) / x.KZ;
    e.beam_table = (const S *)a->beam_table;
    e.beam_row0 = a->beam_row0;
    e.cap = a->capture ? (S *)a->capture : nullptr;
    e.cap_sample0 = a->cap_sample0;
    e.nsteps = a->nsteps;
    e.tile0 = (int)tiles;
    tiles += (int64_t)x.ntx * x.nty * x.ntz;
    e.thr0 = b->traj_threads;
    if (a->traj_E) {
      e.tE = (S *)a->traj_E; e.tH = (S *)a->traj_H;
      for (int q = 0; q < 3; ++q) { e.tlo[q] = a->traj_lo[q]; e.thi[q] = a->traj_hi[q]; }
      e.traj_slot0 = a->traj_slot0;
      const wf::TrajBox tb = wf::traj_box(a->traj_lo, a->traj_hi);
      b->traj_threads += 3 * (tb.nE + tb.nH);
    }
    if (b->lut && !b->fr) lds = std::max(lds, (size_t)a->nmat * 9 * sizeof(S));
    b->maxsteps = std::max(b->maxsteps, a->nsteps);
  }
  if (tiles > 2147483647LL) return fail(WF_ERANGE, "wf_batch_create: too many tiles");
  b->ntiles = (int)tiles;
  b->lds = lds;
  WF_HIP(hipMalloc(&b->desc_dev, sizeof(wf::BatchDesc<S>) * n));
  WF_HIP(hipMemcpy(b->desc_dev, d.data(), sizeof(wf::BatchDesc<S>) * n, hipMemcpyHostToDevice));
  WF_HIP(hipMalloc((void **)&b->step_dev, sizeof(int)));
  return WF_OK;
}

// launch one batch step (the step number is read on the device)
template <typename S>
int batch_launch(const wf_batch *b, hipStream_t st) {
  const auto *d = (const wf::BatchDesc<S> *)b->desc_dev;
  const dim3 grid((unsigned)b->ntiles), block(64 * (BATCH_ROWS + 1));
#define WF_B_LAUNCH_FR(JM, CO, DE, VW, FR)                                                   \
  do {                                                                                       \
    if (b->lut)                                                                              \
      hipLaunchKernelGGL((wf::k_fused4_batch<S, JM, CO, true, DE, BATCH_ROWS, false, VW, FR>), \
                         grid, block, b->lds, st, d, b->ninst, b->ntiles,                    \
                         (const int *)b->step_dev);                                          \
    else                                                                                     \
      hipLaunchKernelGGL((wf::k_fused4_batch<S, JM, CO, false, DE, BATCH_ROWS, true, VW, FR>), \
                         grid, block, b->lds, st, d, b->ninst, b->ntiles,                    \
                         (const int *)b->step_dev);                                          \
  } while (0)
#define WF_B_LAUNCH(JM, CO, DE, VW)                                                          \
  do {                                                                                       \
    if (b->fr) {                                                                             \
      if constexpr (VW != 1) WF_B_LAUNCH_FR(JM, CO, DE, VW, true);                           \
    } else {                                                                                 \
      WF_B_LAUNCH_FR(JM, CO, DE, VW, false);                                                 \
    }                                                                                        \
  } while (0)
#define WF_B_LAUNCH2(JM, CO, VW)                                                             \
  do {                                                                                       \
    if (b->desc) WF_B_LAUNCH(JM, CO, true, VW);                                              \
    else WF_B_LAUNCH(JM, CO, false, VW);                                                     \
  } while (0)
#define WF_B_LAUNCH3(VW)                                                                     \
  do {                                                                                       \
    if (b->cond) WF_B_LAUNCH2(0, true, VW);                                                  \
    else if (b->jm == WF_J_LINE) WF_B_LAUNCH2(1, false, VW);                                 \
    else WF_B_LAUNCH2(0, false, VW);                                                         \
  } while (0)
  if (b->vw == 1) WF_B_LAUNCH3(1);
  else WF_B_LAUNCH3(16 / (int)sizeof(S));
#undef WF_B_LAUNCH3
#undef WF_B_LAUNCH2
#undef WF_B_LAUNCH
#undef WF_B_LAUNCH_FR
  WF_LAUNCH_CHECK();
  if (b->traj_threads > 0) {
    hipLaunchKernelGGL(wf::k_traj_gather_batch<S>, dim3(nblocks(b->traj_threads, 256)),
                       dim3(256), 0, st, d, b->ninst, b->traj_threads,
                       (const int *)b->step_dev);
    WF_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(wf::k_batch_next, dim3(1), dim3(1), 0, st, b->step_dev);
  WF_LAUNCH_CHECK();
  return WF_OK;
}

int batch_step_once(const wf_batch *b, hipStream_t st) {
  return b->dtype == WF_F64 ? batch_launch<double>(b, st) : batch_launch<float>(b, st);
}

}  // namespace

extern "C" {

int wf_batch_create(wf_handle *const *handles, const wf_step_args *args, int32_t n,
                    int32_t cells_per_lane, int32_t kz, wf_batch **out) {
  if (!handles || !args || !out) return fail(WF_EINVAL, "wf_batch_create: null argument");
  *out = nullptr;
  if (n < 1) return fail(WF_EINVAL, "wf_batch_create: empty batch");
  if (kz < 0) return fail(WF_EINVAL, "wf_batch_create: kz < 0");
  const wf_handle *h0 = handles[0];
  if (!h0) return fail(WF_EINVAL, "wf_batch_create: null handle");
  const int dtype = h0->dtype;
  const int vnat = dtype == WF_F64 ? 2 : 4;
  const wf_step_args &a0 = args[0];
  const int fr = a0.full_rows > 0 ? 1 : 0;
  const int vw = (cells_per_lane == 0 || fr) ? vnat : cells_per_lane;
  if (vw != 1 && vw != vnat)
    return fail(WF_EINVAL, "wf_batch_create: cells_per_lane must be 0, 1 or %d", vnat);
  if (fr && cells_per_lane != 0 && cells_per_lane != vnat)
    return fail(WF_EINVAL, "wf_batch_create: whole-row tiles use %d cells per lane", vnat);
  for (int i = 0; i < n; ++i) {
    const wf_handle *h = handles[i];
    const wf_step_args *a = &args[i];
    if (!h) return fail(WF_EINVAL, "wf_batch_create: null handle %d", i);
    char who[64];
    snprintf(who, sizeof(who), "wf_batch_create[%d]", i);
    const int rc = check_step_args(h, a, who);
    if (rc != WF_OK) return rc;
    if (a->path != 1) return fail(WF_EINVAL, "%s: batch items use the fused path (path 1)", who);
    if (a->mon_acc) return fail(WF_EINVAL, "%s: batch items carry no field monitor", who);
    const bool lut = a->code && a->mat && a->nmat > 0 && a->metric2d;
    if (lut != (a0.code && a0.mat && a0.nmat > 0 && a0.metric2d))
      return fail(WF_EINVAL, "%s: all batch items use material codes, or none", who);
    if (lut && !a->zero_masked)
      return fail(WF_EINVAL, "%s: batch items with material codes need zero_masked = 1", who);
    if (lut && a->code_escape)
      return fail(WF_EINVAL, "%s: batch items use material codes without the escape code", who);
    if (h->dtype != dtype || a->conductive != a0.conductive || a->jmode != a0.jmode ||
        h->g.m1desc != h0->g.m1desc)
      return fail(WF_EINVAL, "%s: dtype, conductivity, J mode and boundary order must match "
                  "across the batch", who);
    if (h->g.Nx % vw != 0)
      return fail(WF_EINVAL, "%s: Nx=%d not a multiple of %d cells per lane", who, h->g.Nx, vw);
    if ((a->full_rows > 0 ? 1 : 0) != fr)
      return fail(WF_EINVAL, "%s: all batch items use whole-row tiles, or none", who);
    if ((uint64_t)3 * (uint64_t)h->g.N * (dtype == WF_F64 ? 8 : 4) >= (1ull << 32))
      return fail(WF_ERANGE, "%s: grid too large for a batch", who);
  }
  wf_batch *b = new wf_batch();
  b->dtype = dtype; b->ninst = n; b->vw = vw;
  b->cond = a0.conductive; b->jm = a0.jmode; b->desc = h0->g.m1desc;
  b->lut = (a0.code && a0.mat && a0.nmat > 0 && a0.metric2d) ? 1 : 0;
  b->fr = fr;
  const int rc = dtype == WF_F64 ? batch_build<double>(handles, args, n, vw, kz, b)
                                 : batch_build<float>(handles, args, n, vw, kz, b);
  if (rc != WF_OK) {
    wf_batch_destroy(b);
    return rc;
  }
  *out = b;
  return WF_OK;
}

int wf_batch_run(wf_batch *b, int32_t step0, int32_t nsteps, int32_t use_graph, void *stream) {
  if (!b) return fail(WF_EINVAL, "wf_batch_run: null batch");
  if (step0 < 0 || nsteps < 0) return fail(WF_EINVAL, "wf_batch_run: negative step range");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(wf::k_batch_set, dim3(1), dim3(1), 0, st, b->step_dev, (int)step0);
  WF_LAUNCH_CHECK();
  int done = 0;
  if (use_graph && nsteps >= BATCH_GRAPH_STEPS) {
    if (!b->exec) {
      WF_HIP(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
      WF_HIP(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
      int rc = WF_OK;
      for (int s = 0; s < BATCH_GRAPH_STEPS && rc == WF_OK; ++s) rc = batch_step_once(b, b->cap_stream);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
      if (rc != WF_OK) return rc;
      if (e != hipSuccess) return fail(WF_EHIP, "wf_batch_run: capture: %s", hipGetErrorString(e));
      b->graph = g;
      WF_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      b->graph_steps = BATCH_GRAPH_STEPS;
    }
    for (; done + b->graph_steps <= nsteps; done += b->graph_steps)
      WF_HIP(hipGraphLaunch(b->exec, st));
  }
  for (; done < nsteps; ++done) {
    const int rc = batch_step_once(b, st);
    if (rc != WF_OK) return rc;
  }
  return WF_OK;
}

int wf_batch_destroy(wf_batch *b) {
  if (!b) return WF_OK;
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  if (b->cap_stream) hipStreamDestroy(b->cap_stream);
  if (b->desc_dev) hipFree(b->desc_dev);
  if (b->step_dev) hipFree(b->step_dev);
  delete b;
  return WF_OK;
}

}  // extern "C"
