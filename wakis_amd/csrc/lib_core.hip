// libwakis_fit: C ABI of the MI355X FIT engine (declared in include/wakis_fit.h).
// Handles, the exact step path, captures, wake post-processing and reductions.
#include "lib_internal.h"

#include <mutex>
#include "monitor.h"
#include "reduce.h"
#include "wake.h"

namespace {
thread_local std::string g_err;

// Scratch of wf_count_nonfinite_noalloc: the per-workgroup partial counts
// (8 KiB). Each GPU the library runs on holds its own copy until the process
// ends; every call on a GPU shares it, so calls that overlap on one GPU race.
static __device__ __align__(256) long long scratch[WF_NOALLOC_PARTIALS];

// Device address of `scratch` on each device, resolved with that device current
// on its first call there (a host-side use of the name `scratch` is a host
// stand-in, not the device address). Fixed-size table: no allocation.
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
long long *g_scratch_ptr[MAX_DEVICES] = {};
}  // namespace

int noalloc_scratch(long long **out) {
  int dev = -1;
  WF_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES)
    return fail(WF_ERANGE, "no-allocation scratch: device %d beyond the %d-device table", dev,
                MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(WF_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                  hipGetErrorString(e));
    g_scratch_ptr[dev] = (long long *)p;
  }
  *out = g_scratch_ptr[dev];
  return WF_OK;
}

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

extern "C" {

int wf_version(void) { return 1; }

const char *wf_last_error(void) { return g_err.c_str(); }

int wf_create(const wf_grid *gd, wf_handle **out) {
  if (!gd || !out) return fail(WF_EINVAL, "wf_create: null argument");
  *out = nullptr;
  if (gd->Nx < 2 || gd->Ny < 2 || gd->Nz < 2)
    return fail(WF_EINVAL, "wf_create: need at least 2 cells per axis (got %d,%d,%d)", gd->Nx,
                gd->Ny, gd->Nz);
  if (gd->dtype != WF_F64 && gd->dtype != WF_F32)
    return fail(WF_EINVAL, "wf_create: dtype must be WF_F64 or WF_F32");
  if (!gd->Lx || !gd->Ly || !gd->Lz || !gd->tLx || !gd->tLy || !gd->tLz)
    return fail(WF_EINVAL, "wf_create: null metric vector");
  const int64_t N = (int64_t)gd->Nx * gd->Ny * gd->Nz;
  wf_handle *h = new wf_handle();
  h->g.Nx = gd->Nx; h->g.Ny = gd->Ny; h->g.Nz = gd->Nz;
  h->g.N = N; h->g.sy = gd->Nx; h->g.sz = (int64_t)gd->Nx * gd->Ny;
  h->g.pecz = gd->pec_zero_faces & 63;
  h->g.pmcz = gd->pmc_zero_faces & 63;
  h->g.m1desc = gd->m1_descending ? 1 : 0;
  h->dtype = gd->dtype;
  hipGetDevice(&h->device);
  h->ncu = 0;
  if (hipDeviceGetAttribute(&h->ncu, hipDeviceAttributeMultiprocessorCount, h->device) !=
          hipSuccess || h->ncu < 1)
    h->ncu = 1;
  const int n[6] = {gd->Nx, gd->Ny, gd->Nz, gd->Nx, gd->Ny, gd->Nz};
  const double *src[6] = {gd->Lx, gd->Ly, gd->Lz, gd->tLx, gd->tLy, gd->tLz};
  const size_t esz = gd->dtype == WF_F64 ? 8 : 4;
  const size_t tot = (size_t)(2 * (gd->Nx + gd->Ny + gd->Nz));
  std::vector<unsigned char> host(tot * esz);
  size_t off[6], o = 0;
  for (int v = 0; v < 6; ++v) {
    off[v] = o;
    for (int i = 0; i < n[v]; ++i) {
      if (esz == 8) ((double *)host.data())[o + i] = src[v][i];
      else ((float *)host.data())[o + i] = (float)src[v][i];  // numpy astype: RNE
    }
    o += n[v];
  }
  hipError_t e = hipMalloc(&h->metric_buf, tot * esz);
  if (e != hipSuccess) {
    delete h;
    return fail(WF_EHIP, "wf_create: hipMalloc: %s", hipGetErrorString(e));
  }
  e = hipMemcpy(h->metric_buf, host.data(), tot * esz, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    hipFree(h->metric_buf);
    delete h;
    return fail(WF_EHIP, "wf_create: hipMemcpy: %s", hipGetErrorString(e));
  }
  e = hipMalloc((void **)&h->red_buf, sizeof(double) * 4 * wf::RED_BLOCKS);
  if (e != hipSuccess) {
    hipFree(h->metric_buf);
    delete h;
    return fail(WF_EHIP, "wf_create: hipMalloc: %s", hipGetErrorString(e));
  }
  if (esz == 8) {
    const double *b = (const double *)h->metric_buf;
    h->md = {b + off[0], b + off[1], b + off[2], b + off[3], b + off[4], b + off[5]};
  } else {
    const float *b = (const float *)h->metric_buf;
    h->mf = {b + off[0], b + off[1], b + off[2], b + off[3], b + off[4], b + off[5]};
  }
  *out = h;
  return WF_OK;
}

int wf_destroy(wf_handle *h) {
  if (!h) return WF_OK;
  if (h->metric_buf) hipFree(h->metric_buf);
  if (h->red_buf) hipFree(h->red_buf);
  delete h;
  return WF_OK;
}

int wf_zero_ghosts(wf_handle *h, void *E, void *H, void *stream) {
  if (!h || !E || !H) return fail(WF_EINVAL, "wf_zero_ghosts: null argument");
  hipStream_t st = (hipStream_t)stream;
  if (h->dtype == WF_F64)
    hipLaunchKernelGGL(wf::k_zero_ghosts<double>, dim3(nblocks(h->g.N, 256)), dim3(256), 0, st,
                       h->g, (double *)E, (double *)H);
  else
    hipLaunchKernelGGL(wf::k_zero_ghosts<float>, dim3(nblocks(h->g.N, 256)), dim3(256), 0, st,
                       h->g, (float *)E, (float *)H);
  WF_LAUNCH_CHECK();
  return WF_OK;
}

}  // extern "C"

namespace {

template <typename S, typename Cm>
int step_general(wf_handle *h, const wf_step_args *a, const wf::MetricDev<S> &m,
                 hipStream_t st) {
  const wf::GridDev &g = h->g;
  if ((g.Ny + wf::EX_TY - 1) / wf::EX_TY > 65535)
    return fail(WF_ERANGE, "wf_step: Ny = %d is too large for the exact path", g.Ny);
  const dim3 blk(wf::EX_TX, wf::EX_TY);
  const dim3 grd((unsigned)((g.Nx + wf::EX_TX - 1) / wf::EX_TX),
                 (unsigned)((g.Ny + wf::EX_TY - 1) / wf::EX_TY), (unsigned)std::min(g.Nz, 65535));
  wf::EStepAux x;
  x.jmode = a->jmode; x.ixs = a->ixs; x.iys = a->iys;
  x.cap_ix0 = a->cap_ix0; x.cap_iy0 = a->cap_iy0; x.cap_nx = a->cap_nx; x.cap_ny = a->cap_ny;
  x.cap_samples = a->cap_samples;
  const Cm dt = (Cm)a->dt;
  for (int s = 0; s < a->nsteps; ++s) {
    hipLaunchKernelGGL((wf::k_h_exact<S, Cm>), grd, blk, 0, st, g, m, (S *)a->H,
                       (const S *)a->E, (const S *)a->a, dt);
    WF_LAUNCH_CHECK();
    const S *row = a->beam_table ? (const S *)a->beam_table + (a->beam_row0 + s) * g.Nz : nullptr;
    x.cap_sample = a->cap_sample0 + s;
    S *cap = (S *)a->capture;
#define WF_E_LAUNCH(JM, CO)                                                                  \
  hipLaunchKernelGGL((wf::k_e_exact<S, Cm, JM, CO>), grd, blk, 0, st, g, m,                \
                     (S *)a->E, (const S *)a->H, (S *)a->J, (const S *)a->b,                 \
                     (const S *)a->ieps, (const S *)a->sigma, dt, x, row, cap)
    if (a->conductive) WF_E_LAUNCH(2, true);
    else if (a->jmode == WF_J_FULL) WF_E_LAUNCH(2, false);
    else if (a->jmode == WF_J_LINE) WF_E_LAUNCH(1, false);
    else WF_E_LAUNCH(0, false);
#undef WF_E_LAUNCH
    WF_LAUNCH_CHECK();
    if (a->traj_E) {
      traj_store<S>(h, a, (const S *)a->E, (const S *)a->H, a->traj_slot0 + s + 1, st);
      WF_LAUNCH_CHECK();
    }
    if (a->mon_acc) {
      monitor_store<S>(h, a, (const S *)a->E, s, st);
      WF_LAUNCH_CHECK();
    }
  }
  return WF_OK;
}

}  // namespace

// Workgroups of `kern` that one CU runs at once (cached per handle).
int fused_occupancy(wf_handle *h, const void *kern, int threads, size_t lds) {
  for (const auto &e : h->occ)
    if (e.kern == kern && e.threads == threads && e.lds == lds) return e.per_cu;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, lds) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  h->occ.push_back({kern, threads, lds, per_cu});
  return per_cu;
}

// z-planes per workgroup. About 512 workgroups (long z marches, measured best),
// KZ in [8, 64]; then, since each workgroup's march is a serial chain whose
// time hardly depends on how many others run, a last round of workgroups that
// would fill at most a quarter of the machine is avoided by the smallest
// longer KZ that needs one round less.
int auto_kz(int64_t nxy, int Nz, int64_t conc) {
  const auto clampkz = [](int64_t k) { return (int)(k < 8 ? 8 : (k > 64 ? 64 : k)); };
  const int kz0 = clampkz((Nz * nxy + 511) / 512);
  const auto rounds = [&](int kz) { return (nxy * ((Nz + kz - 1) / kz) + conc - 1) / conc; };
  const int64_t nwg = nxy * ((Nz + kz0 - 1) / kz0);
  const int64_t r0 = rounds(kz0);
  if (r0 < 2 || nwg - (r0 - 1) * conc > conc / 4) return kz0;
  for (int kz = kz0 + 1; kz <= 64; ++kz)
    if (rounds(kz) < r0) return kz;
  return kz0;
}

// Argument checks shared by wf_step and wf_batch_create.
// trajectory box check: NULL if fine, else the reason
const char *check_traj(const wf_handle *h, const void *tE, const void *tH, const int32_t *lo,
                       const int32_t *hi) {
  if (!tE || !tH) return "needs both traj_E and traj_H";
  const int n[3] = {h->g.Nx, h->g.Ny, h->g.Nz};
  for (int d = 0; d < 3; ++d)
    if (lo[d] < 1 || hi[d] > n[d] || lo[d] >= hi[d])
      return "box must satisfy 1 <= lo < hi <= grid size on every axis";
  return nullptr;
}

int check_step_args(const wf_handle *h, const wf_step_args *a, const char *who) {
  if (!a->E || !a->H || !a->a || !a->b) return fail(WF_EINVAL, "%s: null field/coefficient", who);
  if (a->nsteps < 0) return fail(WF_EINVAL, "%s: nsteps < 0", who);
  if (a->two_step != 0 && a->two_step != 1 && a->two_step != 4 && a->two_step != 8)
    return fail(WF_EINVAL, "%s: two_step must be 0, 1, 4 or 8", who);
  if (a->jmode < WF_J_NONE || a->jmode > WF_J_FULL) return fail(WF_EINVAL, "%s: bad jmode", who);
  if (a->conductive && (!a->sigma || !a->J || a->jmode != WF_J_FULL))
    return fail(WF_EINVAL, "%s: conductive step needs J, sigma and jmode WF_J_FULL", who);
  if (a->jmode != WF_J_NONE && !a->ieps) return fail(WF_EINVAL, "%s: J term needs ieps", who);
  if (a->jmode != WF_J_NONE && !a->J && !a->beam_table)
    return fail(WF_EINVAL, "%s: J term needs J or a beam table", who);
  if (a->jmode == WF_J_FULL && !a->J) return fail(WF_EINVAL, "%s: WF_J_FULL needs J", who);
  if ((a->jmode == WF_J_LINE || a->beam_table) &&
      (a->ixs < 0 || a->ixs >= h->g.Nx || a->iys < 0 || a->iys >= h->g.Ny))
    return fail(WF_ERANGE, "%s: beam line (%d,%d) outside the grid", who, a->ixs, a->iys);
  if (a->beam_table && (a->beam_row0 < 0 || a->beam_row0 + a->nsteps > a->beam_rows))
    return fail(WF_ERANGE, "%s: beam rows [%lld,%lld) outside table of %lld rows", who,
                (long long)a->beam_row0, (long long)(a->beam_row0 + a->nsteps),
                (long long)a->beam_rows);
  if (a->beam_table && !a->J) return fail(WF_EINVAL, "%s: beam table needs J", who);
  if (a->capture) {
    if (a->cap_nx < 1 || a->cap_ny < 1 || a->cap_ix0 < 0 || a->cap_iy0 < 0 ||
        a->cap_ix0 + a->cap_nx > h->g.Nx || a->cap_iy0 + a->cap_ny > h->g.Ny)
      return fail(WF_ERANGE, "%s: capture window outside the grid", who);
    if (a->cap_sample0 < 0 || a->cap_samples < 1)
      return fail(WF_ERANGE, "%s: bad capture sample range", who);
  }
  if (a->traj_E) {
    const char *why = check_traj(h, a->traj_E, a->traj_H, a->traj_lo, a->traj_hi);
    if (why) return fail(WF_ERANGE, "%s: trajectory %s", who, why);
    if (a->traj_slot0 < 0 || a->traj_slot0 + a->nsteps + 1 > a->traj_slots)
      return fail(WF_ERANGE, "%s: trajectory slots [%lld,%lld] outside a buffer of %lld", who,
                  (long long)(a->traj_slot0 + 1), (long long)(a->traj_slot0 + a->nsteps),
                  (long long)a->traj_slots);
  }
  if (a->mon_acc && (!a->mon_phase || a->mon_nf < 1 || a->mon_row0 < 0 ||
                     (a->mon_phase_f32 != 0 && a->mon_phase_f32 != 1)))
    return fail(WF_EINVAL, "%s: the field monitor needs a phase table, mon_nf >= 1, "
                "mon_row0 >= 0 and mon_phase_f32 0 or 1", who);
  if (a->path == 1) {
    if (!a->E2 || !a->H2 || a->E2 == a->E || a->H2 == a->H || a->E2 == a->H2)
      return fail(WF_EINVAL, "%s: fused path needs distinct E2/H2 buffers", who);
    if (a->jmode == WF_J_FULL && !a->conductive)
      return fail(WF_EINVAL, "%s: fused path has no general J read (use path 0)", who);
    if (a->conductive && (!a->J || !a->ieps || !a->sigma))
      return fail(WF_EINVAL, "%s: fused conductive step needs J, ieps and sigma", who);
    if (a->kz < 0) return fail(WF_EINVAL, "%s: kz < 0", who);
    if (a->cells_per_lane != 0 && a->cells_per_lane != 1 && a->cells_per_lane != 2 &&
        a->cells_per_lane != 4)
      return fail(WF_EINVAL, "%s: cells_per_lane must be 0, 1, 2 or 4", who);
    if (a->code && (!a->mat || !a->metric2d || a->nmat < 1 || a->nmat > 256))
      return fail(WF_EINVAL, "%s: material code needs a table of 1..256 materials "
                  "and the 2-D metric tables", who);
    if (a->code && a->code_escape && (a->nmat > 255 || !a->a || !a->b || !a->ieps || !a->sigma))
      return fail(WF_EINVAL, "%s: escape code 255 needs a table of at most 255 materials "
                  "and the coefficient arrays a, b, ieps, sigma", who);
    if (a->compute_f64)
      return fail(WF_EINVAL, "%s: FP64 compute on FP32 storage uses path 0", who);
  } else if (a->path != 0) {
    return fail(WF_EINVAL, "%s: path %d not available", who, a->path);
  }
  return WF_OK;
}

extern "C" {

int wf_step(wf_handle *h, wf_step_args *a, void *stream) {
  if (!h || !a) return fail(WF_EINVAL, "wf_step: null argument");
  const int rc = check_step_args(h, a, "wf_step");
  if (rc != WF_OK) return rc;
  a->result_swapped = 0;
  a->two_step_used = 0;
  hipStream_t st = (hipStream_t)stream;
  if (a->path == 1) {
    if (h->dtype == WF_F64) return step_fused_f64(h, a, st);
    return step_fused_f32(h, a, st);
  }
  if (h->dtype == WF_F64) return step_general<double, double>(h, a, h->md, st);
  if (a->compute_f64) return step_general<float, double>(h, a, h->mf, st);
  return step_general<float, float>(h, a, h->mf, st);
}

int wf_trajectory_gather(wf_handle *h, const void *E, const void *H, void *traj_E, void *traj_H,
                         const int32_t *lo, const int32_t *hi, int64_t slots, int64_t slot,
                         void *stream) {
  if (!h || !E || !H || !lo || !hi) return fail(WF_EINVAL, "wf_trajectory_gather: null argument");
  const char *why = check_traj(h, traj_E, traj_H, lo, hi);
  if (why) return fail(WF_ERANGE, "wf_trajectory_gather: %s", why);
  if (slot < 0 || slot >= slots) return fail(WF_ERANGE, "wf_trajectory_gather: slot out of range");
  wf_step_args a;
  memset(&a, 0, sizeof(a));
  a.traj_E = traj_E; a.traj_H = traj_H;
  for (int d = 0; d < 3; ++d) { a.traj_lo[d] = lo[d]; a.traj_hi[d] = hi[d]; }
  hipStream_t st = (hipStream_t)stream;
  if (h->dtype == WF_F64) traj_store<double>(h, &a, (const double *)E, (const double *)H, slot, st);
  else traj_store<float>(h, &a, (const float *)E, (const float *)H, slot, st);
  WF_LAUNCH_CHECK();
  return WF_OK;
}

int wf_capture(wf_handle *h, const void *E, void *capture, int32_t ix0, int32_t iy0, int32_t nx,
               int32_t ny, int64_t samples, int64_t sample, void *stream) {
  if (!h || !E || !capture) return fail(WF_EINVAL, "wf_capture: null argument");
  if (nx < 1 || ny < 1 || ix0 < 0 || iy0 < 0 || ix0 + nx > h->g.Nx || iy0 + ny > h->g.Ny)
    return fail(WF_ERANGE, "wf_capture: window outside the grid");
  if (sample < 0 || sample >= samples) return fail(WF_ERANGE, "wf_capture: sample out of range");
  const int64_t tot = (int64_t)nx * ny * h->g.Nz;
  hipStream_t st = (hipStream_t)stream;
  if (h->dtype == WF_F64)
    hipLaunchKernelGGL(wf::k_capture<double>, dim3(nblocks(tot, 256)), dim3(256), 0, st, h->g,
                       (const double *)E, (double *)capture, ix0, iy0, nx, ny, samples, sample);
  else
    hipLaunchKernelGGL(wf::k_capture<float>, dim3(nblocks(tot, 256)), dim3(256), 0, st, h->g,
                       (const float *)E, (float *)capture, ix0, iy0, nx, ny, samples, sample);
  WF_LAUNCH_CHECK();
  return WF_OK;
}

int wf_abc_low(wf_handle *h, void *E, void *H, int32_t low_faces, void *stream) {
  if (!h || !E || !H) return fail(WF_EINVAL, "wf_abc_low: null argument");
  hipStream_t st = (hipStream_t)stream;
  for (int axis = 0; axis < 3; ++axis) {
    if (!((low_faces >> axis) & 1)) continue;
    const int64_t na = axis == 0 ? h->g.Ny : h->g.Nx;
    const int64_t nb = axis == 2 ? h->g.Ny : h->g.Nz;
    if (h->dtype == WF_F64)
      hipLaunchKernelGGL(wf::k_abc_face<double>, dim3(nblocks(na * nb, 256)), dim3(256), 0, st,
                         h->g, axis, (double *)E, (double *)H);
    else
      hipLaunchKernelGGL(wf::k_abc_face<float>, dim3(nblocks(na * nb, 256)), dim3(256), 0, st,
                         h->g, axis, (float *)E, (float *)H);
    WF_LAUNCH_CHECK();
  }
  return WF_OK;
}

int wf_wake_wp(const void *capture, int32_t dtype, int32_t nlines, int64_t cap_samples,
               int32_t Nz, int32_t k0, int32_t nz, const double *z, const double *dz,
               const double *s, int64_t ns, double v, double zref_v, double t0, double ti,
               double dt, int64_t nt, int64_t ncols, int32_t counter_moving, double qe12,
               double *WP3d, void *stream) {
  if (!capture || !z || !dz || !s || !WP3d) return fail(WF_EINVAL, "wf_wake_wp: null argument");
  if (nlines < 1 || Nz < 1 || k0 < 0 || nz < 0 || k0 + nz > Nz || ns < 0 || cap_samples < 1)
    return fail(WF_ERANGE, "wf_wake_wp: inconsistent sizes");
  if (nt < 0 || nt > cap_samples) return fail(WF_ERANGE, "wf_wake_wp: nt outside the captured samples");
  if (ncols < nt) return fail(WF_ERANGE, "wf_wake_wp: ncols < nt");
  if (dtype != WF_F64 && dtype != WF_F32) return fail(WF_EINVAL, "wf_wake_wp: bad dtype");
  if (ns == 0) return WF_OK;
  hipStream_t st = (hipStream_t)stream;
#define WF_WP_LAUNCH(S, CM)                                                                   \
  hipLaunchKernelGGL((wf::k_wake_wp<S, CM>), dim3(nblocks(ns, 256)), dim3(256), 0, st,        \
                     (const S *)capture, nlines, cap_samples, Nz, k0, nz, z, dz, s, ns, v,    \
                     zref_v, t0, ti, dt, nt, ncols, qe12, WP3d)
  if (dtype == WF_F64) {
    if (counter_moving) WF_WP_LAUNCH(double, true);
    else WF_WP_LAUNCH(double, false);
  } else {
    if (counter_moving) WF_WP_LAUNCH(float, true);
    else WF_WP_LAUNCH(float, false);
  }
#undef WF_WP_LAUNCH
  WF_LAUNCH_CHECK();
  return WF_OK;
}

int wf_wake_trans(const double *WP3d, int64_t ns, double ds, double dxsum, double dysum,
                  double *WPx, double *WPy, void *stream) {
  if (!WP3d || !WPx || !WPy) return fail(WF_EINVAL, "wf_wake_trans: null argument");
  if (ns < 0) return fail(WF_ERANGE, "wf_wake_trans: ns < 0");
  if (ns == 0) return WF_OK;
  hipLaunchKernelGGL(wf::k_wake_trans, dim3(nblocks(ns, 256)), dim3(256), 0, (hipStream_t)stream,
                     WP3d, ns, ds, dxsum, dysum, WPx, WPy);
  WF_LAUNCH_CHECK();
  return WF_OK;
}

int wf_dft(const double *x, int32_t nsig, int64_t len, int64_t N, int64_t nbins, double *out,
           void *stream) {
  if (!x || !out) return fail(WF_EINVAL, "wf_dft: null argument");
  if (nsig < 1 || len < 0 || N < 1 || nbins < 0 || nbins > N || len > N)
    return fail(WF_ERANGE, "wf_dft: inconsistent sizes");
  if (nbins == 0) return WF_OK;
  if (nbins > 2147483647LL || nsig > 65535) return fail(WF_ERANGE, "wf_dft: too many bins");
  hipLaunchKernelGGL(wf::k_dft, dim3((unsigned)nbins, (unsigned)nsig), dim3(256), 0,
                     (hipStream_t)stream, x, len, N, nbins, out);
  WF_LAUNCH_CHECK();
  return WF_OK;
}

int wf_energy(wf_handle *h, const void *E, const void *Hprev, const void *H, const void *ieps,
              const void *imu, const void *metric2d, double *out, void *stream) {
  if (!h || !E || !Hprev || !H || !ieps || !imu || !metric2d || !out)
    return fail(WF_EINVAL, "wf_energy: null argument");
  hipStream_t st = (hipStream_t)stream;
  const int64_t want = (h->g.N + wf::RED_THREADS - 1) / wf::RED_THREADS;
  const int nb = (int)(want < wf::RED_BLOCKS ? want : wf::RED_BLOCKS);
  if (h->dtype == WF_F64)
    hipLaunchKernelGGL(wf::k_energy<double>, dim3(nb), dim3(wf::RED_THREADS), 0, st, h->g,
                       h->md, (const double *)E, (const double *)Hprev, (const double *)H,
                       (const double *)ieps, (const double *)imu, (const double *)metric2d,
                       h->red_buf);
  else
    hipLaunchKernelGGL(wf::k_energy<float>, dim3(nb), dim3(wf::RED_THREADS), 0, st, h->g,
                       h->mf, (const float *)E, (const float *)Hprev, (const float *)H,
                       (const float *)ieps, (const float *)imu, (const float *)metric2d,
                       h->red_buf);
  WF_LAUNCH_CHECK();
  hipLaunchKernelGGL(wf::k_combine2, dim3(1), dim3(wf::RED_THREADS), 0, st,
                     (const double *)h->red_buf, nb, out);
  WF_LAUNCH_CHECK();
  return WF_OK;
}

int wf_count_nonfinite(wf_handle *h, const void *x, int64_t n, int64_t *out, void *stream) {
  if (!h || !out || (!x && n > 0)) return fail(WF_EINVAL, "wf_count_nonfinite: null argument");
  if (n < 0) return fail(WF_ERANGE, "wf_count_nonfinite: negative length");
  hipStream_t st = (hipStream_t)stream;
  const int64_t want = (n + wf::RED_THREADS - 1) / wf::RED_THREADS;
  const int nb = (int)(want < 1 ? 1 : (want < wf::RED_BLOCKS ? want : wf::RED_BLOCKS));
  long long *part = (long long *)h->red_buf;  // RED_BLOCKS*4 doubles: room for nb counts
  if (h->dtype == WF_F64)
    hipLaunchKernelGGL(wf::k_count_nonfinite<double>, dim3(nb), dim3(wf::RED_THREADS), 0, st,
                       (const double *)x, n, part);
  else
    hipLaunchKernelGGL(wf::k_count_nonfinite<float>, dim3(nb), dim3(wf::RED_THREADS), 0, st,
                       (const float *)x, n, part);
  WF_LAUNCH_CHECK();
  hipLaunchKernelGGL(wf::k_sum_counts, dim3(1), dim3(wf::RED_THREADS), 0, st,
                     (const long long *)part, nb, (long long *)out);
  WF_LAUNCH_CHECK();
  return WF_OK;
}

int wf_count_nonfinite_noalloc(const void *x, int32_t dtype, int64_t n, int32_t blocks,
                               int64_t *out, void *stream) {
  if (!out || (!x && n > 0)) return fail(WF_EINVAL, "wf_count_nonfinite_noalloc: null argument");
  if (n < 0) return fail(WF_ERANGE, "wf_count_nonfinite_noalloc: negative length");
  if (dtype != WF_F64 && dtype != WF_F32)
    return fail(WF_EINVAL, "wf_count_nonfinite_noalloc: dtype must be WF_F64 or WF_F32");
  if (blocks < 0 || blocks > WF_NOALLOC_PARTIALS)
    return fail(WF_ERANGE, "wf_count_nonfinite_noalloc: %d workgroups, the scratch holds the "
                "partial counts of at most %d", (int)blocks, WF_NOALLOC_PARTIALS);
  long long *part = nullptr;
  const int rc = noalloc_scratch(&part);
  if (rc != WF_OK) return rc;
  const int64_t want = (n + wf::RED_THREADS - 1) / wf::RED_THREADS;
  const int nb = blocks > 0 ? (int)blocks
                            : (int)(want < 1 ? 1 : (want < WF_NOALLOC_PARTIALS ? want
                                                                            : WF_NOALLOC_PARTIALS));
  hipStream_t st = (hipStream_t)stream;
  if (dtype == WF_F64)
    hipLaunchKernelGGL(wf::k_count_nonfinite<double>, dim3(nb), dim3(wf::RED_THREADS), 0, st,
                       (const double *)x, n, part);
  else
    hipLaunchKernelGGL(wf::k_count_nonfinite<float>, dim3(nb), dim3(wf::RED_THREADS), 0, st,
                       (const float *)x, n, part);
  WF_LAUNCH_CHECK();
  hipLaunchKernelGGL(wf::k_sum_counts, dim3(1), dim3(wf::RED_THREADS), 0, st,
                     (const long long *)part, nb, (long long *)out);
  WF_LAUNCH_CHECK();
  return WF_OK;
}

int wf_loss_factor(const double *WP, const double *lam, int64_t n, double ds, double *out,
                   void *stream) {
  if (!WP || !lam || !out) return fail(WF_EINVAL, "wf_loss_factor: null argument");
  if (n < 0) return fail(WF_ERANGE, "wf_loss_factor: negative length");
  hipLaunchKernelGGL(wf::k_dot, dim3(1), dim3(wf::RED_THREADS), 0, (hipStream_t)stream, WP, lam,
                     n, -ds, out);
  WF_LAUNCH_CHECK();
  return WF_OK;
}

}  // extern "C"

extern "C" {

int wf_field_monitor_update(const void *E, int32_t dtype, int64_t N, const void *phase,
                            int32_t phase_f32, int32_t nfreq, double *acc, void *stream) {
  if (!E || !phase || !acc) return fail(WF_EINVAL, "wf_field_monitor_update: null argument");
  if (N < 1 || nfreq < 0) return fail(WF_ERANGE, "wf_field_monitor_update: sizes");
  if (dtype != WF_F64 && dtype != WF_F32)
    return fail(WF_EINVAL, "wf_field_monitor_update: dtype must be WF_F64 or WF_F32");
  if (phase_f32 != 0 && phase_f32 != 1)
    return fail(WF_EINVAL, "wf_field_monitor_update: phase_f32 must be 0 or 1");
  if (nfreq == 0) return WF_OK;
  monitor_launch(E, dtype == WF_F32, N, phase, phase_f32 != 0, nfreq, acc, (hipStream_t)stream);
  WF_LAUNCH_CHECK();
  return WF_OK;
}

}  // extern "C"
