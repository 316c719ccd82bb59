// libwakis_fit: C ABI of the MI355X FIT engine (declared in include/wakis_fit.h).
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/wakis_fit.h"
#include "fit_common.h"
#include "step_general.h"
#include "step_fused.h"
#include "step_fused4.h"
#include "wake.h"
#include "reduce.h"
#include "adjoint.h"
#include "adjoint_fused.h"
#include "monitor.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

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

}  // namespace

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

// trajectory slot `slot` of a->traj_E / traj_H from E, H (after a step)
template <typename S>
void traj_store(const wf_handle *h, const wf_step_args *a, const S *E, const S *H, int64_t slot,
                hipStream_t st) {
  const wf::TrajBox b = wf::traj_box(a->traj_lo, a->traj_hi);
  const int64_t tot = 3 * (b.nE + b.nH);
  hipLaunchKernelGGL((wf::k_traj_gather<S>), dim3(nblocks(tot, 256)), dim3(256), 0, st, h->g,
                     E, H, b, (S *)a->traj_E + slot * 3 * b.nE, (S *)a->traj_H + slot * 3 * b.nH);
}

template <typename S, typename Cm>
int step_general(wf_handle *h, const wf_step_args *a, const wf::MetricDev<S> &m,
                 hipStream_t st) {
  const wf::GridDev &g = h->g;
  const unsigned nb = nblocks(g.N, 256);
  wf::EStepAux x;
  x.jmode = a->jmode; x.ixs = a->ixs; x.iys = a->iys;
  x.cap_ix0 = a->cap_ix0; x.cap_iy0 = a->cap_iy0; x.cap_nx = a->cap_nx; x.cap_ny = a->cap_ny;
  x.cap_samples = a->cap_samples;
  const Cm dt = (Cm)a->dt;
  for (int s = 0; s < a->nsteps; ++s) {
    hipLaunchKernelGGL((wf::k_h_general<S, Cm>), dim3(nb), dim3(256), 0, st, g, m, (S *)a->H,
                       (const S *)a->E, (const S *)a->a, dt);
    WF_LAUNCH_CHECK();
    const S *row = a->beam_table ? (const S *)a->beam_table + (a->beam_row0 + s) * g.Nz : nullptr;
    x.cap_sample = a->cap_sample0 + s;
    S *cap = (S *)a->capture;
#define WF_E_LAUNCH(JM, CO)                                                                  \
  hipLaunchKernelGGL((wf::k_e_general<S, Cm, JM, CO>), dim3(nb), dim3(256), 0, st, g, m,    \
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
  }
  return WF_OK;
}

constexpr int FTX = 64, FTY = 4;

template <typename S>
using Fused4Fn = void (*)(wf::GridDev, wf::MetricDev<S>, const S *, const S *, S *, S *, S *,
                          wf::Coefs<S>, S, wf::FusedAux, const S *, S *);

// k_fused4, or k_fused4_wide (64-bit offsets) for arrays past 4 GiB
template <typename S, int JM, bool CO, bool LU, bool DE, int TY, bool MK, int V, bool FR,
          bool ES, bool WIDE>
Fused4Fn<S> fused4_ptr() {
  if constexpr (WIDE) return &wf::k_fused4_wide<S, JM, CO, LU, DE, TY, MK, V>;
  else return &wf::k_fused4<S, JM, CO, LU, DE, TY, MK, V, FR, ES>;
}

// The k_fused4 instantiation for the step's J mode, conductivity, coefficient
// form, boundary order and masking.
template <typename S, int TY, int V, bool FR, bool WIDE = false>
Fused4Fn<S> pick_fused4(bool lut, bool cond, bool jline, bool desc, bool mk, bool esc = false) {
#define WF_PICK(JM, CO, LU, ES)                                                             \
  do {                                                                                      \
    if (desc) return mk ? fused4_ptr<S, JM, CO, LU, true, TY, true, V, FR, ES, WIDE>()      \
                        : fused4_ptr<S, JM, CO, LU, true, TY, !LU, V, FR, ES, WIDE>();      \
    return mk ? fused4_ptr<S, JM, CO, LU, false, TY, true, V, FR, ES, WIDE>()               \
              : fused4_ptr<S, JM, CO, LU, false, TY, !LU, V, FR, ES, WIDE>();               \
  } while (0)
  if constexpr (!WIDE) {
    if (lut && esc) {
      if (cond) WF_PICK(0, true, true, true);
      if (jline) WF_PICK(1, false, true, true);
      WF_PICK(0, false, true, true);
    }
  }
  if (lut) {
    if (cond) WF_PICK(0, true, true, false);
    if (jline) WF_PICK(1, false, true, false);
    WF_PICK(0, false, true, false);
  }
  if (cond) WF_PICK(0, true, false, false);
  if (jline) WF_PICK(1, false, false, false);
  WF_PICK(0, false, false, false);
#undef WF_PICK
}

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

// Full-row tiles: dynamic LDS bytes of a tile of `rows` own rows (E and H rows
// of Nx + 2V slots, then the material table).
inline size_t full_rows_lds(int Nx, int V, int rows, size_t sz, size_t table) {
  return (size_t)(6 * rows + 9) * (size_t)(Nx + 2 * V) * sz + table;
}
constexpr int FR_BOUND = 7;            // full-row kernels: up to 64 * 8 threads
constexpr size_t FR_LDS_MAX = 65536;   // and 64 KiB of LDS per workgroup

template <typename S, int TY, int V, bool FR = false, bool WIDE = false>
int step_fused4(wf_handle *h, wf_step_args *a, const wf::MetricDev<S> &m, hipStream_t st,
                int fr_rows = 0) {
  const wf::GridDev &g = h->g;
  wf::FusedAux x;
  x.ixs = (a->jmode == WF_J_LINE || a->beam_table || a->conductive) ? a->ixs : -1;
  x.iys = (x.ixs >= 0) ? a->iys : -1;
  x.cap_ix0 = a->cap_ix0; x.cap_iy0 = a->cap_iy0; x.cap_nx = a->cap_nx; x.cap_ny = a->cap_ny;
  x.cap_samples = a->cap_samples;
  const bool lut = a->code && a->mat && a->nmat > 0 && a->metric2d;
  const size_t table = lut ? (size_t)a->nmat * 9 * sizeof(S) : 0;
  size_t lds = table;
  int threads = 64 * (TY + 1);
  if constexpr (FR) {
    // whole rows per tile: Q = Nx / V lanes per row, fr_rows own rows + the halo row
    x.rows = fr_rows;
    x.qrow = g.Nx / V;
    x.ntx = 1;
    x.nty = (g.Ny + fr_rows - 1) / fr_rows;
    threads = ((fr_rows + 1) * x.qrow + 63) / 64 * 64;
    lds = full_rows_lds(g.Nx, V, fr_rows, sizeof(S), table);
    if (fr_rows < 1 || threads > 64 * (TY + 1) || lds > FR_LDS_MAX)
      return fail(WF_EINVAL, "wf_step: full-row tile of %d rows does not fit (%d threads, "
                  "%zu bytes of LDS)", fr_rows, threads, lds);
  } else {
    constexpr int Wc = 64 * V;
    x.rows = TY;
    x.qrow = 64;
    x.ntx = (g.Nx + Wc - 1) / Wc;
    x.nty = (g.Ny + TY - 1) / TY;
  }
  const Fused4Fn<S> kern = pick_fused4<S, TY, V, FR, WIDE>(lut, a->conductive != 0,
                                                      a->jmode == WF_J_LINE, g.m1desc != 0,
                                                      !(lut && a->zero_masked),
                                                      lut && a->code_escape);
  // optional box: launch only the tiles that meet it
  const bool box = a->box_hi[0] > a->box_lo[0] && a->box_hi[1] > a->box_lo[1] &&
                   a->box_hi[2] > a->box_lo[2];
  x.tx0 = x.ty0 = x.tz0 = 0;
  if (box) {
    const int W = FR ? g.Nx : 64 * V, R = FR ? fr_rows : TY;
    const int lx = std::max(a->box_lo[0], 0), hx = std::min(a->box_hi[0], g.Nx);
    const int ly = std::max(a->box_lo[1], 0), hy = std::min(a->box_hi[1], g.Ny);
    if (hx <= lx || hy <= ly || std::max(a->box_lo[2], 0) >= std::min(a->box_hi[2], g.Nz))
      return fail(WF_EINVAL, "wf_step: box outside the grid");
    x.tx0 = lx / W; x.ntx = (hx - 1) / W - x.tx0 + 1;
    x.ty0 = ly / R; x.nty = (hy - 1) / R - x.ty0 + 1;
  }
  const int lz = box ? std::max(a->box_lo[2], 0) : 0, hz = box ? std::min(a->box_hi[2], g.Nz) : g.Nz;
  if (a->kz > 0) {
    x.KZ = a->kz;
  } else {
    const int64_t conc = (int64_t)fused_occupancy(h, (const void *)kern, threads, lds) * h->ncu;
    if (box) {
      // a small box has few tiles: short z chunks spread its march over the CUs
      // (measured: recomputing 16-step segments on a 62-plane box of 11 tile
      // rows took half the time with 2 planes per workgroup than with 8)
      const int64_t kz = ((int64_t)(hz - lz) * x.ntx * x.nty + conc - 1) / conc;
      x.KZ = (int)std::min<int64_t>(64, std::max<int64_t>(2, kz));
    } else {
      x.KZ = auto_kz((int64_t)x.ntx * x.nty, g.Nz, conc);
    }
  }
  x.ntz = (g.Nz + x.KZ - 1) / x.KZ;
  if (box) {
    x.tz0 = lz / x.KZ; x.ntz = (hz - 1) / x.KZ - x.tz0 + 1;
  }
  const int64_t nb = (int64_t)x.ntx * x.nty * x.ntz;
  if (nb > 2147483647LL) return fail(WF_ERANGE, "wf_step: grid too large for the fused path");
  const S dt = (S)a->dt;
  S *Ebuf[2] = {(S *)a->E, (S *)a->E2};
  S *Hbuf[2] = {(S *)a->H, (S *)a->H2};
  wf::Coefs<S> cf;
  cf.a = (const S *)a->a; cf.b = (const S *)a->b;
  cf.ieps = (const S *)a->ieps; cf.sigma = (const S *)a->sigma;
  cf.code = (const uint8_t *)a->code; cf.mat = (const S *)a->mat; cf.nmat = a->nmat;
  if (lut) {
    const S *t = (const S *)a->metric2d;
    const int64_t yz = (int64_t)g.Ny * g.Nz, xz = (int64_t)g.Nx * g.Nz, xy = (int64_t)g.Nx * g.Ny;
    cf.iAx = t; cf.iAy = t + yz; cf.iAz = t + yz + xz;
    cf.itAx = t + yz + xz + xy; cf.itAy = cf.itAx + yz; cf.itAz = cf.itAy + xz;
  }
  int cur = 0;
  for (int s = 0; s < a->nsteps; ++s) {
    const S *row = a->beam_table ? (const S *)a->beam_table + (a->beam_row0 + s) * g.Nz : nullptr;
    x.cap_sample = a->cap_sample0 + s;
    x.skip_j = (s + 1 < a->nsteps) ? 1 : 0;
    S *cap = (S *)a->capture;
    hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(threads), lds, st, g, m, Ebuf[cur],
                       Hbuf[cur], Ebuf[1 - cur], Hbuf[1 - cur], (S *)a->J, cf, dt, x, row, cap);
    WF_LAUNCH_CHECK();
    if (a->traj_E) {
      traj_store<S>(h, a, Ebuf[1 - cur], Hbuf[1 - cur], a->traj_slot0 + s + 1, st);
      WF_LAUNCH_CHECK();
    }
    cur = 1 - cur;
  }
  a->result_swapped = cur;
  return WF_OK;
}

// Cells per lane for a row of Nx cells: the largest of vn, vn/2, ..., 1 that
// divides Nx and leaves the fewest idle lanes in the row's 64-lane segments.
inline int auto_cells_per_lane(int Nx, int vn) {
  int best = 1;
  double beff = -1.0;
  for (int v = vn; v >= 1; v >>= 1) {
    if (Nx % v) continue;
    const int64_t seg = 64 * (int64_t)v;
    const double eff = (double)Nx / (double)(((Nx + seg - 1) / seg) * seg);
    if (eff > beff + 1e-12) { beff = eff; best = v; }
  }
  return best;
}

template <typename S, typename Cm>
int step_fused(wf_handle *h, wf_step_args *a, const wf::MetricDev<S> &m, hipStream_t st) {
  const wf::GridDev &g = h->g;
  // x-vectorised kernel: Nx a multiple of the lane width (16 bytes) and every
  // 3N array within 4 GiB
  const bool idx32v = (uint64_t)3 * (uint64_t)g.N * sizeof(S) < (1ull << 32);
  constexpr int vn = 16 / (int)sizeof(S);  // cells per lane of a 16-byte access
  // rows per workgroup measured best: FP64 3; FP32 8, or 7 when the step is
  // conductive or reads coefficient arrays (8-row tiles then spill registers)
  const bool lut = a->code && a->mat && a->nmat > 0 && a->metric2d;
  const int rdef = sizeof(S) == 8 ? 3 : ((lut && !a->conductive) ? 8 : 7);
  const int rows = a->variant ? a->variant : rdef;
  // whole-row tiles (16-byte lanes) with full_rows = N >= 1 rows per workgroup;
  // row segments otherwise (which shape is faster is measured by the caller:
  // no static rule fitted the measurements)
  if (a->full_rows > 0) {
    if (!idx32v || g.Nx % vn != 0)
      return fail(WF_EINVAL, "wf_step: full-row tiles need Nx a multiple of %d and every 3N "
                  "array within 4 GiB", vn);
    if (a->cells_per_lane != 0 && a->cells_per_lane != vn)
      return fail(WF_EINVAL, "wf_step: full-row tiles use %d cells per lane (cells_per_lane "
                  "%d)", vn, a->cells_per_lane);
    if (a->variant == 1)
      return fail(WF_EINVAL, "wf_step: full-row tiles and the scalar kernel (variant 1) "
                  "exclude each other");
    return step_fused4<S, FR_BOUND, vn, true>(h, a, m, st, a->full_rows);
  }
  // cells per lane: vn for every row count; 1 and 2 (FP32) with the default rows
  int V = a->cells_per_lane;
  if (V == 0) {
    V = vn;
    if (rows == rdef) V = auto_cells_per_lane(g.Nx, vn);
  }
  if (idx32v && a->variant != 1 && g.Nx % V == 0) {
    // variant = rows per workgroup (0: the measured best for the type)
    if (V == vn) {
      switch (rows) {
        case 2: return step_fused4<S, 2, vn>(h, a, m, st);
        case 3: return step_fused4<S, 3, vn>(h, a, m, st);
        case 7: return step_fused4<S, 7, vn>(h, a, m, st);
        case 8:
          if constexpr (sizeof(S) == 4) return step_fused4<S, 8, vn>(h, a, m, st);
          break;
        default: break;
      }
      return fail(WF_EINVAL, "wf_step: variant (rows per workgroup) not built for this dtype");
    }
    if (rows == rdef) {
      if constexpr (vn == 4) {
        if (rows == 8) {
          if (V == 1) return step_fused4<S, 8, 1>(h, a, m, st);
          if (V == 2) return step_fused4<S, 8, 2>(h, a, m, st);
        } else {
          if (V == 1) return step_fused4<S, 7, 1>(h, a, m, st);
          if (V == 2) return step_fused4<S, 7, 2>(h, a, m, st);
        }
      } else {
        if (V == 1) return step_fused4<S, 3, 1>(h, a, m, st);
      }
    }
    return fail(WF_EINVAL, "wf_step: cells_per_lane %d not built with %d rows per workgroup",
                V, rows);
  }
  // 3N arrays past 4 GiB: the same tile code with 64-bit offsets, default row
  // segments (7 rows FP32: the 8-row tile has no registers left for the wider
  // addresses), 16-byte lanes
  if (!idx32v && a->variant == 0 && g.Nx % vn == 0 &&
      (a->cells_per_lane == 0 || a->cells_per_lane == vn) && !(lut && a->code_escape)) {
    if constexpr (sizeof(S) == 8) return step_fused4<S, 3, vn, false, true>(h, a, m, st);
    else return step_fused4<S, 7, vn, false, true>(h, a, m, st);
  }
  if (a->cells_per_lane && a->variant != 1)
    return fail(WF_EINVAL, "wf_step: cells_per_lane %d needs Nx a multiple of it and every "
                "3N array within 4 GiB (or %d with the default rows)", a->cells_per_lane, vn);
  if (lut && a->code_escape)
    return fail(WF_EINVAL, "wf_step: the escape code needs the x-vectorised fused kernel "
                "(variant 0 and every 3N array within 4 GiB)");
  wf::FusedAux x{};
  x.ixs = (a->jmode == WF_J_LINE || a->beam_table || a->conductive) ? a->ixs : -1;
  x.iys = (x.ixs >= 0) ? a->iys : -1;
  x.cap_ix0 = a->cap_ix0; x.cap_iy0 = a->cap_iy0; x.cap_nx = a->cap_nx; x.cap_ny = a->cap_ny;
  x.cap_samples = a->cap_samples;
  x.KZ = a->kz > 0 ? a->kz : 32;
  x.ntx = (g.Nx + FTX - 1) / FTX;
  x.nty = (g.Ny + FTY - 1) / FTY;
  x.ntz = (g.Nz + x.KZ - 1) / x.KZ;
  const int64_t nb = (int64_t)x.ntx * x.nty * x.ntz;
  if (nb > 2147483647LL) return fail(WF_ERANGE, "wf_step: grid too large for the fused path");
  const Cm dt = (Cm)a->dt;
  S *Ebuf[2] = {(S *)a->E, (S *)a->E2};
  S *Hbuf[2] = {(S *)a->H, (S *)a->H2};
  wf::Coefs<S> cf;
  cf.a = (const S *)a->a; cf.b = (const S *)a->b;
  cf.ieps = (const S *)a->ieps; cf.sigma = (const S *)a->sigma;
  cf.code = (const uint8_t *)a->code; cf.mat = (const S *)a->mat; cf.nmat = a->nmat;
  if (lut) {
    const S *t = (const S *)a->metric2d;
    const int64_t yz = (int64_t)g.Ny * g.Nz, xz = (int64_t)g.Nx * g.Nz, xy = (int64_t)g.Nx * g.Ny;
    cf.iAx = t; cf.iAy = t + yz; cf.iAz = t + yz + xz;
    cf.itAx = t + yz + xz + xy; cf.itAy = cf.itAx + yz; cf.itAz = cf.itAy + xz;
  }
  const size_t lds = lut ? (size_t)a->nmat * 9 * sizeof(S) : 0;
  // 32-bit byte offsets when every 3N array fits in 4 GiB
  const bool idx32 = (uint64_t)3 * (uint64_t)g.N * sizeof(S) < (1ull << 32);
  int cur = 0;
  for (int s = 0; s < a->nsteps; ++s) {
    const S *row = a->beam_table ? (const S *)a->beam_table + (a->beam_row0 + s) * g.Nz : nullptr;
    x.cap_sample = a->cap_sample0 + s;
    S *cap = (S *)a->capture;
#define WF_F_LAUNCH3(I, JM, CO, LU, DE)                                                       \
  hipLaunchKernelGGL((wf::k_fused<S, Cm, I, JM, CO, LU, DE, FTX, FTY>), dim3((unsigned)nb),  \
                     dim3(FTX * FTY), lds, st, g, m, Ebuf[cur], Hbuf[cur], Ebuf[1 - cur],     \
                     Hbuf[1 - cur], (S *)a->J, cf, dt, x, row, cap)
#define WF_F_LAUNCH2(I, JM, CO, LU)                                                          \
  if (g.m1desc) WF_F_LAUNCH3(I, JM, CO, LU, true);                                           \
  else WF_F_LAUNCH3(I, JM, CO, LU, false)
#define WF_F_LAUNCH(JM, CO, LU)                                                              \
  do {                                                                                       \
    WF_F_LAUNCH2(int64_t, JM, CO, LU);                                                       \
  } while (0)
    if (lut) {
      if (a->conductive) WF_F_LAUNCH(0, true, true);
      else if (a->jmode == WF_J_LINE) WF_F_LAUNCH(1, false, true);
      else WF_F_LAUNCH(0, false, true);
    } else {
      if (a->conductive) WF_F_LAUNCH(0, true, false);
      else if (a->jmode == WF_J_LINE) WF_F_LAUNCH(1, false, false);
      else WF_F_LAUNCH(0, false, false);
    }
#undef WF_F_LAUNCH
#undef WF_F_LAUNCH2
#undef WF_F_LAUNCH3
    WF_LAUNCH_CHECK();
    if (a->traj_E) {
      traj_store<S>(h, a, Ebuf[1 - cur], Hbuf[1 - cur], a->traj_slot0 + s + 1, st);
      WF_LAUNCH_CHECK();
    }
    cur = 1 - cur;
  }
  a->result_swapped = cur;
  return WF_OK;
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

}  // namespace

extern "C" {

int wf_step(wf_handle *h, wf_step_args *a, void *stream) {
  if (!h || !a) return fail(WF_EINVAL, "wf_step: null argument");
  const int rc = check_step_args(h, a, "wf_step");
  if (rc != WF_OK) return rc;
  a->result_swapped = 0;
  hipStream_t st = (hipStream_t)stream;
  if (a->path == 1) {
    if (h->dtype == WF_F64) return step_fused<double, double>(h, a, h->md, st);
    return step_fused<float, float>(h, a, h->mf, st);
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
               const double *s, int64_t ns, double v, double zmin_v, double t0, double ti,
               double dt, int64_t nt, double qe12, double *WP3d, void *stream) {
  if (!capture || !z || !dz || !s || !WP3d) return fail(WF_EINVAL, "wf_wake_wp: null argument");
  if (nlines < 1 || Nz < 1 || k0 < 0 || nz < 0 || k0 + nz > Nz || ns < 0 || cap_samples < 1)
    return fail(WF_ERANGE, "wf_wake_wp: inconsistent sizes");
  if (nt > cap_samples) return fail(WF_ERANGE, "wf_wake_wp: nt > captured samples");
  if (ns == 0) return WF_OK;
  hipStream_t st = (hipStream_t)stream;
  // the reference's Ezt has len(t) columns; python negative indices wrap there
  const int64_t ncols = nt;
  if (dtype == WF_F64)
    hipLaunchKernelGGL(wf::k_wake_wp<double>, dim3(nblocks(ns, 256)), dim3(256), 0, st,
                       (const double *)capture, nlines, cap_samples, Nz, k0, nz, z, dz, s, ns, v,
                       zmin_v, t0, ti, dt, nt, ncols, qe12, WP3d);
  else if (dtype == WF_F32)
    hipLaunchKernelGGL(wf::k_wake_wp<float>, dim3(nblocks(ns, 256)), dim3(256), 0, st,
                       (const float *)capture, nlines, cap_samples, Nz, k0, nz, z, dz, s, ns, v,
                       zmin_v, t0, ti, dt, nt, ncols, qe12, WP3d);
  else
    return fail(WF_EINVAL, "wf_wake_wp: bad dtype");
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

// ---------------------------------------------------------------------------
// Batches: one launch per step over many grids.
// ---------------------------------------------------------------------------
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
constexpr int BATCH_GRAPH_STEPS = 16;

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

// ---------------------------------------------------------------------------
// Reverse mode (discrete adjoint of the step and of the wake post-processing)
// ---------------------------------------------------------------------------
namespace {

wf::AdjAux adj_aux(const wf_adjoint_args *a) {
  wf::AdjAux x;
  memset(&x, 0, sizeof(x));
  x.conductive = a->conductive;
  x.ixs = a->ixs; x.iys = a->iys;
  for (int d = 0; d < 3; ++d) { x.r0[d] = a->r0[d]; x.r1[d] = a->r1[d]; }
  x.cap_ix0 = a->cap_ix0; x.cap_iy0 = a->cap_iy0; x.cap_nx = a->cap_nx; x.cap_ny = a->cap_ny;
  x.cap_samples = a->cap_samples; x.cap_sample = a->cap_sample;
  x.traj = a->traj;
  return x;
}

constexpr int ADJ_ROWS = 13;  // owned rows per workgroup of the single-pass reverse kernel
                               // (16 waves with the halo rows and column; measured best)

template <typename S, typename A>
wf::AdjStep<S, A> adj_step_params(const wf_handle *h, const wf_adjoint_args *a,
                                  const wf::MetricDev<S> &m, const wf::AdjAux &x,
                                  const wf::AdjTile &t) {
  wf::AdjStep<S, A> p;
  p.g = h->g; p.m = m;
  p.Win = (const A *)a->W; p.lHin = (const A *)a->lH;
  p.Wout = (A *)a->W2; p.lHout = (A *)a->lH2;
  p.lJr = a->lJ_region;
  p.a = (const S *)a->a; p.b = (const S *)a->b;
  p.ieps = (const S *)a->ieps; p.sigma = (const S *)a->sigma; p.kappa = (const S *)a->kappa;
  p.dt = a->dt; p.x = x; p.gcap = a->gcap; p.t = t;
  p.rW = (A *)a->reg_W; p.rJ = a->reg_lJ; p.rslot = a->reg_slot;
  p.code = nullptr; p.mat = nullptr; p.nmat = 0; p.esc = 0;
  p.iAx = p.iAy = p.iAz = p.itAx = p.itAy = p.itAz = nullptr;
  if (a->code && a->mat && a->nmat > 0 && a->code_metric2d) {
    const wf::GridDev &g = h->g;
    const S *t = (const S *)a->code_metric2d;
    const int64_t yz = (int64_t)g.Ny * g.Nz, xz = (int64_t)g.Nx * g.Nz, xy = (int64_t)g.Nx * g.Ny;
    p.code = (const uint8_t *)a->code; p.mat = (const S *)a->mat;
    p.nmat = a->nmat; p.esc = a->code_escape;
    p.iAx = t; p.iAy = t + yz; p.iAz = t + yz + xz;
    p.itAx = t + yz + xz + xy; p.itAy = p.itAx + yz; p.itAz = p.itAy + xz;
  }
  return p;
}

// static LDS of k_adj_fused (P and Q planes in FP64) for TY own rows
constexpr size_t adj_static_lds(int TY) {
  return (size_t)(3 * (TY + 3) + 3 * (TY + 2)) * (wf::ADJ_TX + 3) * sizeof(double);
}

template <typename S, typename A, int TY>
int adjoint_fused(wf_handle *h, const wf_adjoint_args *a, const wf::MetricDev<S> &m,
                  const wf::AdjAux &x, hipStream_t st) {
  const wf::GridDev &g = h->g;
  wf::AdjTile t;
  t.ntx = (g.Nx + wf::ADJ_TX - 1) / wf::ADJ_TX;
  t.nty = (g.Ny + TY - 1) / TY;
  // material codes for FP32 coefficients when the table fits beside the P/Q
  // planes in 64 KiB of LDS; otherwise the coefficient arrays, with identical
  // results (measured at 128x128x256: FP32 88 against 96 us per step; with
  // FP64 coefficients the code path was 1.5x slower than the arrays)
  const size_t table = (a->code && a->mat && a->nmat > 0 && a->code_metric2d)
                           ? (size_t)a->nmat * 9 * sizeof(S) : 0;
  const bool lut = sizeof(S) == 4 && table > 0 && adj_static_lds(TY) + table <= 65536;
  void (*kern)(wf::AdjStep<S, A>) = &wf::k_adj_fused<S, A, TY, false>;
  if constexpr (sizeof(S) == 4) {
    if (lut) kern = &wf::k_adj_fused<S, A, TY, true>;
  }
  const size_t lds = lut ? table : 0;
  const int threads = 64 * (TY + 3);
  const int64_t conc = (int64_t)fused_occupancy(h, (const void *)kern, threads, lds) * h->ncu;
  t.KZ = auto_kz((int64_t)t.ntx * t.nty, g.Nz, conc);
  t.ntz = (g.Nz + t.KZ - 1) / t.KZ;
  const int64_t nbf = (int64_t)t.ntx * t.nty * t.ntz;
  if (nbf > 2147483647LL) return fail(WF_ERANGE, "wf_adjoint_step: grid too large");
  wf::AdjStep<S, A> p = adj_step_params<S, A>(h, a, m, x, t);
  if (!lut) p.code = nullptr;
  hipLaunchKernelGGL(kern, dim3((unsigned)nbf), dim3(threads), lds, st, p);
  WF_LAUNCH_CHECK();
  return WF_OK;
}

template <typename S>
int adjoint_step(wf_handle *h, const wf_adjoint_args *a, const wf::MetricDev<S> &m,
                 hipStream_t st) {
  const wf::GridDev &g = h->g;
  const wf::AdjAux x = adj_aux(a);
  const int64_t nreg = (int64_t)(a->r1[0] - a->r0[0]) * (a->r1[1] - a->r0[1]) *
                       (a->r1[2] - a->r0[2]);
  if (nreg > 0 && a->g_ieps) {
    if (a->fields_f32)
      hipLaunchKernelGGL((wf::k_adj_grad<S, float>), dim3(nblocks(nreg, 256)), dim3(256), 0, st,
                         g, m, (const float *)a->W, a->lJ_region, (const S *)a->E0,
                         (const S *)a->E1, (const S *)a->H1, (const S *)a->beam_row,
                         (const S *)a->sigma, (const S *)a->metric2d, a->dt, x, a->g_ieps,
                         a->g_sigma);
    else
      hipLaunchKernelGGL((wf::k_adj_grad<S, double>), dim3(nblocks(nreg, 256)), dim3(256), 0,
                         st, g, m, (const double *)a->W, a->lJ_region, (const S *)a->E0,
                         (const S *)a->E1, (const S *)a->H1, (const S *)a->beam_row,
                         (const S *)a->sigma, (const S *)a->metric2d, a->dt, x, a->g_ieps,
                         a->g_sigma);
    WF_LAUNCH_CHECK();
  }
  if (a->fused) {
    const int rows = a->fused == 1 ? ADJ_ROWS : a->fused;
    if (a->fields_f32) {
      if (rows == 5) return adjoint_fused<S, float, 5>(h, a, m, x, st);
      if (rows == 13) return adjoint_fused<S, float, 13>(h, a, m, x, st);
    } else {
      if (rows == 5) return adjoint_fused<S, double, 5>(h, a, m, x, st);
      if (rows == 13) return adjoint_fused<S, double, 13>(h, a, m, x, st);
    }
    return fail(WF_EINVAL, "wf_adjoint_step: fused must be 0, 1, 5 or 13 (rows per workgroup)");
  }
  const unsigned nb = nblocks(g.N, 256);
  hipLaunchKernelGGL((wf::k_adj_h<S>), dim3(nb), dim3(256), 0, st, g, m, a->lH,
                     (const double *)a->W, (const S *)a->b, a->dt);
  WF_LAUNCH_CHECK();
  hipLaunchKernelGGL((wf::k_adj_e<S>), dim3(nb), dim3(256), 0, st, g, m, a->W, a->lJ_region,
                     (const double *)a->lH, (const S *)a->a, (const S *)a->ieps,
                     (const S *)a->sigma, a->dt, x, a->gcap);
  WF_LAUNCH_CHECK();
  return WF_OK;
}

// the checks of wf_adjoint_step (also applied to each grid of a reverse batch)
int check_adjoint_args(const wf_handle *h, const wf_adjoint_args *a, const char *who) {
  if (!a->W || !a->lH || !a->a || !a->b || !a->ieps)
    return fail(WF_EINVAL, "%s: null adjoint field or coefficient", who);
  if (a->conductive && !a->sigma)
    return fail(WF_EINVAL, "%s: conductive step needs sigma", who);
  for (int d = 0; d < 3; ++d) {
    const int n = d == 0 ? h->g.Nx : (d == 1 ? h->g.Ny : h->g.Nz);
    if (a->r0[d] < 0 || a->r1[d] > n || a->r0[d] > a->r1[d])
      return fail(WF_ERANGE, "%s: gradient region outside the grid", who);
  }
  const bool region = (a->r1[0] > a->r0[0]) && (a->r1[1] > a->r0[1]) && (a->r1[2] > a->r0[2]);
  if (region && a->g_ieps) {
    if (!a->E0 || !a->E1 || !a->H1 || !a->metric2d)
      return fail(WF_EINVAL, "%s: gradients need E0, E1, H1 and metric2d", who);
    if (a->traj && (a->r0[0] < 1 || a->r0[1] < 1 || a->r0[2] < 1))
      return fail(WF_ERANGE, "%s: trajectory regions start at cell 1 or later", who);
    if (a->conductive && (!a->g_sigma || !a->lJ_region))
      return fail(WF_EINVAL, "%s: conductive gradients need g_sigma and "
                  "lJ_region", who);
  }
  if (a->fields_f32 && !a->fused)
    return fail(WF_EINVAL, "%s: FP32 adjoint fields need the fused kernel", who);
  if (a->kappa && !a->fused)
    return fail(WF_EINVAL, "%s: kappa is read by the fused kernel only", who);
  if (a->reg_W && (!a->fused || !a->reg_lJ || a->g_ieps || a->reg_slot < 0))
    return fail(WF_EINVAL, "%s: reg_W needs the fused kernel, reg_lJ, a slot "
                "and g_ieps = NULL", who);
  if (a->fused) {
    if ((uint64_t)3 * (uint64_t)h->g.N * sizeof(double) >= (1ull << 32))
      return fail(WF_ERANGE, "%s: grid too large for the fused kernel", who);
    if (!a->W2 || !a->lH2 || a->W2 == a->W || a->lH2 == a->lH || a->W2 == a->lH2 ||
        a->W2 == a->lH || a->lH2 == a->W)
      return fail(WF_EINVAL, "%s: fused step needs distinct W2, lH2 buffers", who);
    if (region && (a->r0[0] < 1 || a->r0[1] < 1))
      return fail(WF_ERANGE, "%s: fused step needs the region at cell 1 or "
                  "later in x and y", who);
  }
  if (a->gcap && (a->cap_nx < 1 || a->cap_ny < 1 || a->cap_ix0 < 0 || a->cap_iy0 < 0 ||
                  a->cap_ix0 + a->cap_nx > h->g.Nx || a->cap_iy0 + a->cap_ny > h->g.Ny ||
                  a->cap_sample < 0))
    return fail(WF_ERANGE, "%s: capture window outside the grid", who);
  if (a->reg_W && a->reg_slot >= a->reg_slots)
    return fail(WF_ERANGE, "%s: reg_slot %lld outside %lld slots", who, (long long)a->reg_slot,
                (long long)a->reg_slots);
  return WF_OK;
}


}  // namespace

extern "C" {

int wf_adjoint_step(wf_handle *h, const wf_adjoint_args *a, void *stream) {
  if (!h || !a) return fail(WF_EINVAL, "wf_adjoint_step: null argument");
  const int rc = check_adjoint_args(h, a, "wf_adjoint_step");
  if (rc != WF_OK) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (h->dtype == WF_F64) return adjoint_step<double>(h, a, h->md, st);
  return adjoint_step<float>(h, a, h->mf, st);
}

int wf_adjoint_grad(wf_handle *h, const wf_adjoint_grad_args *a, void *stream) {
  if (!h || !a) return fail(WF_EINVAL, "wf_adjoint_grad: null argument");
  if (!a->E || !a->H || !a->reg_W || !a->reg_lJ || !a->metric2d || !a->g_ieps)
    return fail(WF_EINVAL, "wf_adjoint_grad: null array");
  if (a->conductive && (!a->sigma || !a->g_sigma))
    return fail(WF_EINVAL, "wf_adjoint_grad: conductive gradients need sigma and g_sigma");
  if (a->nsteps < 0 || a->E_stride < 0 || a->H_stride < 0)
    return fail(WF_EINVAL, "wf_adjoint_grad: negative size");
  for (int d = 0; d < 3; ++d) {
    const int n = d == 0 ? h->g.Nx : (d == 1 ? h->g.Ny : h->g.Nz);
    if (a->r0[d] < 0 || a->r1[d] > n || a->r0[d] >= a->r1[d])
      return fail(WF_ERANGE, "wf_adjoint_grad: region outside the grid or empty");
  }
  if (a->traj && (a->r0[0] < 1 || a->r0[1] < 1 || a->r0[2] < 1))
    return fail(WF_ERANGE, "wf_adjoint_grad: trajectory regions start at cell 1 or later");
  if (a->nsteps == 0) return WF_OK;
  wf::AdjAux x;
  memset(&x, 0, sizeof(x));
  x.conductive = a->conductive; x.ixs = a->ixs; x.iys = a->iys; x.traj = a->traj;
  for (int d = 0; d < 3; ++d) { x.r0[d] = a->r0[d]; x.r1[d] = a->r1[d]; }
  const int64_t nreg = (int64_t)(a->r1[0] - a->r0[0]) * (a->r1[1] - a->r0[1]) *
                       (a->r1[2] - a->r0[2]);
  const unsigned nb = nblocks(3 * nreg, 256);
  hipStream_t st = (hipStream_t)stream;
#define WF_GRAD_SEQ(S, A, M)                                                                  \
  hipLaunchKernelGGL((wf::k_adj_grad_seq<S, A>), dim3(nb), dim3(256), 0, st, h->g, M,         \
                     (const S *)a->E, a->E_stride, (const S *)a->H, a->H_stride,              \
                     (const A *)a->reg_W, a->reg_lJ, (const S *)a->beam, (const S *)a->sigma,  \
                     (const S *)a->metric2d, a->dt, x, a->nsteps, a->g_ieps, a->g_sigma)
  if (h->dtype == WF_F64) {
    if (a->fields_f32) WF_GRAD_SEQ(double, float, h->md);
    else WF_GRAD_SEQ(double, double, h->md);
  } else {
    if (a->fields_f32) WF_GRAD_SEQ(float, float, h->mf);
    else WF_GRAD_SEQ(float, double, h->mf);
  }
#undef WF_GRAD_SEQ
  WF_LAUNCH_CHECK();
  return WF_OK;
}

int wf_wake_wp_adjoint(const double *gI, int32_t nlines, int64_t cap_samples, int32_t Nz,
                       int32_t k0, int32_t nz, const double *z, const double *dz, const double *s,
                       int64_t ns, double v, double zmin_v, double t0, double ti, double dt,
                       int64_t nt, double *gcap, void *stream) {
  if (!gI || !z || !dz || !s || !gcap) return fail(WF_EINVAL, "wf_wake_wp_adjoint: null argument");
  if (nlines < 1 || nz < 0 || k0 < 0 || k0 + nz > Nz || nt > cap_samples || dt <= 0)
    return fail(WF_ERANGE, "wf_wake_wp_adjoint: inconsistent sizes");
  const int64_t tot = (int64_t)nlines * nz;
  if (tot == 0) return WF_OK;
  hipLaunchKernelGGL(wf::k_wake_wp_adjoint, dim3(nblocks(tot, 64)), dim3(64), 0,
                     (hipStream_t)stream, gI, nlines, cap_samples, Nz, k0, nz, z, dz, s, ns, v,
                     zmin_v, t0, ti, dt, nt, gcap);
  WF_LAUNCH_CHECK();
  return WF_OK;
}

int wf_idft_real(const double *w, int64_t nbins, int64_t N, int64_t len, double *out,
                 void *stream) {
  if (!w || !out) return fail(WF_EINVAL, "wf_idft_real: null argument");
  if (nbins < 0 || N < 1 || len < 0 || nbins > N) return fail(WF_ERANGE, "wf_idft_real: sizes");
  if (len == 0) return WF_OK;
  hipLaunchKernelGGL(wf::k_idft_real, dim3(nblocks(len, 256)), dim3(256), 0, (hipStream_t)stream,
                     w, nbins, N, len, out);
  WF_LAUNCH_CHECK();
  return WF_OK;
}

}  // extern "C"

extern "C" {

int wf_field_monitor_update(const void *E, int32_t dtype, int64_t N, const double *phase,
                            int32_t nfreq, double *acc, void *stream) {
  if (!E || !phase || !acc) return fail(WF_EINVAL, "wf_field_monitor_update: null argument");
  if (N < 1 || nfreq < 0) return fail(WF_ERANGE, "wf_field_monitor_update: sizes");
  if (dtype != WF_F64 && dtype != WF_F32)
    return fail(WF_EINVAL, "wf_field_monitor_update: dtype must be WF_F64 or WF_F32");
  if (nfreq == 0) return WF_OK;
  const dim3 grid(nblocks(3 * N, 256));
  if (dtype == WF_F64)
    hipLaunchKernelGGL(wf::k_monitor<double>, grid, dim3(256), 0, (hipStream_t)stream,
                       (const double *)E, N, phase, nfreq, acc);
  else
    hipLaunchKernelGGL(wf::k_monitor<float>, grid, dim3(256), 0, (hipStream_t)stream,
                       (const float *)E, N, phase, nfreq, acc);
  WF_LAUNCH_CHECK();
  return WF_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Reverse mode over a batch of grids
// ---------------------------------------------------------------------------
struct wf_batch_adjoint {
  int dtype = 0, f32 = 0, ninst = 0, ntiles = 0, maxsteps = 0;
  void *desc_dev = nullptr;   // AdjBatchDesc<S, A>[ninst]
  int *step_dev = nullptr;    // reverse step read by the batch kernel
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

constexpr int BATCH_ADJ_ROWS = 5;  // rows per workgroup of the batched reverse kernel

template <typename S, typename A>
int batch_adj_build(wf_handle *const *hs, const wf_adjoint_args *as, const int32_t *nt, int n,
                    wf_batch_adjoint *b) {
  std::vector<wf::AdjBatchDesc<S, A>> d(n);
  int64_t tiles = 0;
  for (int i = 0; i < n; ++i) {
    const wf_handle *h = hs[i];
    const wf_adjoint_args *a = &as[i];
    wf::AdjTile t;
    t.ntx = (h->g.Nx + wf::ADJ_TX - 1) / wf::ADJ_TX;
    t.nty = (h->g.Ny + BATCH_ADJ_ROWS - 1) / BATCH_ADJ_ROWS;
    t.KZ = h->g.Nz < 32 ? h->g.Nz : 32;
    t.ntz = (h->g.Nz + t.KZ - 1) / t.KZ;
    const wf::MetricDev<S> *m;
    if constexpr (sizeof(S) == 8) m = &h->md;
    else m = &h->mf;
    d[i].p = adj_step_params<S, A>(h, a, *m, adj_aux(a), t);
    d[i].Nt = nt[i];
    d[i].tile0 = (int)tiles;
    tiles += (int64_t)t.ntx * t.nty * t.ntz;
    b->maxsteps = std::max(b->maxsteps, (int)nt[i]);
  }
  if (tiles > 2147483647LL) return fail(WF_ERANGE, "wf_batch_adjoint_create: too many tiles");
  b->ntiles = (int)tiles;
  WF_HIP(hipMalloc(&b->desc_dev, sizeof(wf::AdjBatchDesc<S, A>) * n));
  WF_HIP(hipMemcpy(b->desc_dev, d.data(), sizeof(wf::AdjBatchDesc<S, A>) * n,
                   hipMemcpyHostToDevice));
  WF_HIP(hipMalloc((void **)&b->step_dev, sizeof(int)));
  return WF_OK;
}

template <typename S, typename A>
int batch_adj_launch(const wf_batch_adjoint *b, hipStream_t st) {
  hipLaunchKernelGGL((wf::k_adj_fused_batch<S, A, BATCH_ADJ_ROWS>), dim3((unsigned)b->ntiles),
                     dim3(64 * (BATCH_ADJ_ROWS + 3)), 0, st,
                     (const wf::AdjBatchDesc<S, A> *)b->desc_dev, b->ninst, b->ntiles,
                     (const int *)b->step_dev);
  WF_LAUNCH_CHECK();
  hipLaunchKernelGGL(wf::k_batch_next, dim3(1), dim3(1), 0, st, b->step_dev);
  WF_LAUNCH_CHECK();
  return WF_OK;
}

int batch_adj_step_once(const wf_batch_adjoint *b, hipStream_t st) {
  if (b->dtype == WF_F64)
    return b->f32 ? batch_adj_launch<double, float>(b, st) : batch_adj_launch<double, double>(b, st);
  return b->f32 ? batch_adj_launch<float, float>(b, st) : batch_adj_launch<float, double>(b, st);
}

}  // namespace

extern "C" {

int wf_batch_adjoint_create(wf_handle *const *handles, const wf_adjoint_args *args,
                            const int32_t *nsteps, int32_t n, wf_batch_adjoint **out) {
  if (!handles || !args || !nsteps || !out)
    return fail(WF_EINVAL, "wf_batch_adjoint_create: null argument");
  *out = nullptr;
  if (n < 1) return fail(WF_EINVAL, "wf_batch_adjoint_create: empty batch");
  if (!handles[0]) return fail(WF_EINVAL, "wf_batch_adjoint_create: null handle");
  const int dtype = handles[0]->dtype, f32 = args[0].fields_f32;
  for (int i = 0; i < n; ++i) {
    const wf_handle *h = handles[i];
    const wf_adjoint_args *a = &args[i];
    char who[64];
    snprintf(who, sizeof(who), "wf_batch_adjoint_create[%d]", i);
    if (!h) return fail(WF_EINVAL, "%s: null handle", who);
    const int rc = check_adjoint_args(h, a, who);
    if (rc != WF_OK) return rc;
    if (!a->fused || !a->reg_W || a->g_ieps)
      return fail(WF_EINVAL, "%s: batch items use the fused kernel with region slots "
                  "(fused, reg_W, reg_lJ, g_ieps = NULL)", who);
    if (nsteps[i] < 0 || nsteps[i] > a->reg_slots)
      return fail(WF_ERANGE, "%s: %d steps for %lld region slots", who, (int)nsteps[i],
                  (long long)a->reg_slots);
    if (a->gcap && nsteps[i] > a->cap_samples)
      return fail(WF_ERANGE, "%s: %d steps for %lld capture samples", who, (int)nsteps[i],
                  (long long)a->cap_samples);
    if (h->dtype != dtype || a->fields_f32 != f32)
      return fail(WF_EINVAL, "%s: dtype and fields_f32 must match across the batch", who);
  }
  wf_batch_adjoint *b = new wf_batch_adjoint();
  b->dtype = dtype; b->f32 = f32; b->ninst = n;
  int rc;
  if (dtype == WF_F64)
    rc = f32 ? batch_adj_build<double, float>(handles, args, nsteps, n, b)
             : batch_adj_build<double, double>(handles, args, nsteps, n, b);
  else
    rc = f32 ? batch_adj_build<float, float>(handles, args, nsteps, n, b)
             : batch_adj_build<float, double>(handles, args, nsteps, n, b);
  if (rc != WF_OK) {
    wf_batch_adjoint_destroy(b);
    return rc;
  }
  *out = b;
  return WF_OK;
}

int wf_batch_adjoint_run(wf_batch_adjoint *b, int32_t r0, int32_t nr, int32_t use_graph,
                         void *stream) {
  if (!b) return fail(WF_EINVAL, "wf_batch_adjoint_run: null batch");
  if (r0 < 0 || nr < 0) return fail(WF_EINVAL, "wf_batch_adjoint_run: negative step range");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(wf::k_batch_set, dim3(1), dim3(1), 0, st, b->step_dev, (int)r0);
  WF_LAUNCH_CHECK();
  int done = 0;
  if (use_graph && nr >= BATCH_GRAPH_STEPS) {
    if (!b->exec) {
      WF_HIP(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
      WF_HIP(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
      int rc = WF_OK;
      for (int s = 0; s < BATCH_GRAPH_STEPS && rc == WF_OK; ++s)
        rc = batch_adj_step_once(b, b->cap_stream);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
      if (rc != WF_OK) return rc;
      if (e != hipSuccess)
        return fail(WF_EHIP, "wf_batch_adjoint_run: capture: %s", hipGetErrorString(e));
      b->graph = g;
      WF_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    for (; done + BATCH_GRAPH_STEPS <= nr; done += BATCH_GRAPH_STEPS)
      WF_HIP(hipGraphLaunch(b->exec, st));
  }
  for (; done < nr; ++done) {
    const int rc = batch_adj_step_once(b, st);
    if (rc != WF_OK) return rc;
  }
  return WF_OK;
}

int wf_batch_adjoint_destroy(wf_batch_adjoint *b) {
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
