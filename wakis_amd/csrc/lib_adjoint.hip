// Reverse mode: the discrete adjoint of the step and of the wake
// post-processing (wf_adjoint_*, wf_wake_wp_adjoint, wf_idft_real), and its
// batched form (wf_batch_adjoint_*).
#include "lib_internal.h"
#include "adjoint_fused.h"

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

// owned rows per workgroup of the single-pass reverse kernel, measured at
// 128x128x256 (the same gradients for every count): FP32 arithmetic 9 rows (12
// waves with the halo rows and column, two workgroups per CU: 46.4 us per
// step against 52.3 with 7, 53.9 with 13, 56.7 with 5); FP64 13 (161.0 us; 7
// and 5 the same, 9 168.0)
constexpr int ADJ_ROWS_F32 = 9, ADJ_ROWS_F64 = 13;

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
template <typename C>
constexpr size_t adj_static_lds(int TY) {
  return (size_t)(3 * (TY + 3) + 3 * (TY + 2)) * (wf::ADJ_TX + 3) * sizeof(C);
}

template <typename S, typename A, int TY, typename C>
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
  // the material table: 12 values per material (ieps, imu, kappa or sigma)
  const size_t table = (a->code && a->mat && a->nmat > 0 && a->code_metric2d)
                           ? (size_t)a->nmat * 12 * sizeof(S) : 0;
  const bool lut = sizeof(S) == 4 && table > 0 && adj_static_lds<C>(TY) + table <= 65536;
  void (*kern)(wf::AdjStep<S, A>) = &wf::k_adj_fused<S, A, TY, false, C>;
  if constexpr (sizeof(S) == 4) {
    if (lut) kern = &wf::k_adj_fused<S, A, TY, true, C>;
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
    const int rows =
        a->fused != 1 ? a->fused : ((a->fields_f32 && a->compute_f32) ? ADJ_ROWS_F32 : ADJ_ROWS_F64);
    if (a->fields_f32 && a->compute_f32) {
      if (rows == 5) return adjoint_fused<S, float, 5, float>(h, a, m, x, st);
      if (rows == 7) return adjoint_fused<S, float, 7, float>(h, a, m, x, st);
      if (rows == 9) return adjoint_fused<S, float, 9, float>(h, a, m, x, st);
      if (rows == 13) return adjoint_fused<S, float, 13, float>(h, a, m, x, st);
    } else if (a->fields_f32) {
      if (rows == 5) return adjoint_fused<S, float, 5, double>(h, a, m, x, st);
      if (rows == 7) return adjoint_fused<S, float, 7, double>(h, a, m, x, st);
      if (rows == 9) return adjoint_fused<S, float, 9, double>(h, a, m, x, st);
      if (rows == 13) return adjoint_fused<S, float, 13, double>(h, a, m, x, st);
    } else {
      if (rows == 5) return adjoint_fused<S, double, 5, double>(h, a, m, x, st);
      if (rows == 7) return adjoint_fused<S, double, 7, double>(h, a, m, x, st);
      if (rows == 9) return adjoint_fused<S, double, 9, double>(h, a, m, x, st);
      if (rows == 13) return adjoint_fused<S, double, 13, double>(h, a, m, x, st);
    }
    return fail(WF_EINVAL, "wf_adjoint_step: fused must be 0, 1, 5, 7, 9 or 13 (rows per workgroup)");
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
  if (a->compute_f32 && !a->fields_f32)
    return fail(WF_EINVAL, "%s: FP32 arithmetic needs FP32 adjoint fields", who);
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
                       int64_t ns, double v, double zref_v, double t0, double ti, double dt,
                       int64_t nt, int32_t counter_moving, double *gcap, void *stream) {
  if (!gI || !z || !dz || !s || !gcap) return fail(WF_EINVAL, "wf_wake_wp_adjoint: null argument");
  if (nlines < 1 || nz < 0 || k0 < 0 || k0 + nz > Nz || nt > cap_samples || dt <= 0)
    return fail(WF_ERANGE, "wf_wake_wp_adjoint: inconsistent sizes");
  const int64_t tot = (int64_t)nlines * nz;
  if (tot == 0) return WF_OK;
  if (counter_moving)
    hipLaunchKernelGGL(wf::k_wake_wp_adjoint<true>, dim3(nblocks(tot, 64)), dim3(64), 0,
                       (hipStream_t)stream, gI, nlines, cap_samples, Nz, k0, nz, z, dz, s, ns, v,
                       zref_v, t0, ti, dt, nt, gcap);
  else
    hipLaunchKernelGGL(wf::k_wake_wp_adjoint<false>, dim3(nblocks(tot, 64)), dim3(64), 0,
                       (hipStream_t)stream, gI, nlines, cap_samples, Nz, k0, nz, z, dz, s, ns, v,
                       zref_v, t0, ti, dt, nt, gcap);
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

// ---------------------------------------------------------------------------
// Reverse mode over a batch of grids
// ---------------------------------------------------------------------------
struct wf_batch_adjoint {
  int dtype = 0, f32 = 0, c32 = 0, ninst = 0, ntiles = 0, maxsteps = 0;
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

template <typename S, typename A, typename C = double>
int batch_adj_launch(const wf_batch_adjoint *b, hipStream_t st) {
  hipLaunchKernelGGL((wf::k_adj_fused_batch<S, A, BATCH_ADJ_ROWS, C>), dim3((unsigned)b->ntiles),
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
    return b->c32 ? batch_adj_launch<double, float, float>(b, st)
                  : (b->f32 ? batch_adj_launch<double, float>(b, st)
                            : batch_adj_launch<double, double>(b, st));
  return b->c32 ? batch_adj_launch<float, float, float>(b, st)
                : (b->f32 ? batch_adj_launch<float, float>(b, st)
                          : batch_adj_launch<float, double>(b, st));
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
  const int dtype = handles[0]->dtype, f32 = args[0].fields_f32, c32 = args[0].compute_f32;
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
    if (h->dtype != dtype || a->fields_f32 != f32 || a->compute_f32 != c32)
      return fail(WF_EINVAL, "%s: dtype, fields_f32 and compute_f32 must match across the "
                  "batch", who);
  }
  wf_batch_adjoint *b = new wf_batch_adjoint();
  b->dtype = dtype; b->f32 = f32; b->c32 = c32; b->ninst = n;
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
