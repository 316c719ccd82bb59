// The fused single-pass step kernels' launch logic (path 1), included by the
// FP32 and FP64 translation units.
#pragma once
#include "lib_internal.h"
#include "step_fused.h"
#include "step_fused4.h"

// Every symbol below stays inside the library.
#pragma GCC visibility push(hidden)

namespace wff {

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

template <typename S, int TY, int V, bool FR, bool WIDE>
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
    if (a->mon_acc) {
      monitor_store<S>(h, a, Ebuf[1 - cur], s, st);
      WF_LAUNCH_CHECK();
    }
    cur = 1 - cur;
  }
  a->result_swapped = cur;
  return WF_OK;
}

// Instantiations of step_fused4 <S, rows, cells per lane, whole rows, 64-bit
// offsets>, spread over translation units that compile in parallel; the
// dispatcher's translation unit sees them as extern templates.
#define WF_FUSED4_F32_A(X) X(float, 8, 4, false, false) X(float, 8, 1, false, false) \
  X(float, 8, 2, false, false)
#define WF_FUSED4_F32_B(X) X(float, 7, 4, false, false) X(float, 7, 1, false, false) \
  X(float, 7, 2, false, false)
#define WF_FUSED4_F32_C(X) X(float, 2, 4, false, false) X(float, 3, 4, false, false)
#define WF_FUSED4_F32_D(X) X(float, 7, 4, true, false) X(float, 7, 4, false, true)
#define WF_FUSED4_F64_A(X) X(double, 3, 2, false, false) X(double, 3, 1, false, false) \
  X(double, 3, 2, false, true)
#define WF_FUSED4_F64_B(X) X(double, 2, 2, false, false) X(double, 7, 2, false, false) \
  X(double, 7, 2, true, false)

#define WF_FUSED4_SIG(S, TY, V, FR, WIDE)                                              \
  int step_fused4<S, TY, V, FR, WIDE>(wf_handle *, wf_step_args *, const wf::MetricDev<S> &, \
                                      hipStream_t, int);
#define WF_FUSED4_EXTERN(S, TY, V, FR, WIDE) extern template WF_FUSED4_SIG(S, TY, V, FR, WIDE)
#define WF_FUSED4_INSTANTIATE(S, TY, V, FR, WIDE) template WF_FUSED4_SIG(S, TY, V, FR, WIDE)

}  // namespace wff

#pragma GCC visibility pop
