// Two leapfrog steps per launch (k_fused4_t2): the applicability rule and the
// launchers (FP32: 8 or 4 rows per workgroup; FP64: 2 rows, two waves each).
#include "lib_fused.h"
#include "step_fused_t2.h"

#pragma GCC visibility push(hidden)

int fused_t2_applies(const wf_handle *h, const wf_step_args *a) {
  const wf::GridDev &g = h->g;
  const bool lut = a->code && a->mat && a->nmat > 0 && a->metric2d;
  const bool box = a->box_hi[0] > a->box_lo[0] && a->box_hi[1] > a->box_lo[1] &&
                   a->box_hi[2] > a->box_lo[2];
  // the wakefield step: a beam line fed from the beam table (with conduction,
  // J = sigma E elsewhere)
  const bool jok = a->beam_table && (a->conductive ? a->jmode == WF_J_FULL
                                                   : a->jmode == WF_J_LINE);
  // the default (1) starts at 2^20 cells: below that the few workgroups of a
  // launch leave the machine idle and one step per launch is faster (64^3:
  // 28.5 against 17.1 us per step; 96x96x160: 32.7 against 37.1). FP64 rows
  // take two waves, the second one idle up to 128 cells: there the default
  // needs rows of 224 cells or more (256x256x128: 134 against 223 us per
  // step; 160x160x320 and the 200-cell lossy pillbox: one step per launch is
  // as fast or faster)
  const bool big = a->two_step != 1 ||
                   (g.N >= (int64_t)1 << 20 && (h->dtype == WF_F32 || g.Nx >= 224));
  const int elem = h->dtype == WF_F64 ? 8 : 4, vn = 16 / elem;  // cells per lane
  return (h->dtype == WF_F32 || h->dtype == WF_F64) && a->two_step != 0 && big &&
         a->nsteps >= 2 && lut && a->zero_masked && !a->code_escape && !a->compute_f64 && jok &&
         g.Nx <= 256 && g.Nx % vn == 0 &&
         (uint64_t)3 * (uint64_t)g.N * (uint64_t)elem < (1ull << 32) && a->variant == 0 &&
         a->full_rows <= 0 && (a->cells_per_lane == 0 || a->cells_per_lane == vn) && !box &&
         !a->traj_E && !a->mon_acc;
}

namespace wff {
namespace {

template <typename S>
using T2Fn = void (*)(wf::GridDev, wf::MetricDev<S>, const S *, const S *, S *, S *, S *,
                      wf::Coefs<S>, S, wf::FusedAux, const S *, const S *, S *);

template <typename S, int TY>
T2Fn<S> pick_t2(bool cond, bool desc) {
  if (cond) return desc ? &wf::k_fused4_t2<S, true, true, TY> : &wf::k_fused4_t2<S, true, false, TY>;
  return desc ? &wf::k_fused4_t2<S, false, true, TY> : &wf::k_fused4_t2<S, false, false, TY>;
}

template <typename S, int TY>
int launch_t2(wf_handle *h, wf_step_args *a, const wf::MetricDev<S> &m, hipStream_t st) {
  const wf::GridDev &g = h->g;
  wf::FusedAux x{};
  x.ixs = a->ixs;
  x.iys = a->iys;
  x.cap_ix0 = a->cap_ix0; x.cap_iy0 = a->cap_iy0; x.cap_nx = a->cap_nx; x.cap_ny = a->cap_ny;
  x.cap_samples = a->cap_samples;
  x.ntx = 1;
  x.nty = (g.Ny + TY - 1) / TY;
  x.rows = TY;
  x.qrow = 64;
  const int threads = 64 * (TY + 4) * (int)(sizeof(S) / 4);
  const T2Fn<S> kern = pick_t2<S, TY>(a->conductive != 0, g.m1desc != 0);
  // dynamic LDS: the material table, then per plane of a chunk (KZ + 3 planes)
  // Lz, tLz, two beam values and iAx, itAx of the TY + 4 rows
  hipFuncAttributes fa{};
  int dev = 0, lds_dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipFuncGetAttributes(&fa, (const void *)kern) != hipSuccess ||
      hipDeviceGetAttribute(&lds_dev, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    return fail(WF_EHIP, "wf_step: two-step kernel attributes");
  const size_t mat = (size_t)a->nmat * 9 * sizeof(S);
  const size_t per_plane = (size_t)(4 + 2 * (TY + 4)) * sizeof(S);
  const int64_t room = (int64_t)lds_dev - (int64_t)fa.sharedSizeBytes - (int64_t)mat;
  const int kz_lds = (int)(room / (int64_t)per_plane) - 3;
  if (kz_lds < 8) return fail(WF_ERANGE, "wf_step: too many materials for the two-step kernel");
  size_t lds = mat + per_plane * (size_t)(std::min(64, kz_lds) + 3);  // occupancy query
  if (a->kz > 0) {
    x.KZ = a->kz;
  } else {
    // each z chunk recomputes two warm-up planes, so the fewest chunks that
    // still fill the machine in one round: at 256x256x512, 64 planes (256
    // workgroups, 0.2444 ms per step) against 32 (two rounds, 0.2477)
    // (as many z chunks per row tile as fit beside the other tiles in one
    // round: at 160x160x320, 20 tiles x 12 chunks = 240 workgroups, where
    // chunks of ceil(Nz * tiles / CUs) planes made 260, a second round for 4)
    const int64_t conc = (int64_t)fused_occupancy(h, (const void *)kern, threads, lds) * h->ncu;
    const int64_t ntz1 = conc / std::max<int64_t>(x.nty, 1);
    const int64_t kz1 = ntz1 > 0 ? ((int64_t)g.Nz + ntz1 - 1) / ntz1 : (int64_t)g.Nz;
    x.KZ = (kz1 <= 64) ? (int)std::max<int64_t>(kz1, 8) : auto_kz(x.nty, g.Nz, conc);
  }
  x.KZ = std::min(x.KZ, kz_lds);
  lds = mat + per_plane * (size_t)(x.KZ + 3);
  x.ntz = (g.Nz + x.KZ - 1) / x.KZ;
  const int64_t nb = (int64_t)x.nty * x.ntz;
  if (nb > 2147483647LL) return fail(WF_ERANGE, "wf_step: grid too large for the fused path");
  wf::Coefs<S> cf;
  cf.a = (const S *)a->a; cf.b = (const S *)a->b;
  cf.ieps = (const S *)a->ieps; cf.sigma = (const S *)a->sigma;
  cf.code = (const uint8_t *)a->code; cf.mat = (const S *)a->mat; cf.nmat = a->nmat;
  {
    const S *t = (const S *)a->metric2d;
    const int64_t yz = (int64_t)g.Ny * g.Nz, xz = (int64_t)g.Nx * g.Nz, xy = (int64_t)g.Nx * g.Ny;
    cf.iAx = t; cf.iAy = t + yz; cf.iAz = t + yz + xz;
    cf.itAx = t + yz + xz + xy; cf.itAy = cf.itAx + yz; cf.itAz = cf.itAy + xz;
  }
  const S dt = (S)a->dt;
  S *Ebuf[2] = {(S *)a->E, (S *)a->E2};
  S *Hbuf[2] = {(S *)a->H, (S *)a->H2};
  const S *table = (const S *)a->beam_table;
  int cur = 0, s = 0;
  for (; s + 1 < a->nsteps; s += 2) {
    const S *row1 = table ? table + (a->beam_row0 + s) * g.Nz : nullptr;
    const S *row2 = table ? table + (a->beam_row0 + s + 1) * g.Nz : nullptr;
    x.cap_sample = a->cap_sample0 + s;
    x.skip_j = (s + 2 < a->nsteps) ? 1 : 0;  // J = sigma E stored by the last launch
    hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(threads), lds, st, g, m, Ebuf[cur],
                       Hbuf[cur], Ebuf[1 - cur], Hbuf[1 - cur], (S *)a->J, cf, dt, x, row1, row2,
                       (S *)a->capture);
    WF_LAUNCH_CHECK();
    cur = 1 - cur;
  }
  if (s < a->nsteps) {  // an odd count ends with one single step on the current buffers
    wf_step_args b = *a;
    b.two_step = 0;
    b.nsteps = 1;
    b.beam_row0 = a->beam_row0 + s;
    b.cap_sample0 = a->cap_sample0 + s;
    if (cur) {
      b.E = a->E2; b.E2 = a->E;
      b.H = a->H2; b.H2 = a->H;
    }
    const int rc = sizeof(S) == 4 ? step_fused_f32(h, &b, st) : step_fused_f64(h, &b, st);
    if (rc != WF_OK) return rc;
    if (b.result_swapped) cur = 1 - cur;
  }
  a->result_swapped = cur;
  a->two_step_used = 1;
  return WF_OK;
}

}  // namespace
}  // namespace wff

// rows per workgroup: 8 by default (12 waves, 168 VGPRs: 0.249 ms per step at
// 256x256x512 against 0.319 with 4 rows in 8 waves), 4 on request
int step_fused_t2_f32(wf_handle *h, wf_step_args *a, hipStream_t st) {
  if (a->two_step == 4) return wff::launch_t2<float, 4>(h, a, h->mf, st);
  return wff::launch_t2<float, 8>(h, a, h->mf, st);
}

// FP64: 2 rows per workgroup (12 waves, two per row: the 2 KB rows of six
// staged rows fill 112 KB of LDS)
int step_fused_t2_f64(wf_handle *h, wf_step_args *a, hipStream_t st) {
  return wff::launch_t2<double, 2>(h, a, h->md, st);
}

#pragma GCC visibility pop
