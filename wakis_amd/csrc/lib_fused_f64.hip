// Fused step (path 1) with FP64 storage: the dispatcher over row counts,
// cells per lane, whole-row tiles and 64-bit offsets. The step_fused4
// instantiations are compiled in lib_fused_f64_*.hip.
#include "lib_fused.h"

#pragma GCC visibility push(hidden)
namespace wff {

WF_FUSED4_F64_A(WF_FUSED4_EXTERN)
WF_FUSED4_F64_B(WF_FUSED4_EXTERN)

namespace {

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
  if (a->two_step && fused_t2_applies(h, a)) return step_fused_t2_f64(h, a, st);
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
    return step_fused4<S, FR_BOUND, vn, true, false>(h, a, m, st, a->full_rows);
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
        case 2: return step_fused4<S, 2, vn, false, false>(h, a, m, st);
        case 3: return step_fused4<S, 3, vn, false, false>(h, a, m, st);
        case 7: return step_fused4<S, 7, vn, false, false>(h, a, m, st);
        case 8:
          if constexpr (sizeof(S) == 4) return step_fused4<S, 8, vn, false, false>(h, a, m, st);
          break;
        default: break;
      }
      return fail(WF_EINVAL, "wf_step: variant (rows per workgroup) not built for this dtype");
    }
    if (rows == rdef) {
      if constexpr (vn == 4) {
        if (rows == 8) {
          if (V == 1) return step_fused4<S, 8, 1, false, false>(h, a, m, st);
          if (V == 2) return step_fused4<S, 8, 2, false, false>(h, a, m, st);
        } else {
          if (V == 1) return step_fused4<S, 7, 1, false, false>(h, a, m, st);
          if (V == 2) return step_fused4<S, 7, 2, false, false>(h, a, m, st);
        }
      } else {
        if (V == 1) return step_fused4<S, 3, 1, false, false>(h, a, m, st);
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
    if (a->mon_acc) {
      monitor_store<S>(h, a, Ebuf[1 - cur], s, st);
      WF_LAUNCH_CHECK();
    }
    cur = 1 - cur;
  }
  a->result_swapped = cur;
  return WF_OK;
}

}  // namespace
}  // namespace wff
#pragma GCC visibility pop

int step_fused_f64(wf_handle *h, wf_step_args *a, hipStream_t st) {
  return wff::step_fused<double, double>(h, a, h->md, st);
}
