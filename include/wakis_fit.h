/*
 * wakis_fit.h - C ABI of the MI355X FIT time-stepping engine (libwakis_fit.so).
 *
 * Plain C: device pointers are void*, streams are void* (a hipStream_t, or
 * NULL for the default stream). Every entry point returns 0 on success and a
 * negative WF_E* code on failure; wf_last_error() gives the message of the
 * last failure on the calling thread. No entry point allocates device memory
 * except wf_create (small metric vectors and a 32 KB reduction workspace); none
 * synchronises the stream except
 * where stated.
 *
 * Field layout (identical to the reference Field, wakis/field.py:19-24, 57-91):
 * a flat array of 3N values, x|y|z component blocks, each block x-fastest,
 * n = i + Nx*j + Nx*Ny*k. All coefficient arrays use the same layout.
 *
 * Reference interfaces each entry point replaces are cited per function
 * (paths relative to the Wakis v0.7.0 tree).
 */
#ifndef WAKIS_FIT_H
#define WAKIS_FIT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WF_OK 0
#define WF_EINVAL (-1)   /* bad argument (maps to ValueError) */
#define WF_ERANGE (-2)   /* index/size out of range (maps to IndexError) */
#define WF_EHIP (-3)     /* HIP runtime error (maps to RuntimeError) */
#define WF_ESTATE (-4)   /* call not valid in this state (maps to RuntimeError) */

#define WF_F64 0
#define WF_F32 1

/* Face bit numbering used by the mask words: xlo, ylo, zlo, xhi, yhi, zhi. */
#define WF_FACE_XLO 0
#define WF_FACE_YLO 1
#define WF_FACE_ZLO 2
#define WF_FACE_XHI 3
#define WF_FACE_YHI 4
#define WF_FACE_ZHI 5

/* J-term modes of the E update (solverFIT3D.py:492-499, "- ieps*J"). */
#define WF_J_NONE 0  /* J identically zero: E += dt*(M2 H)                        */
#define WF_J_LINE 1  /* J nonzero only on the beam line (ixs, iys, :, z)          */
#define WF_J_FULL 2  /* J read everywhere                                         */

typedef struct wf_handle wf_handle;

/* Grid and boundary description.
 * Replaces the operator assembly of SolverFIT3D.__init__
 * (solverFIT3D.py:203-227, 312-322) and BCsMixin._apply_bc_to_C
 * (boundaries.py:14-188): instead of the 3N x 3N CSR matrices the engine keeps
 * 1-D metric vectors and evaluates the curl stencil with the same entries
 * +-fl(a_r*L_c) / +-fl(b_r*tL_c) and the same per-row summation order. */
typedef struct wf_grid {
  int32_t Nx, Ny, Nz;
  int32_t dtype;           /* WF_F64 or WF_F32: storage type of fields/coefficients   */
  int32_t pec_zero_faces;  /* bit f: face f zeroes tangential E columns (PEC/pml)    */
  int32_t pmc_zero_faces;  /* bit f: face f zeroes tangential H rows (0 if no PMC)   */
  int32_t m1_descending;   /* 1: H-row sums run in descending column order          */
  const double *Lx, *Ly, *Lz;    /* host, primal edge lengths (Nx, Ny, Nz values)    */
  const double *tLx, *tLy, *tLz; /* host, dual edge lengths after periodic/ABC edits */
} wf_grid;

/* One leapfrog batch. Replaces SolverFIT3D._one_step (solverFIT3D.py:482-503)
 * called nsteps times, plus Beam.update's J line write (sources.py:108-110) and
 * wakesolve's per-step Ez window save (routines.py:419-425, 440).
 * E/H/J are updated in place. */
typedef struct wf_step_args {
  void *E, *H, *J;        /* device 3N fields                                        */
  const void *a;          /* device 3N: fl(iA*imu), PMC row mask folded in            */
  const void *b;          /* device 3N: fl(ieps*itA), PEC row mask folded in          */
  const void *ieps;       /* device 3N (read when jmode != WF_J_NONE)                 */
  const void *sigma;      /* device 3N (read when conductive)                         */
  double dt;
  int32_t nsteps;
  int32_t jmode;          /* WF_J_*                                                  */
  int32_t conductive;     /* 1: J <- sigma*E after each E update                     */
  int32_t ixs, iys;       /* beam line (for WF_J_LINE and the beam table)            */
  const void *beam_table; /* device [rows][Nz] in the field dtype, or NULL           */
  int64_t beam_row0;      /* table row used by the first step of this batch          */
  int64_t beam_rows;      /* rows in the table                                       */
  void *capture;          /* device [nlines][cap_samples][Nz] field dtype, or NULL   */
  int32_t cap_ix0, cap_iy0, cap_nx, cap_ny; /* Ez window origin and extent            */
  int64_t cap_samples;    /* samples per line in the buffer                          */
  int64_t cap_sample0;    /* sample written by the first step of this batch          */
  int32_t path;           /* 0: exact general two-kernel path; 1: fused fast path    */
  int32_t compute_f64;    /* 1 with WF_F32 storage: dt products and updates in FP64
                             (numpy promotion when dt is a float64 scalar)           */
  void *E2, *H2;          /* fused path: second E/H buffers (ping-pong), 3N each    */
  int32_t kz;             /* fused path: z-planes per workgroup (0 = automatic)      */
  int32_t result_swapped; /* out: 1 if the fields of the last step are in E2/H2      */
  /* Fused path, optional compressed coefficients: instead of reading a/b
   * (and ieps/sigma) per cell, read a uint8 material code per cell and
   * recompute a = fl(iA*imu), b = fl(ieps*itA) from a material table and
   * 2-D metric tables (must reproduce the a/b arrays exactly). */
  const void *code;       /* device uint8[N]                                          */
  const void *mat;        /* device [nmat][9]: imu_xyz, ieps_xyz, sigma_xyz (dtype)   */
  const void *metric2d;   /* device, dtype: iA_x[Ny*Nz] (j+Ny*k), iA_y[Nx*Nz] (i+Nx*k),
                             iA_z[Nx*Ny] (i+Nx*j), then itA_x, itA_y, itA_z likewise
                             (itA after the periodic/ABC copies)                     */
  int32_t nmat;           /* 1..256                                                   */
  int32_t variant;        /* fused kernel choice: 0 auto (x-vectorised kernel with the
                             row count measured fastest: FP64 3; FP32 8, or 7 when
                             conductive or without material codes), 1 scalar kernel,
                             2, 3, 7 or 8 (FP32 only): x-vectorised kernel with
                             that many rows per workgroup. When a 3N array passes
                             4 GiB, 0 takes the x-vectorised kernel with 64-bit
                             offsets (FP64 3 rows, FP32 7; Nx a multiple of
                             16/sizeof(dtype)), and 2..8 (or another Nx) the
                             scalar kernel                                         */
  int32_t zero_masked;    /* fused path: 1 = E is zero on every edge the PEC mask
                             zeroes and H on every face the PMC mask zeroes, no
                             source writes there, and (with `code`) metric2d has
                             the masks folded in (iA_c = 0 on PMC-masked faces,
                             itA_c = 0 on PEC-masked edges). The x-vectorised
                             kernel then applies no masks; results are identical */
  /* Optional region trajectory (reverse mode): after step s of this call,
   * E on the cells [traj_lo, traj_hi) and H on [traj_lo - 1, traj_hi), all
   * three components (box order: x fastest, then y, z, component), are
   * copied to slot traj_slot0 + s + 1 of traj_E / traj_H, device buffers of
   * [traj_slots][3 * box cells] in the field dtype. traj_lo >= 1 on every
   * axis. NULL traj_E: no trajectory. Not available in batches. */
  void *traj_E, *traj_H;
  int32_t traj_lo[3], traj_hi[3];
  int64_t traj_slots, traj_slot0;
  int32_t cells_per_lane; /* fused x-vectorised kernel: cells per lane, 0 = automatic
                             (16 bytes per lane, fewer when the row of Nx cells would
                             leave lanes idle), 1, 2 (FP32) or 16/sizeof(dtype); 1 and
                             2 are built with the automatic row count only. Results
                             are identical for every choice                        */
  int32_t full_rows;      /* fused x-vectorised kernel, 16-byte lanes: N >= 1 =
                             workgroups of N whole rows (thread t: cells
                             V*(t % (Nx/V)) of tile row t / (Nx/V)) instead of
                             64-lane row segments (0 or negative). Pays when Nx is
                             not close to a multiple of the segment (Nx = 200:
                             1.25-1.4x); which is faster is measured, not derived.
                             Needs Nx a multiple of 16/sizeof(dtype), at most 512
                             threads and 64 KiB of LDS per workgroup. In a batch
                             (wf_batch_create), N >= 1 on every item = whole-row
                             tiles with the rows per grid that fill the batch's
                             256-thread workgroup. Results are identical for every
                             choice                                                */
  int32_t box_lo[3], box_hi[3]; /* optional (box_hi > box_lo on every axis): the
                             x-vectorised fused kernel launches only the tiles that
                             meet the cells [box_lo, box_hi); E, H and J outside
                             them are then unspecified afterwards (other paths
                             update every cell). For recomputing a region whose
                             values depend only on the box over the steps of the
                             call (reverse mode: the block plus one cell per step
                             and margin). Zero: the whole grid                    */
  int32_t code_escape;    /* with `code`: 1 = code 255 marks cells whose coefficients
                             are read from a, b, ieps, sigma (cells with their own
                             parameters, e.g. a block being optimised) and the table
                             has at most 255 rows. x-vectorised fused kernel only
                             (not the scalar variant 1, not batches)               */
  /* Optional FieldMonitor accumulation after every step of the call, what a
   * wakesolve with a monitor does with FieldMonitor.update after each step
   * (field_monitors.py:59-102, routines.py:447-448): with mon_acc set, the E that
   * step s produced is added, times mon_phase[mon_row0 + s][f] (re, im), to
   * mon_acc[c][f][n] (complex128 as (re, im) pairs, [3][mon_nf][N]), with the
   * arithmetic of wf_field_monitor_update. mon_phase holds FP64 pairs
   * (complex128 phases), or FP32 pairs with mon_phase_f32 = 1 (complex64).
   * Not in batches. */
  double *mon_acc;
  const void *mon_phase;
  int64_t mon_row0;
  int32_t mon_nf;
  int32_t mon_phase_f32;
  /* Fused path: 1 = advance two steps per launch (temporal blocking) where
   * the two-step kernel applies: FP32 or FP64 storage (not compute_f64),
   * material codes with zero_masked and no escape code, the beam line fed
   * from beam_table (with or without conduction), Nx <= 256 and a multiple of
   * the cells per lane (4 FP32, 2 FP64), row segments (variant 0, full_rows
   * <= 0, cells_per_lane 0 or the cells per lane), no box, no trajectory, no
   * monitor; otherwise one step per launch. An odd step count ends with one
   * single step. Results are identical to one step per launch. 0 = one step
   * per launch; 4 or 8: two steps per launch with that many rows per
   * workgroup (FP64 always 2 rows, two waves per row); 1 = 8 rows (FP64 2) on
   * grids of 2^20 cells or more, FP64 also with Nx >= 224 (otherwise one step
   * per launch is as fast or faster). */
  int32_t two_step;
  int32_t two_step_used;  /* out: 1 if the call ran two steps per launch            */
} wf_step_args;

int wf_version(void);
const char *wf_last_error(void);

/* Replaces SolverFIT3D.__init__'s operator set-up (solverFIT3D.py:203-227). */
int wf_create(const wf_grid *g, wf_handle **out);
int wf_destroy(wf_handle *h);

/* First-step ghost zeroing, SolverFIT3D._set_ghosts_to_0 (solverFIT3D.py:820-840). */
int wf_zero_ghosts(wf_handle *h, void *E, void *H, void *stream);

/* nsteps of SolverFIT3D._one_step (solverFIT3D.py:482-503).
 * path 0 updates E/H/J in place. path 1 (fused) alternates E/H -> E2/H2 ->
 * E/H ...; a->result_swapped tells where the last step's fields are. The
 * fused path drops the flat-index wrap couplings (i=Nx-1 to the next row,
 * j=Ny-1 to the next plane): callers use it only when those terms are zero.
 * With `conductive` it computes J_old = fl(sigma*E_old) (J must hold sigma*E
 * except on the beam line) and writes J = fl(sigma*E_new). */
int wf_step(wf_handle *h, wf_step_args *a, void *stream);

/* Copy the current Ez window into sample `sample` of the capture buffer:
 * wakesolve's save_to_h5(E[xx,yy,zz,'z']) (routines.py:419-425). */
/* One trajectory slot (the layout of wf_step_args.traj_*) from the fields E,
 * H; used for slot 0 (the state before the first step). No reference
 * counterpart. */
int wf_trajectory_gather(wf_handle *h, const void *E, const void *H, void *traj_E, void *traj_H,
                         const int32_t *lo, const int32_t *hi, int64_t slots, int64_t slot,
                         void *stream);

int wf_capture(wf_handle *h, const void *E, void *capture, int32_t ix0, int32_t iy0,
               int32_t nx, int32_t ny, int64_t samples, int64_t sample, void *stream);

/* CPU-semantics ABC face update of emsolve (boundaries.py:313-410,
 * routines.py:216-220): for each low face bit set in `low_faces`, in x, y, z
 * order, plane 0 <- plane 1 for all components of E then H. */
int wf_abc_low(wf_handle *h, void *E, void *H, int32_t low_faces, void *stream);

/* Longitudinal wake potential of every captured line with the reference's
 * line-to-line carry (WakeSolver.calc_long_WP_3d, wakeSolver.py:422-554;
 * calc_long_WP, :310-420 for one line). Output WP3d[nlines][ns] (FP64).
 * Forward (counter_moving = 0, wakeSolver.py:520-533): row k of a line,
 *   it(n,k) = trunc(((z[k]+s[n])/v - zref_v - t0 + ti)/dt), weight dz[k],
 *   with zref_v = zmin/v.
 * Counter-moving (counter_moving = 1, wakeSolver.py:503-517): row nz-1-k,
 *   it(n,k) = trunc(((z[nz-1-k]-s[n])/(-v) - zref_v - t0 + ti)/dt),
 *   weight -dz[k], with zref_v = zmax/(-v).
 * A term is kept if it < nt; a negative it reads column it + ncols of the
 * reference's Ezt (ncols = len(t) >= nt columns, of which 0..nt-1 hold the
 * samples and the rest zeros). Each line ends with WP /= qe12 where
 * qe12 = q*1e12 as the caller computes it. capture is [nlines][cap_samples][Nz]
 * in `dtype`; row k maps to capture z index k + k0; dz is indexed from 0 (as
 * the reference). */
int wf_wake_wp(const void *capture, int32_t dtype, int32_t nlines, int64_t cap_samples,
               int32_t Nz, int32_t k0, int32_t nz, const double *z, const double *dz,
               const double *s, int64_t ns, double v, double zref_v, double t0,
               double ti, double dt, int64_t nt, int64_t ncols, int32_t counter_moving,
               double qe12, double *WP3d, void *stream);

/* Transverse wake potential (WakeSolver.calc_trans_WP, wakeSolver.py:556-629):
 * int_WP[l][n] = numpy-sum(WP3d[l][0:n]) * ds with numpy's chunked pairwise
 * summation, WPx = -(int[2,1]-int[0,1])/dxsum, WPy = -(int[1,2]-int[1,0])/dysum
 * on a 3x3 window (lines i-major). */
int wf_wake_trans(const double *WP3d, int64_t ns, double ds, double dxsum, double dysum,
                  double *WPx, double *WPy, void *stream);

/* Zero-padded DFT bins 0..nbins-1 of real signals (numpy.fft.fft(x, n=N)[mask],
 * wakeSolver.py:691-702, 751-766): X[sig][f] = sum_n x[sig][n] exp(-2 pi i f n/N),
 * twiddle phase reduced exactly as (f*n mod N), compensated FP64 sums.
 * x is [nsig][len], out is [nsig][nbins][2] (re, im). */
int wf_dft(const double *x, int32_t nsig, int64_t len, int64_t N, int64_t nbins,
           double *out, void *stream);

/* FIT field energy of the leapfrog. No reference function computes it; the
 * reference documentation only sketches an energy density
 * (docs/physicsguide.md:181-182). out (device, 2 doubles):
 *   out[0] = 1/2 sum over edges with ieps > 0 and itA > 0 of L/(ieps*itA) E^2
 *   out[1] = 1/2 sum over faces of tL/(imu*iA) Hprev*H
 * (L, tL: the handle's edge lengths; iA, itA: metric2d, laid out as in
 * wf_step_args; weights formed in FP64 from the storage-type values). With
 * E = E^n, Hprev = H^(n-1/2), H = H^(n+1/2) the sum is the discrete energy the
 * lossless leapfrog conserves exactly (up to rounding); with conductivity it
 * decreases. Deterministic FP64 compensated reduction, one workspace per
 * handle: calls on one handle must not overlap. */
int wf_energy(wf_handle *h, const void *E, const void *Hprev, const void *H, const void *ieps,
              const void *imu, const void *metric2d, double *out, void *stream);

/* Loss factor k = -sum_n WP[n]*lam[n]*ds (V/pC for WP in V/pC, lam in 1/m):
 * the integral -int WP(s) lambda(s) ds over the wake's s axis. No reference
 * function computes it. Compensated FP64 sum on the device, out is a device
 * double. */
int wf_loss_factor(const double *WP, const double *lam, int64_t n, double ds, double *out,
                   void *stream);

/* Number of non-finite values (Inf or NaN) among x[0, n), x of the handle's
 * dtype, into the device int64 *out. No reference counterpart: the reference
 * has no failure detection and steps on with NaN/Inf fields. For run health
 * checks between steps (a NaN written into a field, an unstable dt). Uses the
 * handle's reduction workspace: calls on one handle must not overlap. */
int wf_count_nonfinite(wf_handle *h, const void *x, int64_t n, int64_t *out, void *stream);

/* Partial counts held by the scratch of wf_count_nonfinite_noalloc. */
#define WF_NOALLOC_PARTIALS 1024

/* wf_count_nonfinite without a handle and without any allocation. No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc), which therefore cannot create a handle for
 * wf_count_nonfinite's workspace. x holds n values of `dtype` (WF_F32 or
 * WF_F64); the count goes to the device int64 *out, equal to
 * wf_count_nonfinite's. The per-workgroup partial counts live in a static
 * device array of WF_NOALLOC_PARTIALS int64 (8 KiB) that each GPU holds until
 * the process ends: `blocks` workgroups (0: min(ceil(n/256), 1024)), at most
 * WF_NOALLOC_PARTIALS, more is a WF_ERANGE error. Every workgroup writes its
 * own partial before the second kernel reads it, so nothing needs clearing.
 * The first call on a device resolves the array's device address (make it
 * outside graph capture); calls that overlap on one GPU share the array and
 * race. */
int wf_count_nonfinite_noalloc(const void *x, int32_t dtype, int64_t n, int32_t blocks,
                               int64_t *out, void *stream);

/* Batches. No reference counterpart (the reference runs one grid per process);
 * for parameter sweeps over many small grids. Each grid keeps its handle and
 * a fused-path wf_step_args (path 1, E2/H2; material codes with
 * zero_masked = 1 for every grid, or coefficient arrays for every grid; beam
 * table, capture and region trajectory optional), all with the same dtype,
 * conductivity, J mode and boundary order. args[i].nsteps is the number of steps grid i takes
 * part in; beam_row0 and cap_sample0 refer to step 0. One kernel launch per
 * step covers the tiles of every grid (3 rows per workgroup, cells_per_lane 1
 * or the 16-byte width, 0 = 16-byte width; kz z-planes per workgroup, 0 = 32).
 * Even steps read E/H and write E2/H2, odd steps the reverse: after steps
 * [0, S) grid i's fields are in E2/H2 iff min(S, nsteps_i) is odd. Results are
 * bit-identical to wf_step on each grid. wf_batch_create copies the
 * descriptors to the device (the only allocation); the batch must not outlive
 * the handles and buffers it names. */
typedef struct wf_batch wf_batch;
int wf_batch_create(wf_handle *const *handles, const wf_step_args *args, int32_t n,
                    int32_t cells_per_lane, int32_t kz, wf_batch **out);
/* Steps [step0, step0 + nsteps) of every grid. use_graph: replay a captured
 * HIP graph of 16 steps (built on the first such call) instead of launching
 * each step; same results. */
int wf_batch_run(wf_batch *b, int32_t step0, int32_t nsteps, int32_t use_graph, void *stream);
int wf_batch_destroy(wf_batch *b);

/* Reverse mode. No reference counterpart; for gradients of wake figures of
 * merit with respect to per-cell eps_r and sigma. */
typedef struct wf_adjoint_args {
  double *W;              /* device 3N (FP64, or FP32 with fields_f32): adjoint of E
                             plus sigma * adjoint of J (lE + sigma lJ), in at step n+1,
                             out at n (out in W2 with fused)                          */
  double *lH;             /* device 3N adjoint of H, same type and timing as W          */
  double *lJ_region;      /* device FP64 [3][region cells]: adjoint of J on the region
                             (conductive; in at n+1, out at n)                        */
  const void *E0;         /* forward E^n (field dtype)                                    */
  const void *E1, *H1;    /* forward E^{n+1}, H^{n+1} (field dtype)                       */
  const void *beam_row;   /* beam current of step n on the beam line (Nz values), or NULL */
  const void *a, *b;      /* the step's folded coefficients (wf_step_args.a / .b)          */
  const void *ieps, *sigma;
  const void *metric2d;   /* 2-D metric tables without masks (wf_step_args.metric2d layout) */
  double dt;
  int32_t conductive;     /* 1: J = sigma*E is a state (lJ_region used)                    */
  int32_t ixs, iys;       /* beam line (J there is a source); -1 if none                   */
  int32_t r0[3], r1[3];   /* gradient region [r0, r1) in cells                             */
  double *g_ieps;         /* device FP64 [3][region cells] (x fastest), accumulated, or NULL */
  double *g_sigma;        /* same for sigma (conductive)                                   */
  const double *gcap;     /* device FP64 capture gradient [nlines][cap_samples][Nz], or NULL */
  int32_t cap_ix0, cap_iy0, cap_nx, cap_ny;
  int64_t cap_samples;
  int64_t cap_sample;     /* sample n = E^n, added to lE after the step                   */
  int32_t traj;           /* 1: E0, E1 are slots n, n+1 of wf_step's traj_E and H1 slot
                             n+1 of traj_H, with the region as the box (r0 >= 1); 0: full
                             3N fields                                                     */
  int32_t fused;          /* 0: two kernels, W and lH updated in place. 1 (or 5, 7, 9, 13 rows
                             per workgroup): one single-pass kernel reading W, lH and
                             writing W2, lH2 (the caller swaps
                             the pairs for the next step); drops the flat-index wrap terms,
                             so only for set-ups where wf_step's fused path applies, with
                             the region at cell 1 or later in x and y                    */
  double *W2, *lH2;       /* fused: output buffers, 3N each                               */
  int32_t fields_f32;     /* fused only: W, lH, W2, lH2 are FP32 arrays (stored FP32,
                             arithmetic FP64 unless compute_f32); 0: FP64 arrays         */
  void *reg_W;            /* fused only, optional: [slots][3][region cells] in the type of
                             W; slot reg_slot receives the region's W read by this step   */
  double *reg_lJ;         /* with reg_W: [slots][3][region cells]; slot reg_slot receives
                             the region's lJ of step n. With reg_W set, g_ieps must be NULL:
                             the gradient is accumulated afterwards by wf_adjoint_grad    */
  int64_t reg_slot;
  int64_t reg_slots;      /* slots in reg_W (reg_slot < reg_slots)                        */
  const void *kappa;      /* fused only, optional: per edge -dt*ieps*sigma (coefficient
                             dtype), zero on the beam line's z edges; the conduction term
                             becomes kappa*W and ieps is read on the region only        */
  /* fused only, optional: the forward step's material codes (wf_step_args code,
   * mat, nmat, code_escape and metric2d with the masks folded in). The reverse
   * kernel then recomputes a, b, kappa (or sigma) and ieps per cell as the
   * forward folds them, reading one byte per cell; escaped cells read the
   * arrays. Results are identical to the arrays'. */
  const void *code, *mat;
  int32_t nmat, code_escape;
  const void *code_metric2d;
  int32_t compute_f32;    /* fused only, with fields_f32: the step's arithmetic and LDS
                             staging in FP32 instead of FP64 (lJ_region, reg_lJ and the
                             gradients stay FP64)                                        */
} wf_adjoint_args;

/* Reverse of step n of SolverFIT3D._one_step (solverFIT3D.py:482-503): maps
 * the adjoints of (E, H, J)^{n+1} to those of (E, H, J)^n with the transposed
 * stencils (M1^T, M2^T, same masks and flat-index terms as the forward), adds
 * the capture gradient of sample n, and accumulates d/d ieps and d/d sigma of
 * the cells in the region. The J of step n is taken as wakesolve's: beam_row
 * on the beam line, fl(sigma*E^n) elsewhere (0 when not conductive). */
int wf_adjoint_step(wf_handle *h, const wf_adjoint_args *a, void *stream);

/* Parameter gradients of nsteps reverse steps in one pass (no reference
 * counterpart), from the stored slots: forward E of steps n0+s and n0+s+1
 * (E + s*E_stride, E + (s+1)*E_stride), H of step n0+s+1 (H + (s+1)*H_stride),
 * the region's W of reverse step n0+s (reg_W slot s) and lJ of step n0+s+1
 * (reg_lJ slot s+1, zeros after the run's last step), and the beam row of step
 * n0+s (beam + s*Nz). Adds to g_ieps / g_sigma exactly what nsteps per-step
 * calls with gradients would add. */
typedef struct wf_adjoint_grad_args {
  const void *E, *H;      /* field dtype; trajectory boxes (traj = 1) or full 3N fields */
  int64_t E_stride, H_stride; /* elements per slot                                     */
  int32_t traj;
  const void *beam;       /* beam table rows from step n0, or NULL                      */
  const void *sigma, *metric2d;
  const void *reg_W;      /* [nsteps][3][region cells], FP32 if fields_f32 else FP64    */
  const double *reg_lJ;   /* [nsteps + 1][3][region cells]                              */
  int32_t fields_f32;
  int32_t nsteps;
  double dt;
  int32_t conductive, ixs, iys;
  int32_t r0[3], r1[3];
  double *g_ieps, *g_sigma;
} wf_adjoint_grad_args;

int wf_adjoint_grad(wf_handle *h, const wf_adjoint_grad_args *a, void *stream);

/* Reverse mode over many grids, one launch per reverse step (no reference
 * counterpart; the reverse of wf_batch_*). Grid i runs nsteps[i] reverse steps;
 * reverse step r of the batch is step nsteps[i]-1-r of grid i (grids with fewer
 * steps start later). args[i] as for wf_adjoint_step with the fused kernel and
 * the region slots (fused = 1, W/lH and W2/lH2, reg_W/reg_lJ with reg_slots >=
 * nsteps[i], g_ieps = NULL); cap_sample and reg_slot are set per step. Even r
 * reads W/lH and writes W2/lH2, odd r the reverse. dtype, fields_f32 and
 * compute_f32 must match across the batch. The gradients are accumulated afterwards by
 * wf_adjoint_grad, one call per grid. */
typedef struct wf_batch_adjoint wf_batch_adjoint;
int wf_batch_adjoint_create(wf_handle *const *handles, const wf_adjoint_args *args,
                            const int32_t *nsteps, int32_t n, wf_batch_adjoint **out);
/* Reverse steps [r0, r0 + nr); use_graph: replay a captured HIP graph of 16
 * steps (built on the first such call). */
int wf_batch_adjoint_run(wf_batch_adjoint *b, int32_t r0, int32_t nr, int32_t use_graph,
                         void *stream);
int wf_batch_adjoint_destroy(wf_batch_adjoint *b);

/* Transpose of wf_wake_wp's line integral without the line-to-line carry:
 * gcap[l][m][k0+kr] += sum over s indices n with it(n,k) == m of gI[l][n]*w_k
 * (gI: gradient w.r.t. the raw line sums; the carry and the 1/qe12 scaling
 * are the caller's). kr, it and w_k as in wf_wake_wp: kr = k and w_k = dz[k],
 * or with counter_moving kr = nz-1-k and w_k = -dz[k]; zref_v is zmin/v, or
 * zmax/(-v) with counter_moving. Other arguments as wf_wake_wp. No
 * counterpart in the reference (its wake post-processing has no gradient). */
int wf_wake_wp_adjoint(const double *gI, int32_t nlines, int64_t cap_samples, int32_t Nz,
                       int32_t k0, int32_t nz, const double *z, const double *dz, const double *s,
                       int64_t ns, double v, double zref_v, double t0, double ti, double dt,
                       int64_t nt, int32_t counter_moving, double *gcap, void *stream);

/* Transpose of wf_dft for one signal: out[n] = Re sum_f w[f] exp(-2 pi i f n/N)
 * for n < len, w given as [nbins][2] (re, im); exact phase reduction,
 * compensated FP64 sums. */
int wf_idft_real(const double *w, int64_t nbins, int64_t N, int64_t len, double *out,
                 void *stream);

/* One FieldMonitor.update (field_monitors.py:59-102): for the 3N values of E
 * (field dtype) and each of nfreq monitored frequencies,
 * acc[c][f][n] += E_c[n] * phase[f], phase given as device [nfreq][2] (re, im)
 * = exp(-2 pi i f t) evaluated by the caller: FP64 pairs (complex128), or FP32
 * pairs with phase_f32 = 1 (complex64, what numpy evaluates for an FP32 t).
 * The product is numpy's complex multiply in the promoted type (FP32 for FP32
 * E with a complex64 phase, else FP64); acc is device complex128 laid out
 * [3][nfreq][N][2]. */
int wf_field_monitor_update(const void *E, int32_t dtype, int64_t N, const void *phase,
                            int32_t phase_f32, int32_t nfreq, double *acc, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* WAKIS_FIT_H */
