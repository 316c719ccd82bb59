"""Many grids advanced together: one kernel launch per step for all of them.

No reference counterpart (Wakis runs one grid per process). For parameter
sweeps over many small cavities: each grid is an ordinary ``SolverFIT3D``
(its own materials, boundaries, beam and ``WakeSolver``); ``wakesolve_batch``
runs their wakefield simulations in one batched launch per step
(``wf_batch_create``/``wf_batch_run``) with a captured HIP graph for the step
loop, then the per-grid wake post-processing. The results are bit-identical to
running each grid's ``wakesolve`` alone, for every tile choice, which is what
lets ``autotune`` time the candidate tile shapes on the run's own first steps.
"""

import ctypes
import time

import numpy as np
import torch

from . import _lib


class BatchNotApplicable(RuntimeError):
    """A grid of the batch cannot run on the batched fused path."""


def _native_width(dtype):
    return 16 // np.dtype(dtype).itemsize


class _Batch:
    def __init__(self, solvers, args, cells_per_lane, kz, full_rows=False):
        n = len(solvers)
        handles = (ctypes.c_void_p * n)(*[s._handle.value for s in solvers])
        arr = (_lib.WfStepArgs * n)(*args)
        for a in arr:  # whole-row tiles (rows per grid picked by the library), or segments
            a.full_rows = 1 if full_rows else 0
        self.ptr = ctypes.c_void_p()
        _lib.check(_lib.lib().wf_batch_create(handles, arr, n, int(cells_per_lane), int(kz),
                                              ctypes.byref(self.ptr)), "wf_batch_create")
        self._keep = (handles, arr)

    def run(self, step0, nsteps, use_graph):
        _lib.check(_lib.lib().wf_batch_run(self.ptr, int(step0), int(nsteps), int(use_graph),
                                           _lib.stream_ptr()), "wf_batch_run")

    def close(self):
        if self.ptr is not None and self.ptr.value:
            _lib.lib().wf_batch_destroy(self.ptr)
        self.ptr = None

    __del__ = close


def _without_escape(a):
    """Batches take material codes without the escape code: such a grid reads
    its coefficient arrays instead."""
    if a.code and a.code_escape:
        a.code = a.mat = a.metric2d = None
        a.nmat = 0
        a.code_escape = 0


def _candidates(solvers):
    """Tile shapes to try, (cells per lane, z-planes per workgroup, whole rows):
    row segments of 1 or (when every grid allows it) 16 bytes per lane, and
    whole-row workgroups when every row fits one (Nx / width <= 128), each with
    16 and 32 z-planes per workgroup."""
    widths = [(1, False)]
    nat = _native_width(solvers[0].dtype)
    if all(s.Nx % nat == 0 for s in solvers):
        widths.append((nat, False))
        if all(s.Nx // nat <= 128 for s in solvers):
            widths.append((nat, True))
    return [(w, kz, fr) for w, fr in widths for kz in (16, 32)]


def wakesolve_batch(solvers, wakelength, compute_plane="both", cells_per_lane=None, kz=None,
                    use_graph=True, autotune=True, tune_steps=32, verbose=0, full_rows=None):
    """``wakesolve`` of every solver, advanced together (``wakelength``: one
    value or one per solver).

    Each solver needs its ``wake`` (WakeSolver) and must qualify for the fused
    path: with material codes and zero-valued masked fields (the usual state
    of a fresh wakefield run), or for every grid without material codes
    (coefficient arrays, masks applied in the kernel); otherwise
    BatchNotApplicable is raised before anything runs. ``cells_per_lane``/``kz``/
    ``full_rows`` (whole-row workgroups, rows per grid picked) fix the tile shape; with
    ``autotune`` the candidates are timed on successive chunks of
    ``tune_steps`` steps of the run itself and the fastest runs the rest.
    Returns a dict with the chosen shape, the step count and the timing.
    """
    if not solvers:
        raise ValueError("wakesolve_batch: no solvers")
    dtypes = {np.dtype(s.dtype) for s in solvers}
    if len(dtypes) != 1:
        raise BatchNotApplicable("all solvers of a batch need the same dtype")
    wl = list(wakelength) if np.ndim(wakelength) else [wakelength] * len(solvers)
    ctxs = [s._wake_setup(w, compute_plane) for s, w in zip(solvers, wl)]
    args = []
    for s, ctx in zip(solvers, ctxs):
        a = s._wake_args(ctx, 0, ctx["Nt"])
        if not s._prepare_fused(a):
            raise BatchNotApplicable("a grid of the batch does not qualify for the fused path")
        _without_escape(a)
        if a.code and not a.zero_masked:
            raise BatchNotApplicable("batched grids with material codes need zero masked fields")
        args.append(a)
    if len({bool(a.code) for a in args}) != 1:
        raise BatchNotApplicable("all grids of a batch use material codes, or none")
    nmax = max(ctx["Nt"] for ctx in ctxs)
    if cells_per_lane is not None or kz is not None or full_rows is not None or not autotune:
        cands = [(cells_per_lane or 0, kz or 0, bool(full_rows))]
    else:
        cands = _candidates(solvers)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    stream = torch.cuda.current_stream()
    step = 0
    timings = []
    t0 = time.perf_counter()
    if len(cands) > 1:
        for cpl, kzc, fr in cands:
            if step + tune_steps > nmax:
                break
            b = _Batch(solvers, args, cpl, kzc, fr)
            ev0.record(stream)
            b.run(step, tune_steps, False)
            ev1.record(stream)
            torch.cuda.synchronize()
            timings.append((ev0.elapsed_time(ev1), cpl, kzc, fr))
            b.close()
            step += tune_steps
        best = min(timings)[1:] if timings else cands[0]
    else:
        best = cands[0]
    b = _Batch(solvers, args, *best)
    ev0.record(stream)
    b.run(step, nmax - step, use_graph)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms_main = ev0.elapsed_time(ev1)
    main_steps = nmax - step
    b.close()
    wall = time.perf_counter() - t0
    for s, ctx in zip(solvers, ctxs):
        Nt = ctx["Nt"]
        s._finish_run(True, Nt % 2 == 1)
        s.nstep += Nt
        s.simulationTime = wall
        s._wake_finish(ctx, compute_plane)
    cells = sum(s.N * ctx["Nt"] for s, ctx in zip(solvers, ctxs))
    info = dict(shape=dict(cells_per_lane=best[0], kz=best[1], full_rows=bool(best[2])),
                steps=nmax,
                tuning=[dict(ms=t, cells_per_lane=c, kz=k, full_rows=bool(f))
                        for t, c, k, f in timings],
                ms_main=ms_main, main_steps=main_steps, wall_s=wall, cell_updates=cells)
    if verbose:
        print(f"wakesolve_batch: {len(solvers)} grids, {nmax} steps, shape {best}, "
              f"{wall:.2f} s")
    return info


class _AdjointBatch:
    def __init__(self, solvers, args, nsteps):
        n = len(solvers)
        handles = (ctypes.c_void_p * n)(*[s._handle.value for s in solvers])
        arr = (_lib.WfAdjointArgs * n)(*args)
        nst = (ctypes.c_int32 * n)(*[int(k) for k in nsteps])
        self.ptr = ctypes.c_void_p()
        _lib.check(_lib.lib().wf_batch_adjoint_create(handles, arr, nst, n,
                                                      ctypes.byref(self.ptr)),
                   "wf_batch_adjoint_create")
        self._keep = (handles, arr, nst)

    def run(self, r0, nr, use_graph):
        _lib.check(_lib.lib().wf_batch_adjoint_run(self.ptr, int(r0), int(nr), int(use_graph),
                                                   _lib.stream_ptr()), "wf_batch_adjoint_run")

    def close(self):
        if self.ptr is not None and self.ptr.value:
            _lib.lib().wf_batch_adjoint_destroy(self.ptr)
        self.ptr = None

    __del__ = close


def wake_adjoint_batch(adjs, eps_r, sigma=None, cells_per_lane=0, kz=0, use_graph=True,
                       full_rows=None):
    """Loss and gradients of several ``WakeAdjoint`` problems (one solver and
    block each) computed together.

    The forward runs are one batched launch per step (``wf_batch_*``), with
    each block's trajectory captured inside the step loop. The reverse steps
    are one batched launch per step (``wf_batch_adjoint_*``, the single-pass
    reverse kernel over all grids). Each grid's gradients are then accumulated
    in one pass. The results equal each problem's own ``forward`` and
    ``backward`` (trajectory mode, fused reverse kernel).

    ``eps_r``, ``sigma``: one block array per problem (``sigma`` None for
    solvers without conductivity). Returns (losses, gradients), where
    gradients[i] = (dL/d eps_r, dL/d sigma or None). Raises
    BatchNotApplicable before anything runs if a grid cannot take the batched
    paths. ``full_rows``: whole-row workgroups for the forward steps (None:
    when every grid's rows fit one, as the forward batch tuner picks for
    short rows).
    """
    n = len(adjs)
    if n == 0:
        raise ValueError("wake_adjoint_batch: no problems")
    sigma = list(sigma) if sigma is not None else [None] * n
    solvers = [adj.s for adj in adjs]
    if len({np.dtype(s.dtype) for s in solvers}) != 1:
        raise BatchNotApplicable("all solvers of a batch need the same dtype")
    for adj in adjs:
        if min(adj.i0, adj.j0, adj.k0) < 1:
            raise BatchNotApplicable("batched blocks start at cell 1 or later on every axis")
    # ---- forward: parameters, wake set-up, trajectories (slot 0 before the run)
    args, nts = [], []
    for adj, er, sg in zip(adjs, eps_r, sigma):
        s = adj.s
        adj._fwd = None
        adj._set_params(er, sg)
        ctx = s._wake_setup(adj.wakelength, "both")
        Nt = ctx["Nt"]
        a = s._wake_args(ctx, 0, Nt)
        if not s._prepare_fused(a):
            raise BatchNotApplicable("a grid of the batch does not qualify for the fused path")
        _without_escape(a)
        tE = torch.empty((Nt + 1, 3 * adj._nE), dtype=s.E.array.dtype, device=s.device)
        tH = torch.empty((Nt + 1, 3 * adj._nH), dtype=s.E.array.dtype, device=s.device)
        lo, hi = adj._lohi()
        _lib.check(_lib.lib().wf_trajectory_gather(
            s._handle, _lib.ptr(s.E.array), _lib.ptr(s.H.array), _lib.ptr(tE), _lib.ptr(tH),
            lo, hi, Nt + 1, 0, _lib.stream_ptr()), "wf_trajectory_gather")
        a.traj_E, a.traj_H = _lib.ptr(tE), _lib.ptr(tH)
        a.traj_lo[:] = tuple(lo)
        a.traj_hi[:] = tuple(hi)
        a.traj_slots, a.traj_slot0 = Nt + 1, 0
        adj._fwd = dict(ctx=ctx, Nt=Nt, traj=(tE, tH))
        args.append(a)
        nts.append(Nt)
    if len({bool(a.code) for a in args}) != 1:
        raise BatchNotApplicable("all grids of a batch use material codes, or none")
    if any(a.code and not a.zero_masked for a in args):
        raise BatchNotApplicable("batched grids with material codes need zero masked fields")
    nmax = max(nts)
    if full_rows is None:
        nat = _native_width(solvers[0].dtype)
        full_rows = cells_per_lane in (0, nat) and all(
            s.Nx % nat == 0 and s.Nx // nat <= 128 for s in solvers)
    b = _Batch(solvers, args, cells_per_lane, kz, bool(full_rows))
    b.run(0, nmax, use_graph)
    b.close()
    losses = []
    for adj, Nt in zip(adjs, nts):
        s = adj.s
        s._finish_run(True, Nt % 2 == 1)
        s.nstep += Nt
        s._wake_finish(adj._fwd["ctx"], "both")
        losses.append(adj._loss_value())
    # ---- reverse: per-grid state, batched steps, one gradient pass per grid
    Rs, rargs, regs = [], [], []
    for adj, Nt in zip(adjs, nts):
        R = adj._reverse_setup()
        if not R["fused"]:
            raise BatchNotApplicable("the fused reverse kernel does not apply to a grid")
        a = R["a"]
        rW = torch.empty((max(Nt, 1), R["n3r"]), dtype=R["adt"], device=adj.s.device)
        rJ = torch.zeros((Nt + 1, R["n3r"]), dtype=torch.float64, device=adj.s.device)
        a.reg_W, a.reg_lJ, a.reg_slots = _lib.ptr(rW), _lib.ptr(rJ), max(Nt, 1)
        a.g_ieps = a.g_sigma = None
        Rs.append(R)
        rargs.append(a)
        regs.append((rW, rJ))
    rb = _AdjointBatch(solvers, rargs, nts)
    rb.run(0, nmax, use_graph)
    rb.close()
    grads = []
    for adj, R, (rW, rJ), Nt in zip(adjs, Rs, regs, nts):
        tE, tH = adj._fwd["traj"]
        R["grad_pass"](tE.data_ptr(), tE.stride(0), tH.data_ptr(), tH.stride(0), 1,
                       R["table"].data_ptr(), rW, rJ, Nt)
        grads.append(adj._finish_grad(R["g_ieps"], R["g_sig"]))
    torch.cuda.synchronize()
    return losses, grads


# --------------------------------------------------------------------------
# Synthetic sweep: box cavities with a beam pipe and a dielectric block
# --------------------------------------------------------------------------
def random_box_cavities(seed, n=64, cell=1.0e-3):
    """``n`` box-cavity set-ups drawn from ``np.random.default_rng(seed)``.

    Each: a PEC-walled rectangular cavity with sides U[20, 80] mm, a round
    beam pipe of radius U[3, 10] mm along z through both domain ends, and an
    on-axis dielectric block (a third of the cavity in each direction) with
    eps_r U[1, 6]. The domain adds 4 mm around the cavity in x/y and a pipe
    section of half the cavity length on each z side; cells are about
    ``cell`` wide, with Nx, Ny in [32, 96] and Nz in [32, 160], all multiples
    of 4. Returns a list of dicts (x, y, z node axes, vacuum and dielectric
    cell masks, eps_r).
    """
    from .solids import Box, Cylinder
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        a, b, c = rng.uniform(20e-3, 80e-3, 3)
        r = rng.uniform(3e-3, 10e-3)
        eps_r = float(rng.uniform(1.0, 6.0))
        hx, hy = a / 2 + 4e-3, b / 2 + 4e-3
        hz = c / 2 + c / 2
        dims = []
        for half, lo, hi in ((hx, 32, 96), (hy, 32, 96), (hz, 32, 160)):
            m = int(np.clip(round(2 * half / cell / 4) * 4, lo, hi))
            dims.append(m)
        x = np.linspace(-hx, hx, dims[0] + 1)
        y = np.linspace(-hy, hy, dims[1] + 1)
        z = np.linspace(-hz, hz, dims[2] + 1)
        vac = Box(-a / 2, a / 2, -b / 2, b / 2, -c / 2, c / 2) | Cylinder(min(r, a / 2, b / 2))
        diel = Box(-a / 6, a / 6, -b / 6, b / 6, -c / 6, c / 6)
        out.append(dict(x=x, y=y, z=z, vacuum=vac.cell_mask(x, y, z),
                        dielectric=diel.cell_mask(x, y, z), eps_r=eps_r,
                        sides=(a, b, c), r_pipe=r))
    return out


def dielectric_region(spec):
    """((i0, i1), (j0, j1), (k0, k1)): the cell bounding box of a
    ``random_box_cavities`` entry's dielectric block."""
    idx = np.nonzero(spec["dielectric"])
    return tuple((int(v.min()), int(v.max()) + 1) for v in idx)


def cavity_solver(spec, dtype=np.float32, sigmaz=5e-3, q=1e-9, beta=1.0):
    """SolverFIT3D (PEC background, PEC x/y, ABC z) with a WakeSolver for one
    ``random_box_cavities`` entry."""
    from .gridFIT3D import GridFIT3D
    from .solverFIT3D import SolverFIT3D
    from .wakeSolver import WakeSolver
    grid = GridFIT3D(x=spec["x"], y=spec["y"], z=spec["z"],
                     stl_solids={"vac": spec["vacuum"], "diel": spec["dielectric"]},
                     stl_materials={"vac": "vacuum", "diel": [spec["eps_r"], 1.0]}, verbose=0)
    wake = WakeSolver(q=q, sigmaz=sigmaz, beta=beta, skip_cells=4, save=False, verbose=0)
    return SolverFIT3D(grid, wake=wake, bc_low=["pec", "pec", "abc"],
                       bc_high=["pec", "pec", "abc"], use_stl=True, bg="pec", dtype=dtype,
                       verbose=0)
