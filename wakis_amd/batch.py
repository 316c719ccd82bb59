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
    def __init__(self, solvers, args, cells_per_lane, kz):
        n = len(solvers)
        handles = (ctypes.c_void_p * n)(*[s._handle.value for s in solvers])
        arr = (_lib.WfStepArgs * n)(*args)
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


def _candidates(solvers):
    """Tile shapes to try: cells per lane (1, or the 16-byte width when every
    grid allows it) times z-planes per workgroup."""
    widths = [1]
    nat = _native_width(solvers[0].dtype)
    if all(s.Nx % nat == 0 for s in solvers):
        widths.append(nat)
    return [(w, kz) for w in widths for kz in (16, 32)]


def wakesolve_batch(solvers, wakelength, compute_plane="both", cells_per_lane=None, kz=None,
                    use_graph=True, autotune=True, tune_steps=32, verbose=0):
    """``wakesolve`` of every solver, advanced together (``wakelength``: one
    value or one per solver).

    Each solver needs its ``wake`` (WakeSolver) and must qualify for the fused
    path with material codes and zero-valued masked fields (the usual state of
    a fresh wakefield run); otherwise BatchNotApplicable is raised before
    anything runs. ``cells_per_lane``/``kz`` fix the tile shape; with
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
        if not a.code or not a.zero_masked:
            raise BatchNotApplicable("batched grids need material codes and zero masked fields")
        args.append(a)
    nmax = max(ctx["Nt"] for ctx in ctxs)
    if cells_per_lane is not None or kz is not None or not autotune:
        cands = [(cells_per_lane or 0, kz or 0)]
    else:
        cands = _candidates(solvers)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    stream = torch.cuda.current_stream()
    step = 0
    timings = []
    t0 = time.perf_counter()
    if len(cands) > 1:
        for cpl, kzc in cands:
            if step + tune_steps > nmax:
                break
            b = _Batch(solvers, args, cpl, kzc)
            ev0.record(stream)
            b.run(step, tune_steps, False)
            ev1.record(stream)
            torch.cuda.synchronize()
            timings.append((ev0.elapsed_time(ev1), cpl, kzc))
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
    info = dict(shape=dict(cells_per_lane=best[0], kz=best[1]), steps=nmax,
                tuning=[dict(ms=t, cells_per_lane=c, kz=k) for t, c, k in timings],
                ms_main=ms_main, main_steps=main_steps, wall_s=wall, cell_updates=cells)
    if verbose:
        print(f"wakesolve_batch: {len(solvers)} grids, {nmax} steps, shape {best}, "
              f"{wall:.2f} s")
    return info


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
