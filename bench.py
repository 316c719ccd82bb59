"""Throughput of the FIT leapfrog (wakefield step loop) on one MI355X.

Workload (N=1): PEC cubic cavity 50 mm with a round beam pipe r=5 mm along z
in a 60x60x160 mm domain on a uniform 256x256x512 grid (33.5M cells), vacuum
inside, PEC background, PEC on x/y and ABC on z; Gaussian beam q=1e-9 C,
sigma_z=10 mm, beta=1 on axis. A step is one leapfrog step over the whole grid
as wakesolve runs it: beam J line from the precomputed table, H and E
updates, Ez 3x3 window capture. Fields start at zero (synthetic input, no
data files). Timed with HIP events on the stream the kernels run on.

Beside the timed steps, the JSON line carries legs whose length does not
depend on --steps: the accuracy of the FP32 form against the FP64 kernel on
the full 0.5 m wakefield run of this workload (10,179 steps), the lossy
pillbox with PML (timing, and its 1 m wake at beta = 0.8 with a 2 mm source
offset), the reverse-mode cost, the batched 64-cavity sweep, and the CPU
baseline (the oracle, i.e. the reference's scipy algorithm, on the host).

Prints one JSON line. --gpus N>1 exits: this engine runs on one GPU.
"""

import argparse
import contextlib
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("Gcell-updates/s + %HBM peak (FP32 vs FP64) at 256x256x512; adjoint/fwd time")
HBM_PEAK_GBS = 8000.0
VALUES_PER_CELL = 18  # E r/w 3+3, H r/w 3+3, face coeffs 3, edge coeffs 3

# HBM bytes per launch of the step kernel on the default workload, measured
# with rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE over this script
# (profiles/k_fused4_t2_cavity_pipe_256x256x512.md for FP32, two steps per
# launch; profiles/k_fused4_cavity_pipe_256x256x512.md for FP64): FETCH_SIZE is scaled by 2.0,
# the ratio of the bytes a 1 GiB device copy reads to what the counter reports
# for it; WRITE_SIZE matches the copy's bytes and is taken as is. Keyed by
# (grid, dtype, path, variant, kz); other settings report null.
MEASURED_TRAFFIC = {
    # one launch of the two-step kernel (k_fused4_t2) = two steps
    ((256, 256, 512), "f32", "fused", 0, 0): {"fetch_B": 888_720_998, "write_B": 805_661_184,
                                             "steps_per_launch": 2},
    ((256, 256, 512), "f64", "fused", 0, 0): {"fetch_B": 1_814_107_750, "write_B": 1_610_804_429,
                                             "steps_per_launch": 2},
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(grid, dtype, path, variant, kz, steps_per_launch=1):
    """HBM bytes of one launch of the dominant kernel (rocprofv3), or None when
    not measured for this setting."""
    t = MEASURED_TRAFFIC.get((tuple(grid), dtype, path, variant, kz))
    if t is None or t["steps_per_launch"] != steps_per_launch:
        return None
    return t["fetch_B"] + t["write_B"]


def roofline(r, grid, dtype, variant, kz, copy_gbs=None):
    """The roofline object of a time_gpu result. `achieved` counts the
    algorithmic bytes of every cell-update (72 B FP32, 144 B FP64: E and H read
    and written once, six coefficients) over the kernel time. The two-step
    kernel moves E and H once per two steps, so its real traffic per step is
    about half of that and `frac` passes 1; `traffic` (per launch) and
    `actual_GBs` give what it really moves."""
    spl = 2 if r.get("two_step") else 1
    traffic = measured_traffic(grid, dtype, r["path"], variant, kz, spl)
    out = {"bound": "hbm", "achieved": round(r["achieved_GBs"], 1), "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": round(r["achieved_GBs"] / HBM_PEAK_GBS, 4),
           "traffic": traffic, "bytes_per_cell_update": r["bytes_per_cell"],
           "kernel": ("k_fused4_t2" if spl == 2 else "k_fused4") if r["path"] == "fused"
           else "k_h_exact + k_e_exact", "steps_per_launch": spl}
    if traffic is not None:
        out["actual_GBs"] = round(traffic / (spl * r["ms_per_step"] * 1e-3) / 1e9, 1)
    if spl == 2:
        out["note"] = ("two leapfrog steps per launch: E and H cross HBM once per two steps, "
                       "so the per-cell-update algorithmic bytes overstate the real traffic "
                       "and frac exceeds 1; traffic is per launch (two steps), actual_GBs "
                       "the real HBM rate")
    if copy_gbs is not None:
        out["hbm_copy_GBs"] = round(copy_gbs, 1)
    return out


def finite_or_none(x, why):
    """A finite float, or (None, why): the JSON line never carries NaN/Inf."""
    x = float(x)
    return (x, None) if math.isfinite(x) else (None, why)


def rel_l2_np(a, b, what):
    """Relative L2 of a against b as a dict entry: value or null with the reason."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = float(np.linalg.norm(b))
    if nb == 0.0:
        return {"value": None, "why": f"{what}: the reference is identically zero"}
    v, why = finite_or_none(np.linalg.norm(a - b) / nb, f"{what}: not finite")
    return {"value": v} if why is None else {"value": None, "why": why}


def _rel_l2(a, b):
    import torch
    nb = float(torch.linalg.norm(b))
    return float(torch.linalg.norm(a - b)) / nb if nb > 0 else None


def cavity_grid(Nx, Ny, Nz):
    import wakis_amd as wa
    from wakis_amd.solids import Box, Cylinder
    solid = Box(-25e-3, 25e-3, -25e-3, 25e-3, -25e-3, 25e-3) | Cylinder(5e-3)
    return wa.GridFIT3D(-30e-3, 30e-3, -30e-3, 30e-3, -80e-3, 80e-3, Nx, Ny, Nz,
                        stl_solids={"cavity": solid}, stl_materials={"cavity": "vacuum"},
                        verbose=0)


TWO_STEP = [1]  # the solvers' fused_two_step (--two-step)


def build_solver(Nx, Ny, Nz, dtype, path, wake=None):
    import wakis_amd as wa
    s = wa.SolverFIT3D(cavity_grid(Nx, Ny, Nz), wake=wake, bc_low=["pec", "pec", "abc"],
                       bc_high=["pec", "pec", "abc"], use_stl=True, bg="pec", dtype=dtype,
                       verbose=0, path=path)
    s.fused_two_step = TWO_STEP[0]
    return s


def hbm_copy_gbs(dev, nbytes=1 << 30, reps=5):
    """Device-to-device copy of `nbytes` (read + write counted): the HBM rate a
    plain streaming kernel reaches on this GPU, and the reference dispatch for
    rocprofv3 FETCH_SIZE/WRITE_SIZE scaling."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize(dev)
    stream = torch.cuda.current_stream(dev)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        dst.copy_(src)
    e1.record(stream)
    torch.cuda.synchronize(dev)
    ms = e0.elapsed_time(e1)
    del src, dst
    torch.cuda.empty_cache()
    return 2.0 * nbytes * reps / (ms * 1e-3) / 1e9


def time_gpu(Nx, Ny, Nz, dtype, steps, warmup, path, variant=0, kz=0):
    """`warmup` untimed then `steps` timed wakefield steps (beam table, 3x3
    capture) on the cavity/pipe workload, after the workgroup-shape trials."""
    import torch

    import wakis_amd as wa
    from wakis_amd import _lib
    s = build_solver(Nx, Ny, Nz, dtype, path)
    s.fused_variant = variant
    s.fused_kz = kz
    tot = warmup + steps
    beam = wa.Beam(q=1e-9, sigmaz=10e-3, beta=1.0)
    dev = s.device
    table = torch.as_tensor(beam.table(s, tot), device=dev).contiguous()
    cap = torch.zeros((9, tot, s.Nz), dtype=table.dtype, device=dev)
    ixt = int(np.abs(s.x).argmin())
    iyt = int(np.abs(s.y).argmin())
    s._set_ghosts_to_0()
    s.step_0 = False
    s._beam_line = (int(beam.ixs), int(beam.iys))

    def args(n0, k):
        a = s._step_args(k)
        a.ixs, a.iys = int(beam.ixs), int(beam.iys)
        a.jmode = _lib.WF_J_LINE
        a.beam_table, a.beam_row0, a.beam_rows = _lib.ptr(table), n0, tot
        a.capture = _lib.ptr(cap)
        a.cap_ix0, a.cap_iy0, a.cap_nx, a.cap_ny = ixt - 1, iyt - 1, 3, 3
        a.cap_samples, a.cap_sample0 = tot, n0 + 1
        return a

    def run(n0, k):
        if k > 0:
            s._run(args(n0, k))

    s.tune_fused(lambda k: args(0, k))  # workgroup shape by measurement (fields restored)
    run(0, warmup)
    torch.cuda.synchronize(dev)
    stream = torch.cuda.current_stream(dev)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    w0 = time.perf_counter()
    e0.record(stream)
    run(warmup, steps)
    e1.record(stream)
    torch.cuda.synchronize(dev)
    wall = time.perf_counter() - w0
    ms = e0.elapsed_time(e1)
    finite = bool(torch.isfinite(s.E.array).all().item())
    N = s.N
    out = dict(N=N, steps=steps, ms_total=ms, ms_per_step=ms / steps, wall_s=wall,
               path=s.last_path, two_step=bool(s.last_two_step),
               gcell_s=N * steps / (ms * 1e-3) / 1e9, finite=finite,
               bytes_per_cell=VALUES_PER_CELL * np.dtype(dtype).itemsize,
               full_rows=s.fused_full_rows, tuning=_tuning(s))
    out["achieved_GBs"] = N * out["bytes_per_cell"] / (out["ms_per_step"] * 1e-3) / 1e9
    del s, table, cap
    torch.cuda.empty_cache()
    return out


def wake_run(s, wakelength, snap=None):
    """The body of SolverFIT3D.wakesolve on solver `s` (whose WakeSolver is
    set): set-up, shape trials, the step loop (with E, H kept in FP64 after
    `snap` steps when given) and the wake post-processing, the last two
    timed apart."""
    import torch
    ctx = s._wake_setup(wakelength, "both")
    Nt = ctx["Nt"]
    s.tune_fused(lambda k: s._wake_args(ctx, 0, k))
    torch.cuda.synchronize(s.device)
    t0 = time.perf_counter()
    k1 = Nt if snap is None else min(snap, Nt)
    kept = None
    for n0, k in ((0, k1), (k1, Nt - k1)):
        if k > 0:
            s._run(s._wake_args(ctx, n0, k))
            s.nstep += k
        if n0 == 0 and snap is not None:  # copies: the ping-pong buffers get reused
            kept = (s.E.array.to(torch.float64, copy=True),
                    s.H.array.to(torch.float64, copy=True), k1)
    torch.cuda.synchronize(s.device)
    t_loop = time.perf_counter() - t0
    s.fields_finite = s.check_finite(("E", "H"))
    t0 = time.perf_counter()
    s._wake_finish(ctx, "both")
    torch.cuda.synchronize(s.device)
    t_post = time.perf_counter() - t0
    return dict(Nt=Nt, N=s.N, dt=float(s.dt), loop_s=t_loop, post_s=t_post, kept=kept,
                path=s.last_path, full_rows=s.fused_full_rows, finite=bool(s.fields_finite))


def _wake_compare(w32, w64, with_trans=True):
    """FP32 run's wake against the FP64 run's on the common s / f points (the
    two runs' s and f axes come from their own dt: the FP32 dt is the FP64
    one rounded to FP32, so the axes agree to ~1e-7 relative)."""
    n = min(len(w32.s), len(w64.s))
    m = min(len(w32.f), len(w64.f))
    out = {"n_s": int(n), "n_f": int(m),
           "s_axes_max_rel_diff": float(np.max(np.abs(w32.s[:n] - w64.s[:n]))
                                        / np.max(np.abs(w64.s[:n]))),
           "rel_l2_WP_z": rel_l2_np(w32.WP[:n], w64.WP[:n], "WP_z"),
           "rel_l2_abs_Z_z": rel_l2_np(np.abs(w32.Z[:m]), np.abs(w64.Z[:m]), "|Z_z|"),
           "rel_l2_Z_z": rel_l2_np(np.c_[w32.Z[:m].real, w32.Z[:m].imag],
                                   np.c_[w64.Z[:m].real, w64.Z[:m].imag], "Z_z"),
           "loss_factor_V_per_pC": {"fp32": w32.loss_factor, "fp64": w64.loss_factor}}
    if with_trans:
        out["rel_l2_WP_x"] = rel_l2_np(w32.WPx[:n], w64.WPx[:n], "WP_x")
        out["rel_l2_abs_Z_x"] = rel_l2_np(np.abs(w32.Zx[:m]), np.abs(w64.Zx[:m]), "|Z_x|")
    return out


def accuracy_leg(Nx, Ny, Nz, wakelength=0.5, snap=3000):
    """FP32 form against the FP64 kernel (bit-identical to the FP64 oracle at
    every size tested) on the full wakefield run of the default workload:
    WP_z(s) and Z_z(f) up to fmax = v/(3 sigma_z), E and H after `snap` steps
    and at the end. Independent of --steps."""
    import torch

    import wakis_amd as wa
    res, kept = {}, {}
    for dtype in (np.float32, np.float64):
        wake = wa.WakeSolver(q=1e-9, sigmaz=10e-3, beta=1.0, skip_cells=10, save=False,
                             verbose=0)
        s = build_solver(Nx, Ny, Nz, dtype, "auto", wake=wake)
        r = wake_run(s, wakelength, snap)
        E1, H1, k1 = r.pop("kept")
        kept[dtype] = (E1, H1, s.E.array.to(torch.float64, copy=True),
                       s.H.array.to(torch.float64, copy=True))
        r["wake"] = wake
        res[dtype] = r
        log(f"accuracy {np.dtype(dtype).name}: Nt={r['Nt']} loop {r['loop_s']:.2f}s "
            f"post {r['post_s']:.2f}s path {r['path']}")
        del s
        torch.cuda.empty_cache()
    k32, k64 = kept[np.float32], kept[np.float64]
    r32, r64 = res[np.float32], res[np.float64]
    out = {"quantity": "full wakefield run of the timed workload, 0.5 m wakelength (%d FP64 "
                       "steps), FP32 run vs FP64 run" % r64["Nt"],
           "reference": "the FP64 run: the FP64 kernel is bit-identical to the FP64 oracle "
                        "(scipy CSR restatement of the reference) at every size tested",
           "steps": {"fp32": r32["Nt"], "fp64": r64["Nt"]},
           "wake": _wake_compare(r32["wake"], r64["wake"]),
           "fields_at_%d_steps" % snap: {"rel_l2_E": _rel_l2(k32[0], k64[0]),
                                         "rel_l2_H": _rel_l2(k32[1], k64[1])},
           "fields_at_end": {"rel_l2_E": _rel_l2(k32[2], k64[2]),
                             "rel_l2_H": _rel_l2(k32[3], k64[3])},
           "step_loop_s": {"fp32": round(r32["loop_s"], 3), "fp64": round(r64["loop_s"], 3)},
           "wake_postprocessing_s": {"fp32": round(r32["post_s"], 3),
                                     "fp64": round(r64["post_s"], 3)},
           "finite": bool(r32["finite"] and r64["finite"])}
    return out


def pillbox_solver(dtype, dims=(200, 200, 400), sibc=True, sigma_wall=1e4, wake=None):
    """The lossy pillbox with PML: cylinder r = 40 mm, |z| < 30 mm, beam pipe
    r = 10 mm, a 3-cell wall shell of sigma = `sigma_wall` (eps_r = mu_r = 1)
    around the vacuum in a PEC background; 90 x 90 x 180 mm on `dims` cells,
    PEC x/y, 10-layer PML on both z ends. With `sibc` (the reference's
    default) the wall is replaced by its surface cells with sigma = eps =
    1/Z_s, Z_s = sqrt(pi fmax mu0 / sigma), fmax = 1 GHz (the WakeSolver's fmax when
    `wake` is given), and dt is the CFL
    step; without it the wall keeps its volume conductivity and dt is the
    reference's tau clamp (eps0 / sigma)."""
    import wakis_amd as wa
    from wakis_amd.solids import Cylinder
    Nx, Ny, Nz = dims
    x = np.linspace(-45e-3, 45e-3, Nx + 1)
    y = np.linspace(-45e-3, 45e-3, Ny + 1)
    z = np.linspace(-90e-3, 90e-3, Nz + 1)
    vac = (Cylinder(40e-3, -30e-3, 30e-3) | Cylinder(10e-3)).cell_mask(x, y, z)
    grown = vac.copy()
    for _ in range(3):  # 6-neighbour dilation: one wall cell per pass
        g = grown.copy()
        g[1:] |= grown[:-1]
        g[:-1] |= grown[1:]
        g[:, 1:] |= grown[:, :-1]
        g[:, :-1] |= grown[:, 1:]
        g[:, :, 1:] |= grown[:, :, :-1]
        g[:, :, :-1] |= grown[:, :, 1:]
        grown = g
    wall = grown & ~vac
    grid = wa.GridFIT3D(x=x, y=y, z=z, stl_solids={"wall": wall, "vac": vac},
                        stl_materials={"wall": [1.0, 1.0, sigma_wall], "vac": "vacuum"},
                        verbose=0)
    s = wa.SolverFIT3D(grid, wake=wake, bc_low=["pec", "pec", "pml"],
                       bc_high=["pec", "pec", "pml"], use_stl=True, bg="pec", dtype=dtype,
                       verbose=0, n_pml=10, use_sibc=sibc)
    s.fused_two_step = TWO_STEP[0]
    return s


def lossy_pml(dtype, steps=300, warmup=20, dims=(200, 200, 400), keep_fields=False, sibc=True):
    """Times `steps` leapfrog steps of the lossy pillbox (J = sigma E
    everywhere) with a beta = 0.8 beam (q = 1 nC, sigma_z = 15 mm) on axis,
    its table starting at 0.9 of the injection time so the bunch head is in
    the domain."""
    import torch

    import wakis_amd as wa
    from wakis_amd import _lib
    s = pillbox_solver(dtype, dims, sibc)
    beam = wa.Beam(q=1e-9, sigmaz=15e-3, beta=0.8)
    dev = s.device
    tot = warmup + steps
    beam._locate(s)
    t = 0.9 * beam.ti + np.arange(tot) * float(s.dt)
    table = torch.as_tensor(beam._line(s, t[:, None]).astype(s.dtype), device=dev).contiguous()
    s._set_ghosts_to_0()
    s.step_0 = False
    s._beam_line = (int(beam.ixs), int(beam.iys))

    def args(n0, k):
        a = s._step_args(k)
        a.ixs, a.iys = int(beam.ixs), int(beam.iys)
        a.jmode = _lib.WF_J_FULL
        a.beam_table, a.beam_row0, a.beam_rows = _lib.ptr(table), n0, tot
        return a

    def run(n0, k):
        s._run(args(n0, k))

    s.tune_fused(lambda k: args(0, k))  # workgroup shape by measurement (fields restored)
    run(0, warmup)
    torch.cuda.synchronize(dev)
    stream = torch.cuda.current_stream(dev)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    run(warmup, steps)
    e1.record(stream)
    torch.cuda.synchronize(dev)
    ms = e0.elapsed_time(e1) / steps
    b = 21 * np.dtype(dtype).itemsize
    out = dict(ms_per_step=round(ms, 5), gcell_s=round(s.N / (ms * 1e-3) / 1e9, 3),
               bytes_per_cell_update=b, path=s.last_path, dt=float(s.dt),
               achieved_GBs=round(s.N * b / (ms * 1e-3) / 1e9, 1),
               frac=round(s.N * b / (ms * 1e-3) / 8e12, 4),
               full_rows=s.fused_full_rows, tuning=_tuning(s))
    if keep_fields:
        out["E"] = s.E.array.to(torch.float64, copy=True)
        out["H"] = s.H.array.to(torch.float64, copy=True)
    del s, table
    torch.cuda.empty_cache()
    return out


def lossy_wake(dims=(200, 200, 400), wakelength=1.0, xsource=2e-3):
    """The pillbox's wake: a full wakesolve with beta = 0.8, sigma_z = 15 mm,
    the source line `xsource` off axis (the dipolar transverse wake WP_x on
    the axis) and the reference's default surface impedance, FP32 vs FP64."""
    import torch

    import wakis_amd as wa
    res = {}
    for dtype in (np.float32, np.float64):
        wake = wa.WakeSolver(q=1e-9, sigmaz=15e-3, beta=0.8, xsource=xsource, skip_cells=12,
                             save=False, verbose=0)
        s = pillbox_solver(dtype, dims, sibc=True, wake=wake)
        r = wake_run(s, wakelength)
        r.pop("kept")
        r["wake"] = wake
        r["ixs_ixt"] = (int(s._wake_beam.ixs), int(s.ixt))
        res[dtype] = r
        log(f"lossy wake {np.dtype(dtype).name}: Nt={r['Nt']} loop {r['loop_s']:.2f}s "
            f"post {r['post_s']:.2f}s")
        del s
        torch.cuda.empty_cache()
    r32, r64 = res[np.float32], res[np.float64]
    return {"quantity": "full wakesolve, %.1f m wakelength, beta=0.8, source line %.0f mm off "
                        "axis (x index %d; test line on axis, index %d), FP32 run vs FP64 run"
                        % (wakelength, xsource * 1e3, r64["ixs_ixt"][0], r64["ixs_ixt"][1]),
            "sibc": "use_sibc=True (the reference's default); a solver built with a "
                    "WakeSolver takes the wake's fmax = v/(3 sigma_z) = %.3g GHz for the "
                    "surface impedance (solverFIT3D.py:177-178); dt = %.4g s (CFL)"
                    % (r64["wake"].fmax / 1e9, r64["dt"]),
            "steps": {"fp32": r32["Nt"], "fp64": r64["Nt"]},
            "wake": _wake_compare(r32["wake"], r64["wake"]),
            "step_loop_s": {"fp32": round(r32["loop_s"], 3), "fp64": round(r64["loop_s"], 3)},
            "step_loop_gcell_s": {"fp32": round(r32["N"] * r32["Nt"] / r32["loop_s"] / 1e9, 3),
                                  "fp64": round(r64["N"] * r64["Nt"] / r64["loop_s"] / 1e9, 3)},
            "wake_postprocessing_s": {"fp32": round(r32["post_s"], 3),
                                      "fp64": round(r64["post_s"], 3)},
            "finite": bool(r32["finite"] and r64["finite"])}


def lossy_block():
    import torch
    l32 = lossy_pml(np.float32, keep_fields=True)
    l64 = lossy_pml(np.float64, keep_fields=True)
    out = {"workload": "lossy pillbox r=40 mm, 3-cell sigma=1e4 S/m wall, 10-layer PML on z, "
                       "200x200x400 cells, beam sigma_z=15 mm beta=0.8, 300 timed steps",
           "parity": "both conductivity policies run bit-identically to the oracle "
                     "(tests/test_gpu_configs.py, tests/test_gpu_wake_branches.py); the "
                     "volume conductivity (use_sibc=False) is the reference's algorithm "
                     "throughout, the surface-impedance cells (use_sibc=True) follow the "
                     "oracle's restatement of pyvista's gradient, whose edge and corner cells "
                     "are not pinned",
           "sibc": {"dt_policy": "use_sibc=True (the reference default): wall surface cells "
                                 "with sigma = eps = 1/Z_s at fmax = 1 GHz; dt = %.4g s (CFL, "
                                 "no tau clamp)" % l64["dt"],
                    "fp32": {k: v for k, v in l32.items() if k not in ("E", "H", "dt")},
                    "fp64": {k: v for k, v in l64.items() if k not in ("E", "H", "dt")},
                    "rel_l2_E_fp32_vs_fp64": _rel_l2(l32["E"], l64["E"]),
                    "rel_l2_H_fp32_vs_fp64": _rel_l2(l32["H"], l64["H"])}}
    del l32, l64
    torch.cuda.empty_cache()
    log(f"lossy sibc {out['sibc']['fp32']['gcell_s']} / {out['sibc']['fp64']['gcell_s']} Gcell/s")
    v32 = lossy_pml(np.float32, keep_fields=True, sibc=False)
    v64 = lossy_pml(np.float64, keep_fields=True, sibc=False)
    out["volume"] = {"dt_policy": "use_sibc=False: the wall keeps sigma = 1e4 S/m, dt = %.4g s "
                                  "(the reference's tau clamp eps0/sigma)" % v64["dt"],
                     "fp32": {k: v for k, v in v32.items() if k not in ("E", "H", "dt")},
                     "fp64": {k: v for k, v in v64.items() if k not in ("E", "H", "dt")},
                     "rel_l2_E_fp32_vs_fp64": _rel_l2(v32["E"], v64["E"]),
                     "rel_l2_H_fp32_vs_fp64": _rel_l2(v32["H"], v64["H"])}
    del v32, v64
    torch.cuda.empty_cache()
    log(f"lossy volume {out['volume']['fp32']['gcell_s']} Gcell/s")
    out["wake"] = lossy_wake()
    return out


def _tuning(s):
    """ms per step of each workgroup shape tune_fused() timed (-1: row segments,
    N: whole-row workgroups of N rows)."""
    t = s.fused_tuning or {}
    return {str(k): round(v, 4) for k, v in sorted(t.items())}


def batch_sweep(n=64, steps=2000, seed=0, dtype=np.float32):
    """The batched sweep: n box cavities (wakis_amd.batch.random_box_cavities)
    with about `steps` wakefield steps each, run one after another with the
    single-grid kernel and then all together with wakesolve_batch; both timed
    end to end (set-up, stepping, wake post-processing). Then the forward +
    adjoint of each cavity's loss factor with respect to the eps_r of its
    dielectric block, one problem after another (WakeAdjoint) and all together
    (wake_adjoint_batch)."""
    import torch

    from wakis_amd import batch
    specs = batch.random_box_cavities(seed, n=n)

    def make():
        sv = [batch.cavity_solver(sp, dtype=dtype) for sp in specs]
        wl = [steps * float(s.dt) * s.wake.v - s.wake.ti * s.wake.v - (s.z.max() - s.z.min())
              for s in sv]
        return sv, wl

    seq, wl = make()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s, w in zip(seq, wl):
        s.wakesolve(wakelength=w)
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    cells = sum(s.N * s.Nt for s in seq)
    bat, wl = make()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    info = batch.wakesolve_batch(bat, wl)
    torch.cuda.synchronize()
    t_bat = time.perf_counter() - t0
    same = all(np.array_equal(a.wake.WP_3d, b.wake.WP_3d) for a, b in zip(seq, bat))
    nt_range = [int(min(s.Nt for s in seq)), int(max(s.Nt for s in seq))]
    cells_total = int(sum(s.N for s in seq))
    del seq, bat
    torch.cuda.empty_cache()
    log(f"batch forward seq {t_seq:.2f}s batched {t_bat:.2f}s")
    # forward + adjoint of the loss factor: d k_loss / d eps_r over each dielectric block
    from wakis_amd.adjoint import WakeAdjoint

    def make_adj():
        sv, wls = make()
        return [WakeAdjoint(s, batch.dielectric_region(sp), w, w_loss=1.0, w_band=0.0)
                for s, sp, w in zip(sv, specs, wls)]

    adjs = make_adj()
    ers = [np.full(adj.shape, sp["eps_r"]) for adj, sp in zip(adjs, specs)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ref = []
    for adj, er in zip(adjs, ers):
        L = adj.forward(er)
        ref.append((L, adj.backward()[0]))
    torch.cuda.synchronize()
    t_seq_adj = time.perf_counter() - t0
    ad_cells = 2 * sum(adj.s.N * adj.s.Nt for adj in adjs)  # forward + reverse steps
    del adjs
    torch.cuda.empty_cache()
    adjs = make_adj()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses, grads = batch.wake_adjoint_batch(adjs, ers)
    torch.cuda.synchronize()
    t_bat_adj = time.perf_counter() - t0
    same_adj = all(L == Lb and torch.equal(g, gb[0])
                   for (L, g), Lb, gb in zip(ref, losses, grads))
    log(f"batch adjoint seq {t_seq_adj:.2f}s batched {t_bat_adj:.2f}s")
    adjoint = {"quantity": "k_loss and d k_loss / d eps_r of every dielectric-block cell",
               "cell_updates": int(ad_cells),
               "sequential": round(ad_cells / t_seq_adj / 1e9, 3),
               "batched": round(ad_cells / t_bat_adj / 1e9, 3),
               "speedup": round(t_seq_adj / t_bat_adj, 3), "bit_identical": bool(same_adj)}
    return {"metric": "batched wakefield sweep throughput", "unit": "Gcell-updates/s",
            "grids": n, "steps_per_grid": nt_range,
            "cells_total": cells_total, "cell_updates": int(cells),
            "sequential": round(cells / t_seq / 1e9, 3), "batched": round(cells / t_bat / 1e9, 3),
            "speedup": round(t_seq / t_bat, 3),
            "batched_step_loop": round(info["cell_updates"] * info["main_steps"] / info["steps"]
                                       / (info["ms_main"] * 1e-3) / 1e9, 3),
            "shape": info["shape"], "wp_bit_identical": bool(same), "dtype": np.dtype(dtype).name,
            "adjoint": adjoint,
            "data": "synthetic (random_box_cavities, seed %d)" % seed}


def adjoint_cost(dtype=np.float32, dims=(128, 128, 256), insert=(32, 32, 24), steps=4000,
                 seed=0):
    """Reverse-mode cost on the differentiable workload: the 50 mm PEC cavity
    with pipe on `dims`, a dielectric/lossy block of `insert` cells on the axis
    with per-cell eps_r ~ U[1, 10] and sigma ~ U[0, 2] S/m (seeded), `steps`
    wakefield steps at the CFL dt (tau clamp off). Times one plain forward
    (wakesolve), then the adjoint's forward (keeping the block's trajectory)
    plus its backward, each bracketed by device synchronisation; then the
    same gradient with checkpoints and segment recompute."""
    import torch

    import wakis_amd as wa
    from wakis_amd.adjoint import WakeAdjoint
    from wakis_amd.solids import Box, Cylinder
    Nx, Ny, Nz = dims
    x = np.linspace(-30e-3, 30e-3, Nx + 1)
    y = np.linspace(-30e-3, 30e-3, Ny + 1)
    z = np.linspace(-80e-3, 80e-3, Nz + 1)
    cav = (Box(-25e-3, 25e-3, -25e-3, 25e-3, -25e-3, 25e-3) | Cylinder(5e-3)).cell_mask(x, y, z)
    lo = [(n - m) // 2 for n, m in zip(dims, insert)]
    region = tuple((a, a + m) for a, m in zip(lo, insert))
    ins = np.zeros(dims, dtype=bool)
    ins[region[0][0]:region[0][1], region[1][0]:region[1][1], region[2][0]:region[2][1]] = True
    rng = np.random.default_rng(seed)
    er = rng.uniform(1.0, 10.0, insert)
    sg = rng.uniform(0.0, 2.0, insert)

    def solver():
        grid = wa.GridFIT3D(x=x, y=y, z=z, stl_solids={"cav": cav, "ins": ins},
                            stl_materials={"cav": "vacuum", "ins": [2.0, 1.0, 1.0]}, verbose=0)
        wake = wa.WakeSolver(q=1e-9, sigmaz=10e-3, beta=1.0, skip_cells=10, save=False,
                             verbose=0)
        return wa.SolverFIT3D(grid, wake=wake, bc_low=["pec", "pec", "abc"],
                              bc_high=["pec", "pec", "abc"], use_stl=True, bg="pec",
                              dtype=dtype, verbose=0, use_sibc=False, tau_clamp=False)

    s = solver()
    v = s.wake.v
    wl = steps * float(s.dt) * v - s.wake.ti * v - (s.z.max() - s.z.min())
    adj = WakeAdjoint(s, region, wl, w_loss=1.0, w_band=1e-13)
    adj._set_params(er, sg)
    s.wakesolve(wakelength=wl)  # untimed: first-call set-up (material tables, kernels)
    adj._set_params(er, sg)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.wakesolve(wakelength=wl)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    L = adj.forward(er, sg)
    torch.cuda.synchronize()
    t_fc = time.perf_counter() - t0
    t0 = time.perf_counter()
    ge, gs = adj.backward()
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t0
    mode, kernel, adt = adj.used_mode, adj.used_kernel, adj.used_adjoint_dtype
    adj = None
    # the same gradient with checkpoints and segment recompute
    ck = WakeAdjoint(s, region, wl, w_loss=1.0, w_band=1e-13, mode="checkpoint")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Lc = ck.forward(er, sg)
    gec, gsc = ck.backward()
    torch.cuda.synchronize()
    t_ck = time.perf_counter() - t0
    return {"workload": "cavity+pipe %dx%dx%d, %dx%dx%d eps_r/sigma block, %d steps, %s"
                        % (dims + insert + (s.Nt, np.dtype(dtype).name)),
            "mode": mode, "kernel": kernel, "adjoint_fields": adt, "fwd_s": round(t_fwd, 3),
            "fwd_with_states_s": round(t_fc, 3),
            "bwd_s": round(t_bwd, 3), "adjoint_over_fwd": round((t_fc + t_bwd) / t_fwd, 3),
            "checkpoint_mode_over_fwd": round(t_ck / t_fwd, 3),
            "modes_bitwise_equal": bool(Lc == L and torch.equal(ge, gec) and torch.equal(gs, gsc)),
            "loss": float(L), "grad_finite": bool(torch.isfinite(ge).all() and
                                                 torch.isfinite(gs).all())}


def _threads_info():
    try:
        import threadpoolctl
        return [{k: d.get(k) for k in ("user_api", "internal_api", "num_threads", "version")}
                for d in threadpoolctl.threadpool_info()]
    except Exception as e:  # threadpoolctl missing or failing: say so
        return [{"error": repr(e)}]


def cpu_step_rate(Nx, Ny, Nz, dtype, steps, warmup=2):
    """The oracle's step (the reference's scipy CSR assembly and
    (dt*M1)@E / M2@H step, beam line written every step) on the cavity/pipe
    workload: Gcell-updates/s over `steps` steps after `warmup`."""
    from oracle import cases
    from oracle import fit_oracle as fo
    t0 = time.perf_counter()
    s = cases.cavity_pipe_solver(Nx, Ny, Nz, dtype=dtype)
    setup = time.perf_counter() - t0
    beam = fo.OracleBeam(q=1e-9, sigmaz=10e-3, beta=1.0)
    for n in range(warmup):
        beam.update(s, n * s.dt)
        s.one_step()
    t0 = time.perf_counter()
    for n in range(warmup, warmup + steps):
        beam.update(s, n * s.dt)
        s.one_step()
    el = time.perf_counter() - t0
    N = s.N
    del s
    return dict(N=N, steps=steps, seconds=el, setup_s=setup, gcell_s=N * steps / el / 1e9)


def cpu_config1_wakesolve():
    """The reference's CPU-runnable case end to end with the oracle: the
    48x48x128 cavity/pipe wakesolve (0.5 m, beta = 1, FP64), step loop and
    wake post-processing (WP_3d, WP_x/y, Z up to 5 GHz) timed apart."""
    from oracle import cases
    from oracle import fit_oracle as fo
    from oracle import wake_oracle as wo
    t0 = time.perf_counter()
    s = cases.cavity_pipe_solver(48, 48, 128, dtype=np.float64)
    setup = time.perf_counter() - t0
    q, sigmaz, beta, wl = 1e-9, 10e-3, 1.0, 0.5
    ti = wo.wake_ti(sigmaz, beta)
    t0 = time.perf_counter()
    cap = fo.wakesolve_fields(s, wl, q, sigmaz, beta, 0.0, 0.0, 0.0, 0.0, ti)
    t_loop = time.perf_counter() - t0
    t0 = time.perf_counter()
    wo.wake_solve(cap, q, sigmaz, beta, wl, skip_cells=10, fmax=5e9)
    t_post = time.perf_counter() - t0
    return {"workload": "48x48x128 cavity/pipe wakesolve, 0.5 m, %d steps, FP64" % cap["Nt"],
            "setup_s": round(setup, 2), "step_loop_s": round(t_loop, 2),
            "wake_postprocessing_s": round(t_post, 2),
            "gcell_s": round(s.N * cap["Nt"] / t_loop / 1e9, 6)}


def main():
    # the JSON line is the only thing on stdout: the solvers' progress prints go to stderr
    with contextlib.redirect_stdout(sys.stderr):
        line = run(sys.argv[1:])
    if line is not None:
        print(json.dumps(line, allow_nan=False), flush=True)


def run(argv):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--grid", default="256,256,512")
    ap.add_argument("--path", default="auto")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-grid", default="256,256,512",
                    help="grid of the CPU step-rate sample (the timed workload's by default)")
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--variant", type=int, default=0,
                    help="fused kernel: 0 auto, 1 scalar, 2/3/7/8 rows per workgroup")
    ap.add_argument("--kz", type=int, default=0, help="z-planes per workgroup (0 auto)")
    ap.add_argument("--two-step", type=int, default=1, choices=(0, 1, 4, 8),
                    help="two leapfrog steps per launch where that kernel applies (0: one "
                         "step per launch; 4, 8: rows per workgroup)")
    ap.add_argument("--no-exact", action="store_true",
                    help="skip the timing of the exact general path")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the FP32-vs-FP64 full wakefield run")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="print only the reverse-mode cost (FP32 and FP64) as one JSON line")
    ap.add_argument("--adjoint-steps", type=int, default=4000,
                    help="wakefield steps of the reverse-mode workload")
    ap.add_argument("--no-lossy", action="store_true",
                    help="skip the lossy pillbox with PML (200x200x400)")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the reverse-mode cost measurement")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-cavity batched sweep")
    ap.add_argument("--batch", action="store_true",
                    help="only the batched sweep of 64 box cavities (one JSON line)")
    args = ap.parse_args(argv)
    TWO_STEP[0] = args.two_step
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this engine runs on one GPU; runs on several GPUs are not "
              "supported (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.batch:
        return batch_sweep()
    if args.adjoint_only:
        return {"fp32": adjoint_cost(steps=args.adjoint_steps),
                "fp64": adjoint_cost(np.float64, steps=args.adjoint_steps)}
    Nx, Ny, Nz = (int(v) for v in args.grid.split(","))
    import torch
    copy_gbs = hbm_copy_gbs(torch.device("cuda", 0))
    log(f"hbm copy {copy_gbs:.1f} GB/s")
    log(f"fp32 {Nx}x{Ny}x{Nz} steps={args.steps} warmup={args.warmup}")
    r32 = time_gpu(Nx, Ny, Nz, np.float32, args.steps, args.warmup, args.path, args.variant,
                   args.kz)
    log(f"fp32 {r32}")
    r64 = None
    if not args.no_fp64:
        r64 = time_gpu(Nx, Ny, Nz, np.float64, args.steps, args.warmup, args.path,
                       args.variant, args.kz)
        log(f"fp64 {r64}")
    exact = None
    if not args.no_exact:
        exact = {"quantity": "the exact general path (two kernels per step, flat-index wrap "
                             "terms kept: what runs with live wrap terms, e.g. periodic or "
                             "PMC x/y faces with material on them) on the timed workload, "
                             "50 steps after 5",
                 "bytes_per_cell_update_read_written": {"fp32": 96, "fp64": 192}}
        for dt_, key in ((np.float32, "fp32"), (np.float64, "fp64")):
            r = time_gpu(Nx, Ny, Nz, dt_, 50, 5, "general")
            exact[key] = {"ms_per_step": round(r["ms_per_step"], 5),
                          "gcell_s": round(r["gcell_s"], 3), "path": r["path"]}
        log(f"exact path {exact}")
    acc = None
    if not args.no_accuracy and not args.no_fp64:
        acc = accuracy_leg(Nx, Ny, Nz)
        log(f"accuracy {json.dumps(acc['wake'])}")
    lossy = None
    if not args.no_lossy:
        lossy = lossy_block()
        log(f"lossy wake {json.dumps(lossy['wake']['wake'])}")
    adj = adj64 = None
    if not args.no_adjoint:
        adj = adjoint_cost(steps=args.adjoint_steps)
        log(f"adjoint {adj}")
        adj64 = adjoint_cost(np.float64, steps=args.adjoint_steps)
        log(f"adjoint fp64 {adj64}")
    bat = None
    if not args.no_batch:
        bat = batch_sweep()
        log(f"batch {bat['sequential']} -> {bat['batched']} Gcell/s")
    cpu = cfg1 = None
    if not args.no_cpu:
        cx, cy, cz = (int(v) for v in args.cpu_grid.split(","))
        log(f"cpu step rate on {cx}x{cy}x{cz}, {args.cpu_steps} steps")
        cpu = cpu_step_rate(cx, cy, cz, np.float32, args.cpu_steps)
        log(f"cpu {cpu}")
        cfg1 = cpu_config1_wakesolve()
        log(f"cpu config 1 {cfg1}")

    line = {
        "metric": METRIC,
        "value": round(r32["gcell_s"], 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"wakefield leapfrog, PEC cavity 50mm + pipe r=5mm, "
                               f"{Nx}x{Ny}x{Nz} cells, beam on axis, PEC x/y ABC z",
                   "cells": r32["N"], "path": r32["path"], "full_rows": r32["full_rows"]},
        "roofline": roofline(r32, (Nx, Ny, Nz), "f32", args.variant, args.kz, copy_gbs),
    }
    if r64 is not None:
        line["fp64"] = {"value": round(r64["gcell_s"], 3), "unit": "Gcell-updates/s",
                        "ms_per_step": round(r64["ms_per_step"], 5), "steps": r64["steps"],
                        "achieved_GBs": round(r64["achieved_GBs"], 1),
                        "frac": round(r64["achieved_GBs"] / HBM_PEAK_GBS, 4),
                        "traffic": measured_traffic((Nx, Ny, Nz), "f64", r64["path"],
                                                    args.variant, args.kz,
                                                    2 if r64.get("two_step") else 1),
                        "steps_per_launch": 2 if r64.get("two_step") else 1,
                        "kernel": "k_fused4_t2" if r64.get("two_step") else "k_fused4"}
        t64, spl64 = line["fp64"]["traffic"], line["fp64"]["steps_per_launch"]
        if t64:  # the real HBM rate (traffic per launch over the launch time)
            line["fp64"]["actual_GBs"] = round(t64 / (r64["ms_per_step"] * spl64 * 1e-3) / 1e9, 1)
        line["fp32_over_fp64"] = round(r32["gcell_s"] / r64["gcell_s"], 3)
    if exact is not None:
        line["exact_path"] = exact
    if acc is not None:
        line["accuracy"] = acc
    if lossy is not None:
        line["lossy_pml"] = lossy
    if adj is not None:
        line["adjoint"] = adj
        line["adjoint_fp64"] = adj64
    if bat is not None:
        line["batch"] = bat
    if cpu is not None:
        line["cpu_baseline"] = {
            "value": round(cpu["gcell_s"], 6), "unit": "Gcell-updates/s", "cores": 1,
            "kind": "port",
            "sample": f"the oracle's FP32 step (the reference's scipy CSR algorithm) on the "
                      f"timed workload's {args.cpu_grid} cells: {cpu['steps']} steps after 2 "
                      f"warm-up, {cpu['seconds']:.1f} s (set-up {cpu['setup_s']:.1f} s, not "
                      f"timed); scipy's CSR matvec and numpy's elementwise operations run on "
                      f"one thread (no BLAS call in the step); host os.cpu_count()="
                      f"{os.cpu_count()}",
            "threadpools": _threads_info(),
            "config1_wakesolve": cfg1}
    return line


if __name__ == "__main__":
    main()
