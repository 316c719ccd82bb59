"""Throughput of the FIT leapfrog (wakefield step loop) on one MI355X.

Workload (N=1): PEC cubic cavity 50 mm with a round beam pipe r=5 mm along z
in a 60x60x160 mm domain on a uniform 256x256x512 grid (33.5M cells), vacuum
inside, PEC background, PEC on x/y and ABC on z; Gaussian beam q=1e-9 C,
sigma_z=10 mm, beta=1 on axis. A step is one leapfrog step over the whole grid
as wakesolve runs it: beam J line from the precomputed table, H and E
updates, Ez 3x3 window capture. Fields start at zero (synthetic input, no
data files). Timed with HIP events on the stream the kernels run on.

Prints one JSON line. --gpus N>1 exits: this engine runs on one GPU.
"""

import argparse
import json
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
# (profiles/k_fused4_cavity_pipe_256x256x512.md): FETCH_SIZE is scaled by 2.0,
# the ratio of the bytes a 1 GiB device copy reads to what the counter reports
# for it; WRITE_SIZE matches the copy's bytes and is taken as is. Keyed by
# (grid, dtype, path, variant, kz); other settings report null.
MEASURED_TRAFFIC = {
    ((256, 256, 512), "f32", "fused", 0, 0): {"fetch_B": 926_966_374, "write_B": 805_458_637},
    ((256, 256, 512), "f64", "fused", 0, 0): {"fetch_B": 2_184_980_070, "write_B": 1_610_732_749},
}


def measured_traffic(grid, dtype, path, variant, kz):
    t = MEASURED_TRAFFIC.get((tuple(grid), dtype, path, variant, kz))
    return None if t is None else t["fetch_B"] + t["write_B"]


def build_solver(Nx, Ny, Nz, dtype, path):
    import wakis_amd as wa
    from wakis_amd.solids import Box, Cylinder
    solid = Box(-25e-3, 25e-3, -25e-3, 25e-3, -25e-3, 25e-3) | Cylinder(5e-3)
    grid = wa.GridFIT3D(-30e-3, 30e-3, -30e-3, 30e-3, -80e-3, 80e-3, Nx, Ny, Nz,
                        stl_solids={"cavity": solid}, stl_materials={"cavity": "vacuum"},
                        verbose=0)
    return wa.SolverFIT3D(grid, bc_low=["pec", "pec", "abc"], bc_high=["pec", "pec", "abc"],
                          use_stl=True, bg="pec", dtype=dtype, verbose=0, path=path)


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

    def run(n0, k):
        a = s._step_args(k)
        a.ixs, a.iys = int(beam.ixs), int(beam.iys)
        a.jmode = _lib.WF_J_LINE
        a.beam_table, a.beam_row0, a.beam_rows = _lib.ptr(table), n0, tot
        a.capture = _lib.ptr(cap)
        a.cap_ix0, a.cap_iy0, a.cap_nx, a.cap_ny = ixt - 1, iyt - 1, 3, 3
        a.cap_samples, a.cap_sample0 = tot, n0 + 1
        s._run(a)

    run(0, warmup)
    torch.cuda.synchronize(dev)
    used_path = s.last_path
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
    out = dict(N=N, steps=steps, ms_total=ms, ms_per_step=ms / steps, wall_s=wall, path=used_path,
               gcell_s=N * steps / (ms * 1e-3) / 1e9, finite=finite,
               bytes_per_cell=VALUES_PER_CELL * np.dtype(dtype).itemsize)
    out["achieved_GBs"] = N * out["bytes_per_cell"] / (out["ms_per_step"] * 1e-3) / 1e9
    del s, table, cap
    torch.cuda.empty_cache()
    return out


def time_cpu(Nx, Ny, Nz, dtype, steps, warmup):
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
    return dict(N=s.N, steps=steps, seconds=el, setup_s=setup,
                gcell_s=s.N * steps / el / 1e9)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--grid", default="256,256,512")
    ap.add_argument("--path", default="auto")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-grid", default="128,128,256")
    ap.add_argument("--cpu-steps", type=int, default=120)
    ap.add_argument("--variant", type=int, default=0,
                    help="fused kernel: 0 auto, 1 scalar, 2/3/7/8 rows per workgroup")
    ap.add_argument("--kz", type=int, default=0, help="z-planes per workgroup (0 auto)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this engine runs on one GPU; runs on several GPUs are not "
              "supported (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    Nx, Ny, Nz = (int(v) for v in args.grid.split(","))
    import torch
    copy_gbs = hbm_copy_gbs(torch.device("cuda", 0))
    print(f"[bench] hbm copy {copy_gbs:.1f} GB/s", file=sys.stderr, flush=True)
    print(f"[bench] fp32 {Nx}x{Ny}x{Nz} steps={args.steps} warmup={args.warmup}",
          file=sys.stderr, flush=True)
    r32 = time_gpu(Nx, Ny, Nz, np.float32, args.steps, args.warmup, args.path, args.variant,
                   args.kz)
    print(f"[bench] fp32 {r32}", file=sys.stderr, flush=True)
    r64 = None
    if not args.no_fp64:
        k64 = max(1, args.steps // 3)
        r64 = time_gpu(Nx, Ny, Nz, np.float64, k64, args.warmup, args.path, args.variant,
                       args.kz)
        print(f"[bench] fp64 {r64}", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu:
        cx, cy, cz = (int(v) for v in args.cpu_grid.split(","))
        cpu = time_cpu(cx, cy, cz, np.float32, args.cpu_steps, 2)
        print(f"[bench] cpu {cpu}", file=sys.stderr, flush=True)

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
                   "cells": r32["N"], "path": r32["path"]},
        "roofline": {"bound": "hbm", "achieved": round(r32["achieved_GBs"], 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(r32["achieved_GBs"] / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic((Nx, Ny, Nz), "f32", r32["path"], args.variant,
                                                 args.kz),
                     "bytes_per_cell_update": r32["bytes_per_cell"],
                     "hbm_copy_GBs": round(copy_gbs, 1)},
    }
    if r64 is not None:
        line["fp64"] = {"value": round(r64["gcell_s"], 3), "unit": "Gcell-updates/s",
                        "ms_per_step": round(r64["ms_per_step"], 5), "steps": r64["steps"],
                        "achieved_GBs": round(r64["achieved_GBs"], 1),
                        "frac": round(r64["achieved_GBs"] / HBM_PEAK_GBS, 4),
                        "traffic": measured_traffic((Nx, Ny, Nz), "f64", r64["path"],
                                                    args.variant, args.kz)}
        line["fp32_over_fp64"] = round(r32["gcell_s"] / r64["gcell_s"], 3)
    if cpu is not None:
        line["cpu_baseline"] = {
            "value": round(cpu["gcell_s"], 6), "unit": "Gcell-updates/s", "cores": 1,
            "kind": "port",
            "sample": f"oracle scipy CSR FP32 step (reference algorithm) on {args.cpu_grid} "
                      f"cells, {cpu['steps']} steps after 2 warm-up, {cpu['seconds']:.1f}s, "
                      f"host os.cpu_count()={os.cpu_count()}, single-threaded scipy/numpy"}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
