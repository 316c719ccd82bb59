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
    ((256, 256, 512), "f32", "fused", 0, 0): {"fetch_B": 936_804_147, "write_B": 805_472_666},
    ((256, 256, 512), "f64", "fused", 0, 0): {"fetch_B": 2_007_433_830, "write_B": 1_610_869_248},
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


def time_gpu(Nx, Ny, Nz, dtype, steps, warmup, path, variant=0, kz=0, wake_check=False,
             keep_fields=False):
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
        s._run(args(n0, k))

    s.tune_fused(lambda k: args(0, k))  # workgroup shape by measurement (fields restored)
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
               bytes_per_cell=VALUES_PER_CELL * np.dtype(dtype).itemsize,
               full_rows=s.fused_full_rows, tuning=_tuning(s))
    out["achieved_GBs"] = N * out["bytes_per_cell"] / (out["ms_per_step"] * 1e-3) / 1e9
    if wake_check:
        out["WP"] = wake_of_capture(s, cap, tot)
    if keep_fields:  # final fields in FP64, for the FP32-vs-FP64 field check
        out["E"] = s.E.array.to(torch.float64)
        out["H"] = s.H.array.to(torch.float64)
    del s, table, cap
    torch.cuda.empty_cache()
    return out


WAKE_CHECK_LENGTH = 0.1  # m of s axis for the truncated-integral comparison


def wake_of_capture(s, cap, nsamples):
    """WP_z(s) on the beam axis from the 3x3 Ez capture of a timed run, with the
    device WakeSolver kernels (FP64 line integral), over s in
    [-ti*v, WAKE_CHECK_LENGTH): the integral truncated at the last captured
    sample."""
    import wakis_amd as wa
    wake = wa.WakeSolver(q=1e-9, sigmaz=10e-3, beta=1.0, skip_cells=10, save=False,
                         verbose=0)
    wake.wakelength = WAKE_CHECK_LENGTH
    ixt = int(np.abs(s.x).argmin())
    iyt = int(np.abs(s.y).argmin())
    t = np.arange(nsamples, dtype=np.float64) * float(s.dt)
    wake._set_capture(cap, t, s.x[ixt - 1:ixt + 2], s.y[iyt - 1:iyt + 2], s.z, s.dx, s.dy,
                      s.dz, s.dtype, nsamples)
    wake.calc_long_WP_3d()
    return wake.WP


def _brief(r):
    return {k: v for k, v in r.items() if k not in ("WP", "E", "H")}


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


def lossy_pml(dtype, steps=300, warmup=20, dims=(200, 200, 400), keep_fields=False, sibc=True):
    """The lossy pillbox with PML: cylinder r = 40 mm, |z| < 30 mm, beam pipe
    r = 10 mm, a 3-cell wall shell of sigma = 1e4 S/m (eps_r = mu_r = 1) around
    the vacuum in a PEC background; 90 x 90 x 180 mm on `dims` cells, PEC x/y,
    10-layer PML on both z ends; beam q = 1 nC, sigma_z = 15 mm, beta = 0.8 on
    axis, its table starting at 0.9 of the injection time so the bunch head is
    in the domain. With `sibc` (the reference's default) the wall is replaced
    by its surface cells with sigma = eps = 1/Z_s, Z_s = sqrt(pi fmax mu0 /
    sigma), fmax = 1 GHz, and dt is the CFL step; without it the wall keeps its
    volume conductivity and dt is the reference's tau clamp (eps0 / sigma).
    Times `steps` leapfrog steps (J = sigma E everywhere)."""
    import torch

    import wakis_amd as wa
    from wakis_amd import _lib
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
                        stl_materials={"wall": [1.0, 1.0, 1e4], "vac": "vacuum"}, verbose=0)
    s = wa.SolverFIT3D(grid, bc_low=["pec", "pec", "pml"], bc_high=["pec", "pec", "pml"],
                       use_stl=True, bg="pec", dtype=dtype, verbose=0, n_pml=10, use_sibc=sibc)
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
        out["E"] = s.E.array.to(torch.float64)
        out["H"] = s.H.array.to(torch.float64)
    del s, table
    torch.cuda.empty_cache()
    return out


def _tuning(s):
    """ms per step of each workgroup shape tune_fused() timed (-1: row segments,
    N: whole-row workgroups of N rows)."""
    t = s.fused_tuning or {}
    return {str(k): round(v, 4) for k, v in sorted(t.items())}


def _rel_l2(a, b):
    import torch
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


def adjoint_cost(dtype=np.float32, dims=(128, 128, 256), insert=(32, 32, 24), steps=4000,
                 seed=0):
    """Reverse-mode cost on the differentiable workload: the 50 mm PEC cavity
    with pipe on `dims`, a dielectric/lossy block of `insert` cells on the axis
    with per-cell eps_r ~ U[1, 10] and sigma ~ U[0, 2] S/m (seeded), `steps`
    wakefield steps at the CFL dt (tau clamp off). Times one plain forward
    (wakesolve), then the adjoint's forward (keeping the block's trajectory)
    plus its backward, each bracketed by device synchronisation; then the
    same gradient with sqrt(Nt) checkpoints and segment recompute."""
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
    # the same gradient with sqrt(Nt) checkpoints and segment recompute
    ck = WakeAdjoint(s, region, wl, w_loss=1.0, w_band=1e-13, mode="checkpoint")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Lc = ck.forward(er, sg)
    gec, gsc = ck.backward()
    torch.cuda.synchronize()
    t_ck = time.perf_counter() - t0
    return {"workload": "cavity+pipe %dx%dx%d, %dx%dx%d eps_r/sigma block, %d steps, %s"
                        % (dims + insert + (s.Nt, np.dtype(dtype).name)),
            "mode": mode, "kernel": kernel, "adjoint_fields": adt, "fwd_s": round(t_fwd, 3), "fwd_with_states_s": round(t_fc, 3),
            "bwd_s": round(t_bwd, 3), "adjoint_over_fwd": round((t_fc + t_bwd) / t_fwd, 3),
            "checkpoint_mode_over_fwd": round(t_ck / t_fwd, 3),
            "modes_bitwise_equal": bool(Lc == L and torch.equal(ge, gec) and torch.equal(gs, gsc)),
            "loss": float(L), "grad_finite": bool(torch.isfinite(ge).all() and
                                                 torch.isfinite(gs).all())}


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
    ap.add_argument("--no-wake-check", action="store_true",
                    help="skip the FP32-vs-FP64 WP_z comparison (FP64 then runs steps/3)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="print only the reverse-mode cost (FP32 and FP64) as one JSON line")
    ap.add_argument("--adjoint-steps", type=int, default=4000,
                    help="wakefield steps of the reverse-mode workload")
    ap.add_argument("--no-lossy", action="store_true",
                    help="skip the lossy pillbox with PML (200x200x400)")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the reverse-mode cost measurement")
    ap.add_argument("--batch", action="store_true",
                    help="only the batched sweep of 64 box cavities (one JSON line)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this engine runs on one GPU; runs on several GPUs are not "
              "supported (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.batch:
        print(json.dumps(batch_sweep()), flush=True)
        return
    if args.adjoint_only:
        out = {"fp32": adjoint_cost(steps=args.adjoint_steps),
               "fp64": adjoint_cost(np.float64, steps=args.adjoint_steps)}
        print(json.dumps(out), flush=True)
        return
    Nx, Ny, Nz = (int(v) for v in args.grid.split(","))
    import torch
    copy_gbs = hbm_copy_gbs(torch.device("cuda", 0))
    print(f"[bench] hbm copy {copy_gbs:.1f} GB/s", file=sys.stderr, flush=True)
    print(f"[bench] fp32 {Nx}x{Ny}x{Nz} steps={args.steps} warmup={args.warmup}",
          file=sys.stderr, flush=True)
    check = not args.no_fp64 and not args.no_wake_check
    r32 = time_gpu(Nx, Ny, Nz, np.float32, args.steps, args.warmup, args.path, args.variant,
                   args.kz, wake_check=check, keep_fields=check)
    print(f"[bench] fp32 {_brief(r32)}", file=sys.stderr, flush=True)
    r64 = None
    if not args.no_fp64:
        # same number of steps as FP32 when the wake check runs (the FP64 run is the
        # reference for the FP32 wake), else a third of them
        k64 = args.steps if check else max(1, args.steps // 3)
        r64 = time_gpu(Nx, Ny, Nz, np.float64, k64, args.warmup, args.path, args.variant,
                       args.kz, wake_check=check, keep_fields=check)
        print(f"[bench] fp64 {_brief(r64)}", file=sys.stderr, flush=True)
    fields = None
    if r64 is not None and "E" in r32 and "E" in r64:
        fields = {"quantity": "E and H after %d steps (warm-up included), FP32 run vs FP64 run"
                              % (args.warmup + args.steps),
                  "rel_l2_E": _rel_l2(r32["E"], r64["E"]), "rel_l2_H": _rel_l2(r32["H"], r64["H"])}
        for r in (r32, r64):
            r.pop("E"), r.pop("H")
        torch.cuda.empty_cache()
        print(f"[bench] fields {fields}", file=sys.stderr, flush=True)
    lossy = None
    if not args.no_lossy:
        l32 = lossy_pml(np.float32, keep_fields=True)
        l64 = lossy_pml(np.float64, keep_fields=True)
        lossy = {"workload": "lossy pillbox r=40 mm, 3-cell sigma=1e4 S/m wall, 10-layer PML "
                             "on z, 200x200x400 cells, beam sigma_z=15 mm beta=0.8, 320 steps",
                 "dt_policy": "use_sibc=True (the reference default): wall surface cells "
                              "with sigma = eps = 1/Z_s at fmax = 1 GHz; dt = %.4g s (CFL, "
                              "no tau clamp)" % l64["dt"],
                 "fp32": {k: v for k, v in l32.items() if k not in ("E", "H", "dt")},
                 "fp64": {k: v for k, v in l64.items() if k not in ("E", "H", "dt")},
                 "rel_l2_E_fp32_vs_fp64": _rel_l2(l32["E"], l64["E"]),
                 "rel_l2_H_fp32_vs_fp64": _rel_l2(l32["H"], l64["H"])}
        del l32, l64
        torch.cuda.empty_cache()
        print(f"[bench] lossy {lossy}", file=sys.stderr, flush=True)
    adj = adj64 = None
    if not args.no_adjoint:
        adj = adjoint_cost(steps=args.adjoint_steps)
        print(f"[bench] adjoint {adj}", file=sys.stderr, flush=True)
        adj64 = adjoint_cost(np.float64, steps=args.adjoint_steps)
        print(f"[bench] adjoint fp64 {adj64}", file=sys.stderr, flush=True)
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
                   "cells": r32["N"], "path": r32["path"], "full_rows": r32["full_rows"]},
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
        if check:
            n = min(len(r32["WP"]), len(r64["WP"]))  # s axes from dt32 / dt64
            wp32, wp64 = r32["WP"][:n], r64["WP"][:n]
            line["wake_check"] = {
                "quantity": "WP_z(s) on axis, s in [-ti*v, %.2f m), integral truncated at the "
                            "last of %d steps" % (WAKE_CHECK_LENGTH, args.warmup + args.steps),
                "reference": "same run with FP64 fields (the FP64 kernel is bit-identical to "
                             "the FP64 oracle)",
                "rel_l2_fp32_vs_fp64": float(np.linalg.norm(wp32 - wp64) / np.linalg.norm(wp64)),
                "n_s": int(len(wp64))}
    if fields is not None:
        line["field_check"] = fields
    if lossy is not None:
        line["lossy_pml"] = lossy
    if adj is not None:
        line["adjoint"] = adj
        line["adjoint_fp64"] = adj64
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
