"""GPU: the benchmark workloads at their full sizes, through checks that do not
depend on the grid size (the oracle's CPU run would take hours there).

* 256x256x512 cavity + pipe, beam on axis: the fused step (one launch per step,
  32-bit offsets, ~10^5 workgroups) equals the exact general path bit for bit,
  FP32 and FP64; a lossless FP64 run conserves the leapfrog energy.
* 200x200x400 lossy pillbox with PML and SIBC: whole-row workgroups, row
  segments and the general path give bitwise the same fields.
* Grids whose 3N field arrays pass 4 GiB (up to 1024x1024x700 FP32, 3N > 2^31
  elements): the fused path's 64-bit-offset kernels equal the general path.
* 128x128x256 cavity with a 32x32x24 block of per-cell eps_r and sigma: the
  checkpoint mode (segments recomputed on the block's domain of dependence)
  and the trajectory mode give bitwise the same gradients.

The general path is pinned to the oracle bit for bit at small sizes
(tests/test_gpu_parity.py); it shares no indexing or tiling with the fused
kernel.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cavity(dims, dtype, path):
    import wakis_amd as wa
    from wakis_amd.solids import Box, Cylinder
    solid = Box(-25e-3, 25e-3, -25e-3, 25e-3, -25e-3, 25e-3) | Cylinder(5e-3)
    grid = wa.GridFIT3D(-30e-3, 30e-3, -30e-3, 30e-3, -80e-3, 80e-3, *dims,
                        stl_solids={"cavity": solid}, stl_materials={"cavity": "vacuum"},
                        verbose=0)
    return wa.SolverFIT3D(grid, bc_low=["pec", "pec", "abc"], bc_high=["pec", "pec", "abc"],
                          use_stl=True, bg="pec", dtype=dtype, verbose=0, path=path)


def _beam_steps(solvers, n, sigmaz=10e-3, beta=1.0):
    import wakis_amd as wa
    beams = [wa.Beam(q=1e-9, sigmaz=sigmaz, beta=beta) for _ in solvers]
    for k in range(n):
        for s, b in zip(solvers, beams):
            b.update(s, (k + 0.5 * n) * s.dt)  # bunch head inside the domain
            s.one_step()


def _same_fields(a, b):
    for name in ("E", "H", "J"):
        u, v = getattr(a, name).array, getattr(b, name).array
        assert bool((u == v).all().item()), name


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_cavity_256x256x512_fused_equals_general_path(dtype):
    print("256x256x512 fused vs general", np.dtype(dtype).name, flush=True)
    dims = (256, 256, 512)
    f, g = _cavity(dims, dtype, "fused"), _cavity(dims, dtype, "general")
    _beam_steps([f, g], 6)
    assert f.last_path == "fused" and g.last_path == "general"
    assert bool(f.E.array.abs().max().item() > 0)
    _same_fields(f, g)


def test_cavity_256x256x512_lossless_energy_conserved():
    import torch
    print("256x256x512 energy", flush=True)
    s = _cavity((256, 256, 512), np.float64, "fused")
    # a smooth field inside the cavity, zero near the walls and outside
    ix = torch.arange(s.Nx, device=s.device, dtype=torch.float64)
    iy = torch.arange(s.Ny, device=s.device, dtype=torch.float64)
    iz = torch.arange(s.Nz, device=s.device, dtype=torch.float64)
    bump = (torch.sin(np.pi * (iz[:, None, None] - 200) / 112).clamp(min=0)
            * torch.sin(np.pi * (iy[None, :, None] - 40) / 176).clamp(min=0)
            * torch.sin(np.pi * (ix[None, None, :] - 40) / 176).clamp(min=0))
    bump[:200] = 0
    bump[312:] = 0
    E = torch.zeros(3 * s.N, dtype=torch.float64, device=s.device)
    E[2 * s.N:] = bump.reshape(-1)
    s.E.fromarray(E)
    s.one_step()
    s.one_step()
    u0 = s.energy()
    s._run(s._step_args(300))
    u1 = s.energy()
    assert s.last_path == "fused" and u0 > 0
    assert abs(u1 - u0) <= 1e-9 * u0, (u0, u1)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_lossy_pillbox_200x200x400_shapes_and_general_path_agree(dtype):
    import wakis_amd as wa
    from wakis_amd.solids import Cylinder
    print("200x200x400 lossy pillbox shapes", np.dtype(dtype).name, flush=True)
    Nx, Ny, Nz = 200, 200, 400
    x = np.linspace(-45e-3, 45e-3, Nx + 1)
    y = np.linspace(-45e-3, 45e-3, Ny + 1)
    z = np.linspace(-90e-3, 90e-3, Nz + 1)
    vac = (Cylinder(40e-3, -30e-3, 30e-3) | Cylinder(10e-3)).cell_mask(x, y, z)
    grown = vac.copy()
    for _ in range(3):
        gr = grown.copy()
        gr[1:] |= grown[:-1]
        gr[:-1] |= grown[1:]
        gr[:, 1:] |= grown[:, :-1]
        gr[:, :-1] |= grown[:, 1:]
        gr[:, :, 1:] |= grown[:, :, :-1]
        gr[:, :, :-1] |= grown[:, :, 1:]
        grown = gr
    wall = grown & ~vac

    def solver(path, rows):
        grid = wa.GridFIT3D(x=x, y=y, z=z, stl_solids={"wall": wall, "vac": vac},
                            stl_materials={"wall": [1.0, 1.0, 1e4], "vac": "vacuum"},
                            verbose=0)
        s = wa.SolverFIT3D(grid, bc_low=["pec", "pec", "pml"], bc_high=["pec", "pec", "pml"],
                           use_stl=True, bg="pec", dtype=dtype, verbose=0, n_pml=10,
                           path=path)
        s.fused_full_rows = rows
        return s

    runs = [solver("fused", 4), solver("fused", -1), solver("general", 0)]
    _beam_steps(runs, 5, sigmaz=15e-3, beta=0.8)
    assert [s.last_path for s in runs] == ["fused", "fused", "general"]
    assert bool(runs[0].E.array.abs().max().item() > 0)
    _same_fields(runs[0], runs[1])
    _same_fields(runs[0], runs[2])


@pytest.mark.parametrize("dtype,dims,variant", [(np.float32, (512, 512, 1536), 0),
                                                (np.float64, (512, 512, 768), 0),
                                                (np.float32, (1024, 1024, 700), 0),
                                                (np.float32, (512, 512, 1536), 1)])
def test_field_arrays_over_4gib_fused_equals_general_path(dtype, dims, variant):
    """Every 3N field array is 4.83 GB (8.81 GB and 3N > 2^31 elements on the
    1024x1024x700 grid), past the 32-bit byte offsets of the default fused
    kernel: the fused path takes the same tile code with 64-bit offsets
    (k_fused4_wide), or the one-cell-per-lane kernel (variant 1), and equals
    the general path bit for bit, from random E and H."""
    import torch

    import wakis_amd as wa
    print("over 4 GiB fused vs general", np.dtype(dtype).name, dims, variant, flush=True)
    grid = wa.GridFIT3D(-30e-3, 30e-3, -30e-3, 30e-3, -80e-3, 80e-3, *dims, verbose=0)
    # PEC on every face: the flat-index wrap terms are masked, as the fused path needs
    s = wa.SolverFIT3D(grid, bc_low=["pec"] * 3, bc_high=["pec"] * 3, dtype=dtype, verbose=0,
                       path="general")
    assert 3 * s.N * np.dtype(dtype).itemsize > 2 ** 32
    print("solver built", flush=True)
    gen = torch.Generator(device=s.device).manual_seed(5)
    for F in (s.E, s.H):
        F.fromarray(torch.randn(3 * s.N, generator=gen, device=s.device,
                                dtype=F.array.dtype))
    s.one_step()  # first step: ghost cells zeroed
    saved = [F.array.clone() for F in (s.E, s.H, s.J)]
    s.path = "fused"
    s.fused_variant = variant
    s._run(s._step_args(3))
    assert s.last_path == "fused"
    fused = [F.array.clone() for F in (s.E, s.H, s.J)]
    print("fused steps done", flush=True)
    for F, X in zip((s.E, s.H, s.J), saved):
        F.array.copy_(X)
        F.version += 1
    del saved
    s.path = "general"
    s._run(s._step_args(3))
    assert s.last_path == "general"
    for F, X, name in zip((s.E, s.H, s.J), fused, "EHJ"):
        assert bool((F.array == X).all().item()), name
    assert bool(s.E.array.abs().max().item() > 0)


@pytest.mark.parametrize("dtype,dims", [(np.float32, (512, 512, 1536)),
                                        (np.float64, (512, 512, 768))])
def test_beam_over_4gib_fused_equals_general_path(dtype, dims):
    """A beam on the axis of a PEC box whose 3N arrays pass 4 GiB: the
    64-bit-offset kernel with the beam-line current and no boundary masks
    (fields zero on every masked edge) equals the general path bit for bit."""
    import torch

    import wakis_amd as wa
    print("over 4 GiB beam fused vs general", np.dtype(dtype).name, dims, flush=True)
    out = {}
    for path in ("fused", "general"):
        grid = wa.GridFIT3D(-30e-3, 30e-3, -30e-3, 30e-3, -80e-3, 80e-3, *dims, verbose=0)
        s = wa.SolverFIT3D(grid, bc_low=["pec"] * 3, bc_high=["pec"] * 3, dtype=dtype,
                           verbose=0, path=path)
        assert 3 * s.N * np.dtype(dtype).itemsize > 2 ** 32
        beam = wa.Beam(q=1e-9, sigmaz=10e-3, beta=1.0)
        for k in range(6):
            beam.update(s, (k + 1500) * s.dt)  # bunch inside the box
            s.one_step()
        assert s.last_path == path
        out[path] = [F.array.clone() for F in (s.E, s.H, s.J)]
        del s, grid
        torch.cuda.empty_cache()
        print(path, "steps done", flush=True)
    assert bool(out["fused"][0].abs().max().item() > 0)
    for X, Y, name in zip(out["fused"], out["general"], "EHJ"):
        assert bool(torch.equal(X, Y)), name


def test_adjoint_128x128x256_checkpoint_and_trajectory_gradients_agree():
    import wakis_amd as wa
    from wakis_amd.adjoint import WakeAdjoint
    from wakis_amd.solids import Box, Cylinder
    print("128x128x256 adjoint modes", flush=True)
    dims, insert = (128, 128, 256), (32, 32, 24)
    x = np.linspace(-30e-3, 30e-3, dims[0] + 1)
    y = np.linspace(-30e-3, 30e-3, dims[1] + 1)
    z = np.linspace(-80e-3, 80e-3, dims[2] + 1)
    cav = (Box(-25e-3, 25e-3, -25e-3, 25e-3, -25e-3, 25e-3) | Cylinder(5e-3)).cell_mask(x, y, z)
    lo = [(n - m) // 2 for n, m in zip(dims, insert)]
    region = tuple((a, a + m) for a, m in zip(lo, insert))
    ins = np.zeros(dims, dtype=bool)
    ins[tuple(slice(a, b) for a, b in region)] = True
    rng = np.random.default_rng(0)
    er, sg = rng.uniform(1.0, 10.0, insert), rng.uniform(0.0, 2.0, insert)
    out = {}
    for mode in ("trajectory", "checkpoint"):
        grid = wa.GridFIT3D(x=x, y=y, z=z, stl_solids={"cav": cav, "ins": ins},
                            stl_materials={"cav": "vacuum", "ins": [2.0, 1.0, 1.0]}, verbose=0)
        wake = wa.WakeSolver(q=1e-9, sigmaz=10e-3, beta=1.0, skip_cells=10, save=False,
                             verbose=0)
        s = wa.SolverFIT3D(grid, wake=wake, bc_low=["pec", "pec", "abc"],
                           bc_high=["pec", "pec", "abc"], use_stl=True, bg="pec",
                           dtype=np.float32, verbose=0, use_sibc=False, tau_clamp=False)
        adj = WakeAdjoint(s, region, 0.01, w_loss=1.0, w_band=1e-13, mode=mode, segment=16)
        L = adj.forward(er, sg)
        assert adj.used_mode == mode
        out[mode] = (L,) + tuple(g.cpu().numpy() for g in adj.backward())
    (La, ea, sa), (Lb, eb, sb) = out["trajectory"], out["checkpoint"]
    assert La == Lb and np.any(ea != 0) and np.all(np.isfinite(ea))
    assert np.array_equal(ea, eb) and np.array_equal(sa, sb)


def test_conductive_and_box_over_4gib_fused_equals_general_path():
    """512x512x1536 FP32 (4.83 GB per 3N array) with 10 PML layers on z: the
    conductive instantiation of the 64-bit-offset kernel (J = sigma E, stored
    only by the last step of a multi-step call) equals the general path bit for
    bit from random E and H; and a multi-step call limited to a box (the
    checkpoint recompute's launch) reproduces the full-grid run on the box
    minus a margin of one cell per step."""
    import torch

    import wakis_amd as wa
    dims, k = (512, 512, 1536), 3
    print("over 4 GiB conductive and box", dims, flush=True)
    grid = wa.GridFIT3D(-30e-3, 30e-3, -30e-3, 30e-3, -80e-3, 80e-3, *dims, verbose=0)
    s = wa.SolverFIT3D(grid, bc_low=["pec", "pec", "pml"], bc_high=["pec", "pec", "pml"],
                       dtype=np.float32, verbose=0, n_pml=10, path="general")
    assert s.use_conductivity and 3 * s.N * 4 > 2 ** 32
    gen = torch.Generator(device=s.device).manual_seed(9)
    for F in (s.E, s.H):
        F.fromarray(torch.randn(3 * s.N, generator=gen, device=s.device, dtype=F.array.dtype))
    s.one_step()  # general step: ghosts zeroed, J = sigma E from here on
    saved = [F.array.clone() for F in (s.E, s.H, s.J)]
    print("solver built", flush=True)

    def restore():
        for F, X in zip((s.E, s.H, s.J), saved):
            F.array.copy_(X)
            F.version += 1
        s._sigmaE_versions = (s.E.version, s.J.version)

    s.path = "fused"
    s._run(s._step_args(k))
    assert s.last_path == "fused"
    fused = [F.array.clone() for F in (s.E, s.H, s.J)]
    restore()
    s.path = "general"
    s._run(s._step_args(k))
    assert s.last_path == "general"
    for F, X, name in zip((s.E, s.H, s.J), fused, "EHJ"):
        assert bool(torch.equal(F.array, X)), name
    assert bool(fused[2].abs().max().item() > 0)
    print("conductive fused == general", flush=True)
    # box-limited multi-step call: the cells [lo + k + 1, hi - k - 1) equal the full run
    restore()
    s.path = "fused"
    a = s._step_args(k)
    lo, hi = (100, 180, 700), (300, 330, 760)
    a.box_lo[:] = lo
    a.box_hi[:] = hi
    s._run(a)
    assert s.last_path == "fused"
    inner = (slice(None),) + tuple(slice(l + k + 1, h - k - 1) for l, h in
                                   zip(reversed(lo), reversed(hi)))
    for F, X, name in zip((s.E, s.H), fused, "EH"):
        got = F.array.view(3, dims[2], dims[1], dims[0])[inner]
        ref = X.view(3, dims[2], dims[1], dims[0])[inner]
        assert bool(torch.equal(got, ref)), name
    print("box-limited call == full run inside the box", flush=True)
