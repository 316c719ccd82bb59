"""CPU: host-side set-up of wakis_amd.SolverFIT3D (no GPU needed).

The solver is built on CPU torch tensors with the HIP handle creation
skipped, which exercises everything before the first kernel launch:
material folding, boundary bits, dt, and the flat-index wrap proof that
gates the fused kernel.
"""

import numpy as np
import pytest
import scipy.sparse as sp
import torch

import wakis_amd as wa
import wakis_amd.field as wf
from oracle import cases
from oracle import fit_oracle as fo
from tests import gpu_cases as gc
from tests.test_stencil_rules import BCS, model_coeffs


@pytest.fixture
def host_only(monkeypatch):
    monkeypatch.setattr(wf, "default_device", lambda: torch.device("cpu"))
    monkeypatch.setattr(wa.SolverFIT3D, "_create_handle", lambda self, lib: None)
    monkeypatch.setattr(torch.cuda, "synchronize", lambda *a, **k: None)


@pytest.mark.parametrize("bci", range(len(BCS)))
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("conductive", [False, True])
def test_folded_coefficients_match_model(host_only, bci, dtype, conductive):
    case = gc.random_case(BCS[bci], dtype, conductive, seed=11 + bci)
    so = gc.oracle_solver(case)
    sa = gc.amd_solver(case)
    a, b, L1, tL1, pecz, pmcz, m1desc = model_coeffs(so)
    assert np.array_equal(sa._a.numpy(), a) and np.array_equal(sa._b.numpy(), b)
    assert (sa.pec_zero, sa.pmc_zero, sa.m1_descending) == (pecz, pmcz, m1desc)
    assert all(np.array_equal(u, v) for u, v in zip(sa._tL1, tL1))
    assert sa.dt == so.dt and type(sa.dt) is type(so.dt)
    for name in ("ieps", "imu", "sigma"):
        assert np.array_equal(getattr(sa, name).to_numpy(), getattr(so, name))


def _strip_wraps(M, dims):
    """Copy of a CSR operator without the flat-index wrap entries.

    Entries are removed in place so every row keeps scipy's stored column
    order (and hence its summation order).
    """
    Nx, Ny, Nz = dims
    N = Nx * Ny * Nz
    M = M.copy()
    rows = np.repeat(np.arange(M.shape[0]), np.diff(M.indptr))
    r, c = rows % N, M.indices % N
    ri, rj = r % Nx, (r // Nx) % Ny
    ci, cj = c % Nx, (c // Nx) % Ny
    wrap = (np.abs(ri - ci) > 1) | (np.abs(rj - cj) > 1)
    M.data[wrap] = 0
    M.eliminate_zeros()
    assert isinstance(M, sp.csr_matrix)
    return M


def _run(so, M1, M2, steps):
    E, H, J = so.E.copy(), so.H.copy(), so.J.copy()
    for _ in range(steps):
        H = H - so.dt * M1 * E
        E = E + so.dt * (M2 * H - so.ieps * J)
        if so.use_conductivity:
            J = so.sigma * E
    return E, H


def _wrap_cases():
    out = []
    for bci in range(len(BCS)):
        for seed in (1, 2):
            out.append(("random", bci, seed))
    out.append(("cavity", None, 0))
    return out


@pytest.mark.parametrize("kind,bci,seed", _wrap_cases())
def test_wrap_proof_is_sound(host_only, kind, bci, seed):
    if kind == "cavity":
        x, y, z, mask = cases.cavity_pipe_geometry(10, 10, 16)
        bc = (["pec", "pec", "abc"], ["pec", "pec", "abc"])
        case = dict(x=x, y=y, z=z, bc=bc, bg="pec", solids=[("cav", mask, "vacuum")],
                    dtype=np.float64, n_pml=2,
                    fields=[np.zeros(3 * mask.size) for _ in range(3)])
        rng = np.random.default_rng(0)
        f = np.zeros(3 * mask.size)  # fields only inside the vacuum region
        inside = np.tile(np.reshape(mask, -1, order="F"), 3)
        f[inside] = rng.standard_normal(inside.sum())
        case["fields"] = [f.copy(), f.copy() * 0.5, np.zeros_like(f)]
    else:
        case = gc.random_case(BCS[bci], np.float64, seed == 2, seed=seed, dims=(7, 6, 8))
        if seed == 2:  # zero the fields on the outer layer: the proof may then pass
            for F in case["fields"]:
                for c in range(3):
                    V = fo.comp(F, (7, 6, 8), c)
                    V[0], V[-1], V[:, 0], V[:, -1] = 0, 0, 0, 0
    so = gc.oracle_solver(case)
    sa = gc.amd_solver(case)
    so.set_ghosts_to_0()
    so.step_0 = False
    dead = sa._wraps_dead()
    full = _run(so, so.M1, so.M2, 6)
    cut = _run(so, _strip_wraps(so.M1, so.dims), _strip_wraps(so.M2, so.dims), 6)
    same = all(np.array_equal(u, v) for u, v in zip(full, cut))
    if dead:
        assert same, "proof says the wrap terms are dead but dropping them changes the run"
    if kind == "cavity":
        assert dead, "the cavity/pipe set-up should be provably wrap-free"


@pytest.mark.parametrize("conductive", [False, True])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_material_table_reproduces_coefficients(host_only, conductive, dtype):
    x, y, z, mask = cases.cavity_pipe_geometry(20, 20, 32)
    zb = "pml" if conductive else "abc"
    case = dict(x=x, y=y, z=z, bc=(["pec", "pec", zb], ["pec", "pec", zb]),
                bg="pec", solids=[("cav", mask, "vacuum")], dtype=dtype, n_pml=4,
                fields=[np.zeros(3 * mask.size)] * 3)
    sa = gc.amd_solver(case)
    assert sa.use_conductivity == conductive
    assert sa._material_table(), "compressed coefficients rejected"
    assert sa._nmat <= 256 and sa._code.dtype == torch.uint8


def test_material_table_on_random_grid(host_only):
    """Non-uniform grid, three materials: still exact (metric tables are 2-D)."""
    case = gc.random_case((["pec", "pec", "pec"], ["pec", "pec", "pec"]), np.float64, True,
                          seed=4)
    sa = gc.amd_solver(case)
    assert sa._material_table()
    sa.ieps.array[0] += 1.0  # edit without update_tensors -> arrays again
    sa.ieps.version += 1
    assert not sa._material_table()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_adjoint_kernel_choice_past_4gib(host_only, monkeypatch, dtype):
    """Reverse-kernel choice when 3N doubles pass the fused kernel's 4 GiB of
    32-bit offsets (the limit lowered to this grid's size): "auto" takes the
    two-pass kernels with FP64 adjoint fields, also for FP32 solver fields;
    kernel="fused" and FP32 adjoint fields without the fused kernel raise."""
    from wakis_amd import adjoint as adjm
    x, y, z, mask = cases.cavity_pipe_geometry(16, 16, 40)
    ins = np.zeros((16, 16, 40), dtype=bool)
    region = ((6, 10), (6, 10), (17, 23))
    ins[tuple(slice(a, b) for a, b in region)] = True
    grid = wa.GridFIT3D(x=x, y=y, z=z, stl_solids={"cav": mask, "ins": ins},
                        stl_materials={"cav": "vacuum", "ins": [2.0, 1.0, 0.5]}, verbose=0)
    wake = wa.WakeSolver(q=1e-9, sigmaz=10e-3, beta=1.0, save=False, verbose=0)
    s = wa.SolverFIT3D(grid, wake=wake, bc_low=["pec", "pec", "abc"],
                       bc_high=["pec", "pec", "abc"], use_stl=True, bg="pec", dtype=dtype,
                       verbose=0, use_sibc=False, tau_clamp=False)
    adj = adjm.WakeAdjoint(s, region, 0.03)
    assert adj._use_fused_kernel()
    assert adj._adjoint_f32(True) == (dtype == np.float32)
    monkeypatch.setattr(adjm, "FUSED_MAX_BYTES", 3 * s.N * 8)  # this grid is one past it
    assert not adj._use_fused_kernel()
    assert adj._adjoint_f32(False) is False  # FP32 solver fields, FP64 adjoint fields
    with pytest.raises(ValueError):
        adjm.WakeAdjoint(s, region, 0.03, kernel="fused")._use_fused_kernel()
    with pytest.raises(ValueError):
        adjm.WakeAdjoint(s, region, 0.03, adjoint_dtype="float32")._adjoint_f32(False)
    monkeypatch.setattr(adjm, "FUSED_MAX_BYTES", 3 * s.N * 8 + 1)
    assert adjm.WakeAdjoint(s, region, 0.03, kernel="fused")._use_fused_kernel()


@pytest.mark.parametrize("freqs", [[0.7e9, 2.1e9, 4.0e9], np.array([1e9, 3e9], np.float32)])
@pytest.mark.parametrize("dt", [np.float32(1.234567e-12), np.float64(1.234567e-12),
                                1.234567e-12])
def test_monitor_phase_table_is_the_per_update_scalar_expression(freqs, dt):
    """FieldMonitor's vectorised phase table equals, value and dtype, the
    reference's per-update scalar exp(-2j*pi*f*t), t = time_index*dt
    (field_monitors.py:90-97): complex64 for an FP32 dt (or FP32 frequencies
    with a Python float dt), complex128 otherwise."""
    m = wa.FieldMonitor(freqs)
    m.dt, m.time_index = dt, 5
    ph = m._phase_values(1500)
    for k in range(1500):
        t = (5 + k) * m.dt
        for i, fr in enumerate(m.frequencies):
            sc = np.exp(-2j * np.pi * fr * t)
            assert np.asarray(sc).dtype == ph.dtype
            assert sc == ph[k, i], (k, i)
    want = np.complex64 if (isinstance(dt, np.float32) or
                            (not isinstance(dt, np.floating) and
                             np.asarray(freqs).dtype == np.float32)) else np.complex128
    assert ph.dtype == want
