"""GPU: the fused single-pass kernel against the oracle (bit-identical)."""

import numpy as np
import pytest

from oracle import cases
from oracle import fit_oracle as fo
from tests import gpu_cases as gc
from tests.test_stencil_rules import BCS

pytestmark = pytest.mark.gpu

PEC_ALL = (["pec", "pec", "pec"], ["pec", "pec", "pec"])
PMC_MIX = (["pec", "pmc", "pec"], ["pmc", "pec", "pec"])


@pytest.mark.parametrize("dims", [(12, 10, 14), (70, 13, 37), (64, 8, 33)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("conductive", [False, True])
@pytest.mark.parametrize("kz", [0, 5, 1])
@pytest.mark.parametrize("use_code", [True, False])
@pytest.mark.parametrize("variant", [0, 1, 2, 3, 7, 8])
def test_fused_matches_oracle(dims, dtype, conductive, kz, use_code, variant):
    print("fused", dims, np.dtype(dtype).name, conductive, kz, use_code, variant, flush=True)
    case = gc.random_case(PEC_ALL, dtype, conductive, seed=sum(dims) + kz, dims=dims)
    # the fused step has no general J read: J starts at zero (conductive: J == sigma*E
    # is re-established by one exact step)
    case["fields"][2] = np.zeros_like(case["fields"][2])
    so = gc.oracle_solver(case)
    sa = gc.amd_solver(case, path="fused")
    sa.fused_kz = kz
    sa.use_material_code = use_code
    sa.fused_variant = variant
    if conductive:
        so.J[:] = 0
        so.one_step()          # reference step with J = 0
        sa.path = "general"    # the first step (J not sigma*E yet) runs exactly
        sa.one_step()
        sa.path = "fused"
    if variant == 8 and dtype == np.float64 and dims[0] % 2 == 0:
        with pytest.raises(ValueError, match="variant"):  # 8-row tiles: FP32 only
            sa.one_step()
        return
    for _ in range(5):
        so.one_step()
        sa.one_step()
        assert sa.last_path == "fused"
    assert (sa._code is not None) == use_code
    for name in ("E", "H", "J"):
        got, ref = getattr(sa, name).to_numpy(), getattr(so, name)
        assert np.array_equal(got, ref), f"{name}: rel-L2 {gc.rel_l2(got, ref):.3e}"


@pytest.mark.parametrize("dims", [(12, 10, 14), (70, 13, 37), (128, 6, 20), (36, 9, 11)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("conductive", [False, True])
@pytest.mark.parametrize("use_code", [True, False])
@pytest.mark.parametrize("cpl", [0, 1, 2, 4])
def test_fused_cells_per_lane_matches_oracle(dims, dtype, conductive, use_code, cpl):
    """Every cells-per-lane choice of the x-vectorised kernel is bit-identical
    to the oracle (automatic: the one that leaves the fewest idle lanes)."""
    print("fused cells/lane", dims, np.dtype(dtype).name, conductive, use_code, cpl, flush=True)
    case = gc.random_case(PEC_ALL, dtype, conductive, seed=sum(dims) + cpl, dims=dims)
    case["fields"][2] = np.zeros_like(case["fields"][2])
    so = gc.oracle_solver(case)
    sa = gc.amd_solver(case, path="fused")
    sa.use_material_code = use_code
    sa.fused_cells_per_lane = cpl
    if conductive:
        so.J[:] = 0
        so.one_step()
        sa.path = "general"
        sa.one_step()
        sa.path = "fused"
    vn = 16 // np.dtype(dtype).itemsize
    if cpl not in (0, 1, vn) and not (cpl == 2 and vn == 4) or (cpl and dims[0] % cpl):
        with pytest.raises(ValueError, match="cells_per_lane"):
            sa.one_step()
        return
    for _ in range(4):
        so.one_step()
        sa.one_step()
        assert sa.last_path == "fused"
    for name in ("E", "H", "J"):
        got, ref = getattr(sa, name).to_numpy(), getattr(so, name)
        assert np.array_equal(got, ref), f"{name}: rel-L2 {gc.rel_l2(got, ref):.3e}"


@pytest.mark.parametrize("dims", [(12, 10, 14), (64, 8, 33), (70, 13, 37)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("use_code", [True, False])
def test_fused_multistep_conductive_call_matches_oracle(dims, dtype, use_code):
    """One wf_step call of several conductive steps: J = sigma E is stored only
    by the call's last step (the next step recomputes sigma E from E), and a
    beam line without a beam table keeps its J between the steps; E, H and J
    equal the oracle's step-by-step run."""
    print("fused multistep conductive", dims, np.dtype(dtype).name, use_code, flush=True)
    case = gc.random_case(PEC_ALL, dtype, True, seed=sum(dims) + 17, dims=dims)
    case["fields"][2] = np.zeros_like(case["fields"][2])
    so = gc.oracle_solver(case)
    sa = gc.amd_solver(case, path="fused")
    sa.use_material_code = use_code
    line = (dims[0] // 2, dims[1] // 2)
    so.J[:] = 0
    so.one_step()
    sa.path = "general"
    sa.one_step()  # J = sigma E now holds
    sa.path = "fused"
    sa._beam_line = line  # a beam line with no beam table: its J is read back
    for _ in range(6):
        so.one_step()
    sa._run(sa._step_args(6))
    assert sa.last_path == "fused"
    for name in ("E", "H", "J"):
        got, ref = getattr(sa, name).to_numpy(), getattr(so, name)
        assert np.array_equal(got, ref), f"{name}: rel-L2 {gc.rel_l2(got, ref):.3e}"


def _full_rows_fit(dims, dtype, rows):
    """Whether a full-row tile of `rows` rows fits (512 threads, 64 KiB LDS
    without the material table), as the engine decides."""
    vn = 16 // np.dtype(dtype).itemsize
    Nx, Ny = dims[0], dims[1]
    if Nx % vn:
        return False
    q = Nx // vn

    def fits(r):
        thr = -(-(r + 1) * q // 64) * 64
        return thr <= 512 and (6 * r + 9) * (Nx + 2 * vn) * np.dtype(dtype).itemsize <= 60000

    return fits(1) if rows == 1 else fits(rows)


@pytest.mark.parametrize("dims", [(12, 10, 14), (200, 9, 12), (40, 17, 21), (36, 3, 11),
                                  (70, 13, 37)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("conductive", [False, True])
@pytest.mark.parametrize("use_code", [True, False])
@pytest.mark.parametrize("rows", [1, 2, 5, 9])
def test_fused_full_rows_matches_oracle(dims, dtype, conductive, use_code, rows):
    """Workgroups of whole rows (thread -> (row, cells) map, no x halos) are
    bit-identical to the oracle for every row count that fits; the others are
    refused."""
    print("fused full rows", dims, np.dtype(dtype).name, conductive, use_code, rows, flush=True)
    case = gc.random_case(PEC_ALL, dtype, conductive, seed=sum(dims) + 3 * rows, dims=dims)
    case["fields"][2] = np.zeros_like(case["fields"][2])
    so = gc.oracle_solver(case)
    sa = gc.amd_solver(case, path="fused")
    sa.use_material_code = use_code
    sa.fused_full_rows = rows
    if conductive:
        so.J[:] = 0
        so.one_step()
        sa.path = "general"
        sa.one_step()
        sa.path = "fused"
    if not _full_rows_fit(dims, dtype, rows):
        with pytest.raises(ValueError, match="full-row"):
            sa.one_step()
        return
    for _ in range(5):
        so.one_step()
        sa.one_step()
        assert sa.last_path == "fused"
    for name in ("E", "H", "J"):
        got, ref = getattr(sa, name).to_numpy(), getattr(so, name)
        assert np.array_equal(got, ref), f"{name}: rel-L2 {gc.rel_l2(got, ref):.3e}"


@pytest.mark.parametrize("bc", [PEC_ALL, PMC_MIX])
@pytest.mark.parametrize("zero_masked", [False, True])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("rows", [1, 3])
def test_fused_full_rows_masks_match_oracle(bc, zero_masked, dtype, rows):
    """Full-row tiles with PEC/PMC masks applied in the kernel or folded away;
    a multi-step conductive call with a beam line and no beam table."""
    print("fused full rows masks", bc[0], zero_masked, np.dtype(dtype).name, rows, flush=True)
    dims = (40, 11, 17)
    case = gc.random_case(bc, dtype, True, seed=29 + rows, dims=dims)
    case["fields"][2] = np.zeros_like(case["fields"][2])
    if zero_masked:
        _zero_masked_values(case)
    so = gc.oracle_solver(case)
    sa = gc.amd_solver(case, path="fused")
    sa.fused_full_rows = rows
    so.J[:] = 0
    so.one_step()
    sa.path = "general"
    sa.one_step()
    sa.path = "fused"
    sa._beam_line = (dims[0] // 2, dims[1] // 2)
    for _ in range(4):
        so.one_step()
    sa._run(sa._step_args(4))
    assert sa.last_path == "fused"
    assert sa._masked_zero() == zero_masked
    for name in ("E", "H", "J"):
        got, ref = getattr(sa, name).to_numpy(), getattr(so, name)
        assert np.array_equal(got, ref), f"{name}: rel-L2 {gc.rel_l2(got, ref):.3e}"


@pytest.mark.parametrize("conductive", [False, True])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_tune_fused_restores_state_and_keeps_results(conductive, dtype):
    """tune_fused() times every candidate shape, leaves E, H, J as they were,
    and the run that follows equals the oracle bitwise."""
    print("tune fused", conductive, np.dtype(dtype).name, flush=True)
    dims = (40, 21, 30)
    case = gc.random_case(PEC_ALL, dtype, conductive, seed=41, dims=dims)
    case["fields"][2] = np.zeros_like(case["fields"][2])
    so = gc.oracle_solver(case)
    sa = gc.amd_solver(case, path="fused")
    if conductive:
        so.J[:] = 0
        so.one_step()
        sa.path = "general"
        sa.one_step()
        sa.path = "fused"
    before = [getattr(sa, n).to_numpy() for n in ("E", "H", "J")]
    best = sa.tune_fused()
    assert best is not None and best in sa.fused_tuning and -1 in sa.fused_tuning
    assert len(sa.fused_tuning) >= 3 and sa.fused_full_rows == best
    for n, b in zip(("E", "H", "J"), before):
        assert np.array_equal(getattr(sa, n).to_numpy(), b), n
    assert sa.tune_fused() is None  # the choice is now fixed
    so.one_step()
    sa.one_step()  # (first step: ghost zeroing, as in the oracle)
    for _ in range(4):
        so.one_step()
    sa._run(sa._step_args(4))
    assert sa.last_path == "fused"
    for name in ("E", "H", "J"):
        got, ref = getattr(sa, name).to_numpy(), getattr(so, name)
        assert np.array_equal(got, ref), f"{name}: rel-L2 {gc.rel_l2(got, ref):.3e}"


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("conductive", [False, True])
@pytest.mark.parametrize("rows", [0, 3])
def test_fused_escape_codes_match_oracle(dtype, conductive, rows):
    """A block of cells with their own eps_r (and sigma): more than 256
    materials, so those cells get the escape code and read the coefficient
    arrays while the rest use the table; with a beam line inside the block.
    Bit-identical to the oracle, row segments and whole rows."""
    from scipy.constants import epsilon_0
    print("fused escape", np.dtype(dtype).name, conductive, rows, flush=True)
    dims = (24, 20, 30)
    Nx, Ny, Nz = dims
    N = Nx * Ny * Nz
    case = gc.random_case(PEC_ALL, dtype, conductive, seed=23 + rows, dims=dims)
    case["fields"][2] = np.zeros_like(case["fields"][2])
    so = gc.oracle_solver(case)
    sa = gc.amd_solver(case, path="fused")
    rng = np.random.default_rng(5)
    sl = (slice(8, 14), slice(6, 12), slice(10, 18))
    ii, jj, kk = np.meshgrid(np.arange(8, 14), np.arange(6, 12), np.arange(10, 18), indexing="ij")
    flat = ii + Nx * jj + Nx * Ny * kk
    for c, name in enumerate("xyz"):
        ie = (1.0 / (rng.uniform(1.0, 5.0, ii.shape) * epsilon_0)).astype(dtype)
        sa.ieps[sl + (name,)] = ie
        so.ieps[c * N + flat] = ie
        if conductive:
            sg = rng.uniform(0.0, 2.0, ii.shape).astype(dtype)
            sa.sigma[sl + (name,)] = sg
            so.sigma[c * N + flat] = sg
    sa.update_tensors()
    so.assemble()
    so.step_0 = False  # update_tensors ends the first-step ghost zeroing (solverFIT3D.py:414)
    sa.fused_full_rows = rows
    line = (10, 8)
    vals = rng.standard_normal(Nz).astype(dtype)
    if conductive:
        so.J[:] = 0
        so.one_step()
        sa.path = "general"
        sa.one_step()
        sa.path = "fused"
        sa._beam_line = line  # a beam line with no table: its J is read back
        for _ in range(4):
            so.one_step()
        sa._run(sa._step_args(4))
    else:
        sa._write_beam_line(line[0], line[1], vals)
        so.J[2 * N + line[0] + Nx * line[1] + Nx * Ny * np.arange(Nz)] = vals
        for _ in range(4):
            so.one_step()
            sa.one_step()
    assert sa.last_path == "fused" and sa._code_escape and sa._nmat <= 255
    for name in ("E", "H", "J"):
        got, ref = getattr(sa, name).to_numpy(), getattr(so, name)
        assert np.array_equal(got, ref), f"{name}: rel-L2 {gc.rel_l2(got, ref):.3e}"


def _zero_masked_values(case):
    """Zero E on PEC-masked edges (and H on PMC-masked faces) of a random case."""
    dims = tuple(len(a) - 1 for a in (case["x"], case["y"], case["z"]))
    pec, pmc = fo.bc_column_row_masks(case["bc"][0], case["bc"][1], dims)
    if pec is not None:
        case["fields"][0] = np.where(pec == 0, 0, case["fields"][0]).astype(case["dtype"])
    if pmc is not None:
        case["fields"][1] = np.where(pmc == 0, 0, case["fields"][1]).astype(case["dtype"])



@pytest.mark.parametrize("bc", [PEC_ALL, PMC_MIX])
@pytest.mark.parametrize("dims", [(12, 10, 14), (64, 8, 33)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("conductive", [False, True])
@pytest.mark.parametrize("variant", [0, 2, 3, 7])
def test_fused_without_masks_matches_oracle(bc, dims, dtype, conductive, variant):
    """Masked E/H values zero: the kernel runs without in-kernel masks."""
    print("fused no-mask", bc[0], dims, np.dtype(dtype).name, conductive, variant, flush=True)
    case = gc.random_case(bc, dtype, conductive, seed=5 + sum(dims), dims=dims)
    case["fields"][2] = np.zeros_like(case["fields"][2])
    _zero_masked_values(case)
    so = gc.oracle_solver(case)
    sa = gc.amd_solver(case, path="fused")
    sa.fused_variant = variant
    if conductive:
        so.one_step()
        sa.path = "general"
        sa.one_step()
        sa.path = "fused"
    for _ in range(5):
        so.one_step()
        sa.one_step()
    assert sa.last_path == "fused"
    assert sa._masked_zero() and sa._code is not None
    for name in ("E", "H", "J"):
        got, ref = getattr(sa, name).to_numpy(), getattr(so, name)
        assert np.array_equal(got, ref), f"{name}: rel-L2 {gc.rel_l2(got, ref):.3e}"


def test_fused_falls_back_when_wraps_live():
    case = gc.random_case(BCS[2], np.float64, False, seed=3)
    sa = gc.amd_solver(case)
    so = gc.oracle_solver(case)
    for _ in range(3):
        so.one_step()
        sa.one_step()
    assert sa.last_path == "general"
    assert np.array_equal(sa.E.to_numpy(), so.E) and np.array_equal(sa.H.to_numpy(), so.H)
    with pytest.raises(RuntimeError):
        sb = gc.amd_solver(case, path="fused")
        sb.one_step()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("zbc", ["abc", "pml"])
@pytest.mark.parametrize("use_code", [True, False])
@pytest.mark.parametrize("variant", [0, 1, "full_rows"])
def test_fused_beam_loop_matches_oracle(dtype, zbc, use_code, variant):
    """Beam.update + one_step user loop on the cavity/pipe (beam-line J mode;
    with PML on z the conductive fused step), with and without the material
    code table."""
    import wakis_amd as wa
    print("beam loop", np.dtype(dtype).name, zbc, use_code, flush=True)
    x, y, z, mask = cases.cavity_pipe_geometry(24, 24, 40)
    bc_lo, bc_hi = ["pec", "pec", zbc], ["pec", "pec", zbc]
    so = fo.OracleSolver(x, y, z, bc_lo, bc_hi, bg="pec", solids=[(mask, "vacuum")],
                         dtype=dtype, n_pml=4)
    grid = wa.GridFIT3D(x=x, y=y, z=z, stl_solids={"cav": mask},
                        stl_materials={"cav": "vacuum"}, verbose=0)
    sa = wa.SolverFIT3D(grid, bc_low=bc_lo, bc_high=bc_hi, use_stl=True, bg="pec",
                        dtype=dtype, verbose=0, path="fused", n_pml=4)
    sa.use_material_code = use_code
    if variant == "full_rows":
        sa.fused_full_rows = 1
    else:
        sa.fused_variant = variant
        sa.fused_full_rows = -1
    assert sa.dt == so.dt
    bo, ba = fo.OracleBeam(q=1e-9, sigmaz=5e-3), wa.Beam(q=1e-9, sigmaz=5e-3)
    for n in range(80):
        bo.update(so, n * so.dt)
        ba.update(sa, n * sa.dt)
        so.one_step()
        sa.one_step()
    assert sa.last_path == "fused"
    assert (sa._code is not None) == use_code
    for name in ("E", "H", "J"):
        assert np.array_equal(getattr(sa, name).to_numpy(), getattr(so, name)), name
