"""CPU: the general-path stencil rules reproduce the oracle's CSR step bit for bit.

The numpy model in stencil_model.py restates the kernel's term tables, face
masks, flat wraps and summation order; here it is checked against the
oracle's scipy assembly on small grids for many boundary combinations, both
dtypes, with and without conductivity and a live J.
"""

import itertools

import numpy as np
import pytest

from oracle import fit_oracle as fo
from tests import stencil_model as sm

BCS = [
    (["pec", "pec", "pec"], ["pec", "pec", "pec"]),
    (["pec", "pec", "abc"], ["pec", "pec", "abc"]),
    (["periodic", "periodic", "pec"], ["periodic", "periodic", "pml"]),
    (["pmc", "pec", "periodic"], ["magnetic", "pec", "periodic"]),
    (["PMC", "abc", "pec"], ["pmc", "abc", "electric"]),
    (["periodic", "pmc", "abc"], ["periodic", "pmc", "abc"]),
    (["abc", "abc", "abc"], ["abc", "abc", "abc"]),
]


def build(bc, dtype, conductive, seed, dims=(6, 5, 7)):
    rng = np.random.default_rng(seed)
    Nx, Ny, Nz = dims
    x = np.cumsum(np.r_[0.0, rng.uniform(0.8e-3, 1.2e-3, Nx)])
    y = np.cumsum(np.r_[0.0, rng.uniform(0.8e-3, 1.2e-3, Ny)])
    z = np.cumsum(np.r_[0.0, rng.uniform(0.8e-3, 1.2e-3, Nz)])
    diel = rng.random((Nx, Ny, Nz)) < 0.3
    pec = rng.random((Nx, Ny, Nz)) < 0.15
    solids = [(diel, [4.0, 1.5]), (pec, "pec")]
    bg = [1.0, 1.0, 50.0] if conductive else [1.0, 1.0]
    s = fo.OracleSolver(x, y, z, bc[0], bc[1], bg=bg, solids=solids, dtype=dtype, n_pml=2)
    s.set_ghosts_to_0()
    s.step_0 = False
    for F in (s.E, s.H, s.J):
        F[:] = rng.standard_normal(F.shape).astype(dtype)
    return s


def model_coeffs(s):
    S = s.dtype
    N = s.N
    pecz = 0 if s.pec_mask is None else _bits(s, "pec")
    pmcz = 0 if s.pmc_mask is None else _bits(s, "pmc")
    a = s.iA.astype(S) * s.imu
    b = s.ieps * s.itA.astype(S)
    if s.pmc_mask is not None:
        a = a * s.pmc_mask.astype(S)
    if s.pec_mask is not None:
        b = b * s.pec_mask.astype(S)
    L1 = [np.diff(s.x_nodes), np.diff(s.y_nodes), np.diff(s.z_nodes)]
    tL1 = [fo.comp(s.tL, s.dims, "x")[:, 0, 0].copy(), fo.comp(s.tL, s.dims, "y")[0, :, 0].copy(),
           fo.comp(s.tL, s.dims, "z")[0, 0, :].copy()]
    m1desc = int(s.pec_mask is not None) ^ int(s.pmc_mask is not None)
    assert a.dtype == S and b.dtype == S and len(a) == 3 * N
    return a, b, L1, tL1, pecz, pmcz, m1desc


def _bits(s, kind):
    keys = ("electric", "pec", "pml")
    bits = 0
    for ax in range(3):
        lo, hi = s.bc_low[ax], s.bc_high[ax]
        if kind == "pec":
            bits |= (lo.lower() in keys) << ax
            bits |= (hi.lower() in keys) << (3 + ax)
        else:
            bits |= (lo.lower() in keys or lo.lower() == "magnetic" or lo == "pmc") << ax
            bits |= (hi.lower() in keys or hi.lower() == "magnetic" or hi == "pmc") << (3 + ax)
    if kind == "pmc" and s.pec_mask is None:
        for ax in range(3):
            for sh, b in ((ax, s.bc_low[ax]), (3 + ax, s.bc_high[ax])):
                if b.lower() in keys:
                    bits &= ~(1 << sh)
    return bits


@pytest.mark.parametrize("bci,dtype,conductive",
                         list(itertools.product(range(len(BCS)), [np.float64, np.float32],
                                                [False, True])))
def test_model_matches_csr(bci, dtype, conductive):
    s = build(BCS[bci], dtype, conductive, seed=bci * 7 + conductive)
    a, b, L1, tL1, pecz, pmcz, m1desc = model_coeffs(s)
    E, H, J = s.E.copy(), s.H.copy(), s.J.copy()
    for _ in range(3):
        s.one_step()
        sm.step(E, H, J, a, b, s.ieps, s.sigma, s.dt, s.dims, L1, tL1, pecz, pmcz, m1desc,
                jfull=True, cond=s.use_conductivity)
    assert np.array_equal(H, s.H), "H differs from the CSR step"
    assert np.array_equal(E, s.E), "E differs from the CSR step"
    assert np.array_equal(J, s.J)


@pytest.mark.parametrize("bci", range(len(BCS)))
def test_m1_column_order_rule(bci):
    """scipy's M1 rows run descending iff exactly one of the PEC/PMC masks applied."""
    s = build(BCS[bci], np.float64, False, seed=3)
    desc = int(s.pec_mask is not None) ^ int(s.pmc_mask is not None)
    M1, M2 = s.M1, s.M2
    for M, want_desc in ((M1, desc), (M2, 0)):
        for r in range(M.shape[0]):
            cols = M.indices[M.indptr[r]:M.indptr[r + 1]]
            if len(cols) > 1:
                d = np.diff(cols)
                assert np.all(d < 0) if want_desc else np.all(d > 0), (r, cols)
