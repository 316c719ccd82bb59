"""Shared builders for GPU parity tests: the same case as an oracle solver and
as a wakis_amd solver (tests only)."""

import numpy as np

from oracle import fit_oracle as fo


def random_case(bc, dtype, conductive, seed, dims=(12, 10, 14)):
    rng = np.random.default_rng(seed)
    Nx, Ny, Nz = dims
    x = np.cumsum(np.r_[-6e-3, rng.uniform(0.8e-3, 1.2e-3, Nx)])
    y = np.cumsum(np.r_[-5e-3, rng.uniform(0.8e-3, 1.2e-3, Ny)])
    z = np.cumsum(np.r_[-7e-3, rng.uniform(0.8e-3, 1.2e-3, Nz)])
    diel = rng.random(dims) < 0.3
    pec = rng.random(dims) < 0.15
    solids = [("diel", diel, [4.0, 1.5]), ("metal", pec, "pec")]
    bg = [1.0, 1.0, 50.0] if conductive else [1.0, 1.0]
    fields = [rng.standard_normal(3 * Nx * Ny * Nz).astype(dtype) for _ in range(3)]
    return dict(x=x, y=y, z=z, bc=bc, bg=bg, solids=solids, dtype=dtype, fields=fields,
                n_pml=2)


def oracle_solver(case):
    s = fo.OracleSolver(case["x"], case["y"], case["z"], case["bc"][0], case["bc"][1],
                        bg=case["bg"], solids=[(m, mat) for _, m, mat in case["solids"]],
                        dtype=case["dtype"], n_pml=case["n_pml"])
    s.E[:], s.H[:], s.J[:] = case["fields"]
    return s


def amd_solver(case, **kw):
    import wakis_amd as wa
    grid = wa.GridFIT3D(x=case["x"], y=case["y"], z=case["z"],
                        stl_solids={k: m for k, m, _ in case["solids"]},
                        stl_materials={k: mat for k, _, mat in case["solids"]}, verbose=0)
    s = wa.SolverFIT3D(grid, bc_low=case["bc"][0], bc_high=case["bc"][1], use_stl=True,
                       bg=case["bg"], dtype=case["dtype"], n_pml=case["n_pml"], verbose=0, **kw)
    s.E.fromarray(case["fields"][0])
    s.H.fromarray(case["fields"][1])
    if np.any(case["fields"][2]):  # a zero J stays in the "no source" state
        s.J.fromarray(case["fields"][2])
    return s


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))
