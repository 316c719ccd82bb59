"""CPU pins of the SIBC substitution (solverFIT3D.py:445-467, gridFIT3D.py:518-533).

The surface rule (nonzero finite-difference gradient of the solid's 0/1 cell
field) is checked against its closed form on boxes: the marked set is the box
grown by one cell along one axis at a time, less the box shrunk by one cell on
every axis. The material values are checked against sigma = eps = 1/Z_s with
Z_s = sqrt(pi fmax mu / sigma). The engine's host-side marking must equal the
oracle's on random masks. How pyvista's gradient treats edges and corners is
not checked here (pyvista is not available): the rule is parity unpinned.
"""

import numpy as np
import pytest
from scipy.constants import epsilon_0 as eps_0
from scipy.constants import mu_0

from oracle import fit_oracle as fo


def _box_mask(shape, lo, hi):
    m = np.zeros(shape, dtype=bool)
    m[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = True
    return m


def _box_surface_closed_form(shape, lo, hi):
    out = np.zeros(shape, dtype=bool)
    for ax in range(3):
        l, h = list(lo), list(hi)
        l[ax], h[ax] = max(lo[ax] - 1, 0), min(hi[ax] + 1, shape[ax])
        out |= _box_mask(shape, l, h)
    core = _box_mask(shape, [a + 1 for a in lo], [b - 1 for b in hi])
    return out & ~core


@pytest.mark.parametrize("lo,hi", [((3, 4, 5), (9, 8, 12)), ((2, 2, 2), (4, 4, 4)),
                                   ((5, 1, 3), (8, 11, 14))])
def test_surface_of_box_is_the_two_cell_shell(lo, hi):
    shape = (12, 12, 16)
    got = fo.sibc_surface(_box_mask(shape, lo, hi))
    assert np.array_equal(got, _box_surface_closed_form(shape, lo, hi))


def test_surface_at_domain_edge_uses_one_sided_difference():
    # a slab touching x = 0: the first cell compares with its one neighbour
    shape = (10, 6, 6)
    m = _box_mask(shape, (0, 0, 0), (4, 6, 6))
    got = fo.sibc_surface(m)
    want = np.zeros(shape, dtype=bool)
    want[3:5] = True  # last slab layer and the first outside layer
    assert np.array_equal(got, want)
    assert not fo.sibc_surface(np.ones(shape, dtype=bool)).any()
    assert not fo.sibc_surface(np.zeros(shape, dtype=bool)).any()


def test_marked_cells_sit_on_the_solid_boundary():
    rng = np.random.default_rng(5)
    m = rng.random((9, 10, 11)) < 0.4
    got = fo.sibc_surface(m)
    p = np.pad(m, 1, mode="edge")
    c = p[1:-1, 1:-1, 1:-1]
    differs = np.zeros_like(m)
    for ax in range(3):
        for s in (-1, 1):
            differs |= np.roll(p, s, axis=ax)[1:-1, 1:-1, 1:-1] != c
    assert got.any() and not (got & ~differs).any()


def test_engine_surface_marking_equals_oracle():
    import wakis_amd as wa
    rng = np.random.default_rng(11)
    for shape in [(7, 8, 9), (2, 5, 6), (12, 3, 10)]:
        m = rng.random(shape) < 0.5
        x, y, z = (np.linspace(0.0, 1.0, n + 1) for n in shape)
        g = wa.GridFIT3D(x=x, y=y, z=z, stl_solids={"s": m}, stl_materials={"s": "copper"},
                         verbose=0)
        got = g._mark_cells_in_surface("s")
        assert got.dtype == bool and np.array_equal(got, fo.sibc_surface(m))
        assert np.array_equal(g.grid["s"], got)


def test_sibc_material_values_and_time_step():
    shape = (12, 12, 16)
    x, y, z = fo.uniform_axes(-6e-3, 6e-3, -6e-3, 6e-3, 0.0, 16e-3, *shape)
    lo, hi = (3, 3, 4), (9, 9, 12)
    box = _box_mask(shape, lo, hi)
    sig, fmax = 5.8e7, 3e9
    s = fo.OracleSolver(x, y, z, ["pec"] * 3, ["pec"] * 3, solids=[(box, [1.0, 1.0, sig])],
                        use_sibc=True, fmax=fmax)
    plain = fo.OracleSolver(x, y, z, ["pec"] * 3, ["pec"] * 3)
    dn = np.sqrt(2) * 1e-3
    assert sig > 10 / (np.pi * fmax * mu_0 * dn ** 2)
    zs = np.sqrt(np.pi * fmax * mu_0 / sig)
    surf = _box_surface_closed_form(shape, lo, hi)
    for arr, on in ((s.sigma, 1 / zs), (s.ieps, zs)):
        for c in range(3):
            a3 = np.reshape(arr[c * s.N:(c + 1) * s.N], shape, order="F")
            assert np.allclose(a3[surf], on, rtol=1e-15, atol=0)
    for c in range(3):
        sl = slice(c * s.N, (c + 1) * s.N)
        s3 = np.reshape(s.sigma[sl], shape, order="F")
        e3 = np.reshape(s.ieps[sl], shape, order="F")
        e0 = np.reshape(plain.ieps[sl], shape, order="F")
        assert not s3[~surf].any()
        assert np.array_equal(e3[~surf], e0[~surf])
        assert np.allclose(e0, 1 / eps_0, rtol=1e-15)
    # tau = eps / sigma = 1 s on the surface: the CFL step stands
    assert s.dt == plain.dt
    # below the threshold (or without SIBC) the volume conductivity stays
    v = fo.OracleSolver(x, y, z, ["pec"] * 3, ["pec"] * 3, solids=[(box, [1.0, 1.0, sig])],
                        use_sibc=False, fmax=fmax)
    s3 = np.reshape(v.sigma[:v.N], shape, order="F")
    assert np.all(s3[box] == sig) and not s3[~box].any()
