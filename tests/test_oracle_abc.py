"""CPU: pins the oracle's emsolve ABC (``abc_copy_low_faces``) against the
reference's snapshot protocol run literally on numpy.

The reference saves the index-1 plane of every low ABC face and the index -1
plane of every high ABC face (boundaries.py:313-367), keeps two such
snapshots, and after each step writes the one taken two steps earlier back
onto the faces (routines.py:183-221, boundaries.py:369-410). The oracle
claims that on numpy this reduces to "copy the current index-1 plane onto
index 0 of every low ABC face, leave high faces alone": the snapshots are
views of the field arrays (``Field.__getitem__`` reshapes in F order,
field.py:223-262) and ``one_step`` writes those arrays in place
(``Field.fromarray``: ``array[:] = ...``, field.py:128-137). The emulation
below follows the protocol step by step (snapshot dictionaries, n-2 / n-1
rotation, face order) with views taken the way the reference takes them; a
protocol with copied snapshots, a true two-step delay, gives other fields,
so the check tells the two readings apart.
"""

import numpy as np
import pytest

from oracle import fit_oracle as fo

_AXES = ("x", "y", "z")


def _plane(V, axis, idx):
    sl = [slice(None)] * 3
    sl[axis] = idx
    return V[tuple(sl)]


def _get_abc(s, copy):
    """Snapshot dictionaries {axis: {comp+'lo'|'hi': plane}} for E and H."""
    out = []
    for F in (s.E, s.H):
        snap = {}
        for side, idx, bcs in (("lo", 1, s.bc_low), ("hi", -1, s.bc_high)):
            for axis in range(3):
                if bcs[axis].lower() != "abc":
                    continue
                snap.setdefault(axis, {})
                for d in _AXES:
                    p = _plane(fo.comp(F, s.dims, d), axis, idx)
                    snap[axis][d + side] = p.copy() if copy else p
        out.append(snap)
    return out


def _update_abc(s, E_abc, H_abc):
    """Writes the snapshots back: low faces x, y, z, then high faces x, y, z;
    per face and component E then H."""
    for side, idx, bcs in (("lo", 0, s.bc_low), ("hi", -1, s.bc_high)):
        for axis in range(3):
            if bcs[axis].lower() != "abc":
                continue
            for d in _AXES:
                for F, snap in ((s.E, E_abc), (s.H, H_abc)):
                    V = fo.comp(F, s.dims, d)
                    sl = [slice(None)] * 3
                    sl[axis] = idx
                    V[tuple(sl)] = snap[axis][d + side]


def _emsolve_protocol(s, Nt, copy):
    E2, H2 = _get_abc(s, copy)
    E1, H1 = _get_abc(s, copy)
    for _ in range(Nt):
        s.one_step()
        _update_abc(s, E2, H2)
        E2, H2 = E1, H1
        E1, H1 = _get_abc(s, copy)


def _solver(bc_low, bc_high, dtype, seed):
    x, y, z = fo.uniform_axes(-1.0, 1.0, -1.0, 1.0, 0.0, 2.0, 7, 6, 9)
    s = fo.OracleSolver(x, y, z, bc_low, bc_high, bg=[1.0, 1.0], dtype=dtype)
    rng = np.random.default_rng(seed)
    s.E[:] = rng.standard_normal(s.E.shape).astype(dtype)
    s.H[:] = rng.standard_normal(s.H.shape).astype(dtype)
    return s


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("bcs", [
    (["abc", "abc", "abc"], ["abc", "abc", "abc"]),
    (["pec", "pec", "abc"], ["pec", "pec", "abc"]),
    (["abc", "periodic", "pec"], ["pec", "periodic", "abc"]),
])
def test_abc_copy_equals_reference_snapshot_protocol(dtype, bcs):
    Nt = 12
    a = _solver(bcs[0], bcs[1], dtype, 3)
    b = _solver(bcs[0], bcs[1], dtype, 3)
    assert a.activate_abc
    _emsolve_protocol(a, Nt, copy=False)
    fo.emsolve(b, Nt)
    assert np.array_equal(a.E, b.E) and np.array_equal(a.H, b.H)

    # the snapshots really are views of the field arrays
    (E1, _) = _get_abc(a, copy=False)
    axis = next(iter(E1))
    key = next(iter(E1[axis]))
    assert np.shares_memory(E1[axis][key], a.E)


def test_copied_snapshots_give_other_fields():
    """A true two-step delay (copied snapshots) is a different boundary: the
    check above would fail an oracle written that way."""
    bcs = (["pec", "pec", "abc"], ["pec", "pec", "abc"])
    a = _solver(*bcs, np.float64, 5)
    b = _solver(*bcs, np.float64, 5)
    _emsolve_protocol(a, 8, copy=True)
    fo.emsolve(b, 8)
    assert not np.array_equal(a.E, b.E)


def test_high_abc_face_is_left_to_the_step():
    """On a high ABC face the write-back is the identity: the fields equal a
    run with no face copies at all when only high faces are ABC."""
    bcs = (["pec", "pec", "pec"], ["pec", "pec", "abc"])
    a = _solver(*bcs, np.float64, 7)
    b = _solver(*bcs, np.float64, 7)
    assert a.activate_abc
    fo.emsolve(a, 10)
    for _ in range(10):
        b.one_step()
    assert np.array_equal(a.E, b.E) and np.array_equal(a.H, b.H)
