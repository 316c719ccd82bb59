"""GPU: the fused path's wrap check (``SolverFIT3D._wraps_dead``) on boundary
planes against the same test written with gathered index vectors.

The check decides whether the fused kernels may drop the flat-index wrap
terms of the reference operator (i = Nx-1 to (0, j+1, k), j = Ny-1 to
(i, 0, k+1) and the transposes). ``_wraps_dead_now`` forms rows and columns as
two boundary planes shifted by one flat index; ``_gather`` builds every
row's column from explicit (i, j, k) coordinates (``tests/wraps_reference.py``).
Both must agree on every boundary set-up and on field states that make the
terms live or dead.
"""

import numpy as np
import pytest

from tests import gpu_cases as gc
from tests.test_stencil_rules import BCS
from tests.wraps_reference import wraps_dead_gather as _gather

pytestmark = pytest.mark.gpu


def _set(s, E=None, H=None, J=None):
    for F, v in ((s.E, E), (s.H, H), (s.J, J)):
        if v is not None:
            F.fromarray(np.ascontiguousarray(v.reshape(-1)))


@pytest.mark.parametrize("dims", [(12, 10, 14), (7, 5, 6)])
@pytest.mark.parametrize("conductive", [False, True])
def test_wrap_check_planes_equal_gather(conductive, dims):
    print("wrap check", conductive, dims, flush=True)
    outcomes = []
    for bci, bc in enumerate(BCS):
        case = gc.random_case(bc, np.float64, conductive, seed=5 + bci, dims=dims)
        s = gc.amd_solver(case)
        Nx, Ny, Nz = dims
        rng = np.random.default_rng(bci)
        zero = np.zeros((3, Nz, Ny, Nx))
        states = [("random", None)]
        states.append(("zero", dict(E=zero, H=zero, J=zero)))
        for plane in ("i0", "i1", "j0", "j1"):
            for which in ("E", "H"):
                v = zero.copy()
                sl = {"i0": (slice(None), slice(None), slice(None), 0),
                      "i1": (slice(None), slice(None), slice(None), Nx - 1),
                      "j0": (slice(None), slice(None), 0, slice(None)),
                      "j1": (slice(None), slice(None), Ny - 1, slice(None))}[plane]
                v[sl] = rng.standard_normal(v[sl].shape)
                st = dict(E=zero, H=zero, J=zero)
                st[which] = v
                states.append((plane + which, st))
        for name, st in states:
            if st is not None:
                _set(s, **st)
            for line in (None, (0, Ny // 2), (Nx // 2, 0), (0, 0)):
                s._beam_line = line
                got, ref = s._wraps_dead_now(), _gather(s)
                assert got == ref, (bc, name, line)
                outcomes.append(got)
    assert any(outcomes) and not all(outcomes)
