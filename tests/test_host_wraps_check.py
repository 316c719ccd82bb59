"""CPU: the fused path's wrap-term check on shifted boundary planes
(``SolverFIT3D._wraps_dead_now``) against the gathered-index form
(``tests/wraps_reference.py``) on random small set-ups: grids down to one cell
per axis, every PEC and PMC face set, coefficient and field arrays with
scattered zeros, edge lengths with zeros and beam lines on the boundary
planes. The check runs on host tensors here; the GPU test repeats it on real
solvers.
"""

from types import SimpleNamespace

import numpy as np
import torch

from tests.wraps_reference import wraps_dead_gather
from wakis_amd.solverFIT3D import SolverFIT3D


def _stub(rng, dims):
    Nx, Ny, Nz = dims
    N = Nx * Ny * Nz

    def arr(p):  # nonzero with probability p
        v = rng.standard_normal(3 * N) * (rng.random(3 * N) < p)
        return torch.as_tensor(v)

    f = lambda p: SimpleNamespace(array=arr(p))  # noqa: E731
    line = None
    if rng.random() < 0.5:
        line = (int(rng.integers(0, Nx)), int(rng.integers(0, Ny)))
        if rng.random() < 0.5:
            line = (0, line[1]) if rng.random() < 0.5 else (line[0], 0)
    tl = [rng.random(n) * (rng.random(n) < 0.8) for n in dims]
    return SimpleNamespace(Nx=Nx, Ny=Ny, Nz=Nz, N=N, device="cpu", _a=arr(0.6), _b=arr(0.6),
                           E=f(0.2), H=f(0.2), J=f(0.3), ieps=f(0.7), _beam_line=line,
                           pec_zero=int(rng.integers(0, 64)), pmc_zero=int(rng.integers(0, 64)),
                           _tL1=tl)


def test_plane_check_equals_gathered_form():
    rng = np.random.default_rng(2024)
    outcomes = []
    for _ in range(600):
        dims = tuple(int(rng.integers(1, 6)) for _ in range(3))
        s = _stub(rng, dims)
        got, ref = SolverFIT3D._wraps_dead_now(s), wraps_dead_gather(s)
        assert got == ref, (dims, s.pec_zero, s.pmc_zero, s._beam_line)
        outcomes.append(got)
    assert 0.05 < np.mean(outcomes) < 0.95  # both answers occur often
