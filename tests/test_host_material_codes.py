"""CPU: the dense material codes of the fused path's coefficient table.

The code of a cell is the rank of its tuple of nine column values (imu, ieps,
sigma per component) in lexicographic order among the distinct tuples: checked
against that definition by brute force, with the mixed-radix combination forced
to compact at every column and never, and with a column over the code limit.
"""

import numpy as np
import pytest
import torch

from wakis_amd.solverFIT3D import material_codes


def _brute(cols):
    tuples = list(zip(*[c.tolist() for c in cols]))
    order = {t: r for r, t in enumerate(sorted(set(tuples)))}
    return np.array([order[t] for t in tuples]), len(order)


@pytest.mark.parametrize("seed", range(6))
def test_codes_are_lexicographic_ranks(seed, monkeypatch):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 4000))
    cols = [torch.as_tensor(rng.integers(0, int(rng.integers(1, 40)), n) * 0.25 - 3.0,
                            dtype=torch.float32) for _ in range(9)]
    ref, k = _brute(cols)
    code, K = material_codes(cols)
    assert K == k and np.array_equal(code.numpy(), ref)


def test_codes_with_forced_compaction():
    """Columns whose distinct counts multiply past 2^62 (eight columns of 256
    values and one more): the radix product is ranked again on the way."""
    rng = np.random.default_rng(7)
    n = 3000
    cols = [torch.as_tensor(rng.permutation(np.resize(np.arange(256), n)) * 1.5,
                            dtype=torch.float64) for _ in range(9)]
    ref, k = _brute(cols)
    code, K = material_codes(cols)
    assert K == k == n and np.array_equal(code.numpy(), ref)


def test_codes_with_single_valued_columns():
    """Columns with one value (imu, or sigma of a lossless grid) are skipped."""
    rng = np.random.default_rng(11)
    n = 2500
    cols = [torch.full((n,), 1.25), torch.as_tensor(rng.integers(0, 5, n) * 0.5),
            torch.zeros(n), torch.as_tensor(rng.integers(0, 3, n) * 2.0), torch.full((n,), -7.0),
            torch.as_tensor(rng.integers(0, 4, n) * 1.0), torch.zeros(n), torch.zeros(n),
            torch.as_tensor(rng.integers(0, 2, n) * 3.0)]
    ref, k = _brute(cols)
    code, K = material_codes(cols)
    assert K == k and np.array_equal(code.numpy(), ref)
    code, K = material_codes([torch.full((9,), 2.0)] * 9)
    assert K == 1 and not code.any()


def test_column_over_the_limit():
    cols = [torch.arange(300, dtype=torch.float32)] + [torch.zeros(300)] * 8
    code, K = material_codes(cols)
    assert code is None and K == 257
    code, K = material_codes(cols, kmax=300)
    assert K == 300 and torch.equal(code, torch.arange(300))
