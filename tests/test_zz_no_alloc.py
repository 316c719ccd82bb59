"""GPU: wf_count_nonfinite_noalloc (no handle, no allocation: the partial
counts in the library's static device scratch) against wf_count_nonfinite
(the handle's workspace), through the C ABI.

Two calls in a row, the second on a smaller input (the scratch keeps the
first call's partials; the second must not read them), FP32 and FP64, every
workgroup count up to the scratch's 1024 partials; more workgroups and bad
arguments are refused on the host before any launch.
"""

import ctypes

import numpy as np
import pytest
import torch

from oracle import cases

pytestmark = pytest.mark.gpu


def _handle_count(s, x):
    """wf_count_nonfinite through the solver's handle (the existing entry point)."""
    from wakis_amd import _lib
    out = torch.full((1,), -1, dtype=torch.int64, device="cuda")
    _lib.check(_lib.lib().wf_count_nonfinite(s._handle, _lib.ptr(x), x.numel(), _lib.ptr(out),
                                             _lib.stream_ptr()), "wf_count_nonfinite")
    return int(out.item())


def _noalloc_count(x, blocks=0):
    from wakis_amd import _lib
    code = _lib.WF_F64 if x.dtype == torch.float64 else _lib.WF_F32
    out = torch.full((1,), -1, dtype=torch.int64, device="cuda")
    _lib.check(_lib.lib().wf_count_nonfinite_noalloc(_lib.ptr(x), code, x.numel(), blocks,
                                                     _lib.ptr(out), _lib.stream_ptr()),
               "wf_count_nonfinite_noalloc")
    return int(out.item())


def _tiny_solver(dtype):
    import wakis_amd as wa
    x, y, z, mask = cases.cavity_pipe_geometry(8, 8, 8)
    grid = wa.GridFIT3D(x=x, y=y, z=z, stl_solids={"cav": mask},
                        stl_materials={"cav": "vacuum"}, verbose=0)
    return wa.SolverFIT3D(grid, bc_low=["pec", "pec", "abc"], bc_high=["pec", "pec", "abc"],
                          use_stl=True, bg="pec", dtype=dtype, verbose=0)


def _with_nonfinite(n, k, dtype, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(n).astype(dtype)
    idx = rng.choice(n, size=k, replace=False)
    a[idx[: k // 3]] = np.nan
    a[idx[k // 3: 2 * k // 3]] = np.inf
    a[idx[2 * k // 3:]] = -np.inf
    return torch.as_tensor(a, device="cuda"), int(np.count_nonzero(~np.isfinite(a)))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_noalloc_count_equals_handle_count_two_calls(dtype):
    print("noalloc count", np.dtype(dtype).name, flush=True)
    s = _tiny_solver(dtype)
    big, kb = _with_nonfinite(3_000_001, 517, dtype, seed=1)   # 1024 workgroups
    small, ks = _with_nonfinite(1_000, 7, dtype, seed=2)       # 4 workgroups
    for x, k in ((big, kb), (small, ks)):
        got = _noalloc_count(x)
        assert got == _handle_count(s, x) == k
    # the smaller call after the larger one, on a clean input: no stale partials
    clean = torch.zeros(300, dtype=big.dtype, device="cuda")
    assert _noalloc_count(big) == kb
    assert _noalloc_count(clean) == 0 == _handle_count(s, clean)


@pytest.mark.parametrize("blocks", [1, 3, 64, 1023, 1024])
def test_noalloc_count_any_workgroup_count(blocks):
    x, k = _with_nonfinite(200_003, 41, np.float32, seed=blocks)
    assert _noalloc_count(x, blocks) == k


def test_noalloc_count_empty_and_refusals():
    from wakis_amd import _lib
    L = _lib.lib()
    empty = torch.zeros(0, dtype=torch.float64, device="cuda")
    assert _noalloc_count(empty) == 0
    x, _ = _with_nonfinite(1000, 3, np.float64, seed=5)
    out = torch.zeros(1, dtype=torch.int64, device="cuda")
    st = _lib.stream_ptr()
    assert L.wf_count_nonfinite_noalloc(_lib.ptr(x), _lib.WF_F64, 1000, 1025, _lib.ptr(out),
                                        st) == _lib.WF_ERANGE
    assert b"at most 1024" in L.wf_last_error()
    assert L.wf_count_nonfinite_noalloc(_lib.ptr(x), _lib.WF_F64, -1, 0, _lib.ptr(out),
                                        st) == _lib.WF_ERANGE
    assert L.wf_count_nonfinite_noalloc(_lib.ptr(x), 7, 1000, 0, _lib.ptr(out),
                                        st) == _lib.WF_EINVAL
    assert L.wf_count_nonfinite_noalloc(_lib.ptr(x), _lib.WF_F64, 1000, 0, None,
                                        st) == _lib.WF_EINVAL
    with pytest.raises(IndexError):
        _noalloc_count(x, 4096)
