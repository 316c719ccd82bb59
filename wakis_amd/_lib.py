"""ctypes binding of libwakis_fit.so (the C ABI in include/wakis_fit.h).

The library is built in-tree by ``wakis_amd.build``. There is no fallback:
if the shared library is missing or a call fails, an exception is raised.
"""

import ctypes
from ctypes import POINTER, c_double, c_int32, c_int64, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().with_name("libwakis_fit.so")

WF_OK, WF_EINVAL, WF_ERANGE, WF_EHIP, WF_ESTATE = 0, -1, -2, -3, -4
WF_F64, WF_F32 = 0, 1
WF_J_NONE, WF_J_LINE, WF_J_FULL = 0, 1, 2


class WfGrid(ctypes.Structure):
    _fields_ = [
        ("Nx", c_int32), ("Ny", c_int32), ("Nz", c_int32), ("dtype", c_int32),
        ("pec_zero_faces", c_int32), ("pmc_zero_faces", c_int32), ("m1_descending", c_int32),
        ("Lx", POINTER(c_double)), ("Ly", POINTER(c_double)), ("Lz", POINTER(c_double)),
        ("tLx", POINTER(c_double)), ("tLy", POINTER(c_double)), ("tLz", POINTER(c_double)),
    ]


class WfStepArgs(ctypes.Structure):
    _fields_ = [
        ("E", c_void_p), ("H", c_void_p), ("J", c_void_p),
        ("a", c_void_p), ("b", c_void_p), ("ieps", c_void_p), ("sigma", c_void_p),
        ("dt", c_double), ("nsteps", c_int32), ("jmode", c_int32), ("conductive", c_int32),
        ("ixs", c_int32), ("iys", c_int32),
        ("beam_table", c_void_p), ("beam_row0", c_int64), ("beam_rows", c_int64),
        ("capture", c_void_p),
        ("cap_ix0", c_int32), ("cap_iy0", c_int32), ("cap_nx", c_int32), ("cap_ny", c_int32),
        ("cap_samples", c_int64), ("cap_sample0", c_int64),
        ("path", c_int32), ("compute_f64", c_int32),
        ("E2", c_void_p), ("H2", c_void_p), ("kz", c_int32), ("result_swapped", c_int32),
        ("code", c_void_p), ("mat", c_void_p), ("metric2d", c_void_p), ("nmat", c_int32),
        ("variant", c_int32), ("zero_masked", c_int32),
        ("traj_E", c_void_p), ("traj_H", c_void_p),
        ("traj_lo", c_int32 * 3), ("traj_hi", c_int32 * 3),
        ("traj_slots", c_int64), ("traj_slot0", c_int64), ("cells_per_lane", c_int32),
        ("full_rows", c_int32), ("box_lo", c_int32 * 3), ("box_hi", c_int32 * 3),
        ("code_escape", c_int32),
        ("mon_acc", c_void_p), ("mon_phase", c_void_p), ("mon_row0", c_int64),
        ("mon_nf", c_int32), ("mon_phase_f32", c_int32), ("two_step", c_int32),
        ("two_step_used", c_int32),
    ]


class WfAdjointArgs(ctypes.Structure):
    _fields_ = [
        ("W", c_void_p), ("lH", c_void_p), ("lJ_region", c_void_p),
        ("E0", c_void_p), ("E1", c_void_p), ("H1", c_void_p), ("beam_row", c_void_p),
        ("a", c_void_p), ("b", c_void_p), ("ieps", c_void_p), ("sigma", c_void_p),
        ("metric2d", c_void_p), ("dt", c_double), ("conductive", c_int32),
        ("ixs", c_int32), ("iys", c_int32), ("r0", c_int32 * 3), ("r1", c_int32 * 3),
        ("g_ieps", c_void_p), ("g_sigma", c_void_p), ("gcap", c_void_p),
        ("cap_ix0", c_int32), ("cap_iy0", c_int32), ("cap_nx", c_int32), ("cap_ny", c_int32),
        ("cap_samples", c_int64), ("cap_sample", c_int64), ("traj", c_int32),
        ("fused", c_int32), ("W2", c_void_p), ("lH2", c_void_p), ("fields_f32", c_int32),
        ("reg_W", c_void_p), ("reg_lJ", c_void_p), ("reg_slot", c_int64),
        ("reg_slots", c_int64), ("kappa", c_void_p),
        ("code", c_void_p), ("mat", c_void_p), ("nmat", c_int32), ("code_escape", c_int32),
        ("code_metric2d", c_void_p), ("compute_f32", c_int32),
    ]


class WfAdjointGradArgs(ctypes.Structure):
    """wf_adjoint_grad_args of include/wakis_fit.h."""
    _fields_ = [
        ("E", c_void_p), ("H", c_void_p), ("E_stride", c_int64), ("H_stride", c_int64),
        ("traj", c_int32), ("beam", c_void_p), ("sigma", c_void_p), ("metric2d", c_void_p),
        ("reg_W", c_void_p), ("reg_lJ", c_void_p), ("fields_f32", c_int32),
        ("nsteps", c_int32), ("dt", c_double), ("conductive", c_int32), ("ixs", c_int32),
        ("iys", c_int32), ("r0", c_int32 * 3), ("r1", c_int32 * 3),
        ("g_ieps", c_void_p), ("g_sigma", c_void_p),
    ]


EXPORTS = {
    "wf_version": (c_int32, []),
    "wf_last_error": (ctypes.c_char_p, []),
    "wf_create": (c_int32, [POINTER(WfGrid), POINTER(c_void_p)]),
    "wf_destroy": (c_int32, [c_void_p]),
    "wf_zero_ghosts": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "wf_step": (c_int32, [c_void_p, POINTER(WfStepArgs), c_void_p]),
    "wf_capture": (c_int32, [c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_int32, c_int32,
                             c_int64, c_int64, c_void_p]),
    "wf_abc_low": (c_int32, [c_void_p, c_void_p, c_void_p, c_int32, c_void_p]),
    "wf_wake_wp": (c_int32, [c_void_p, c_int32, c_int32, c_int64, c_int32, c_int32, c_int32,
                             c_void_p, c_void_p, c_void_p, c_int64, c_double, c_double,
                             c_double, c_double, c_double, c_int64, c_int64, c_int32,
                             c_double, c_void_p, c_void_p]),
    "wf_wake_trans": (c_int32, [c_void_p, c_int64, c_double, c_double, c_double, c_void_p,
                                c_void_p, c_void_p]),
    "wf_dft": (c_int32, [c_void_p, c_int32, c_int64, c_int64, c_int64, c_void_p, c_void_p]),
    "wf_energy": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                            c_void_p, c_void_p, c_void_p]),
    "wf_loss_factor": (c_int32, [c_void_p, c_void_p, c_int64, c_double, c_void_p, c_void_p]),
    "wf_count_nonfinite": (c_int32, [c_void_p, c_void_p, c_int64, c_void_p, c_void_p]),
    "wf_count_nonfinite_noalloc": (c_int32, [c_void_p, c_int32, c_int64, c_int32, c_void_p,
                                             c_void_p]),
    "wf_batch_create": (c_int32, [POINTER(c_void_p), POINTER(WfStepArgs), c_int32, c_int32,
                                  c_int32, POINTER(c_void_p)]),
    "wf_batch_run": (c_int32, [c_void_p, c_int32, c_int32, c_int32, c_void_p]),
    "wf_batch_destroy": (c_int32, [c_void_p]),
    "wf_adjoint_step": (c_int32, [c_void_p, POINTER(WfAdjointArgs), c_void_p]),
    "wf_adjoint_grad": (c_int32, [c_void_p, POINTER(WfAdjointGradArgs), c_void_p]),
    "wf_batch_adjoint_create": (c_int32, [POINTER(c_void_p), POINTER(WfAdjointArgs),
                                          POINTER(c_int32), c_int32, POINTER(c_void_p)]),
    "wf_batch_adjoint_run": (c_int32, [c_void_p, c_int32, c_int32, c_int32, c_void_p]),
    "wf_batch_adjoint_destroy": (c_int32, [c_void_p]),
    "wf_trajectory_gather": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                       POINTER(c_int32), POINTER(c_int32), c_int64, c_int64,
                                       c_void_p]),
    "wf_wake_wp_adjoint": (c_int32, [c_void_p, c_int32, c_int64, c_int32, c_int32, c_int32,
                                     c_void_p, c_void_p, c_void_p, c_int64, c_double, c_double,
                                     c_double, c_double, c_double, c_int64, c_int32, c_void_p,
                                     c_void_p]),
    "wf_idft_real": (c_int32, [c_void_p, c_int64, c_int64, c_int64, c_void_p, c_void_p]),
    "wf_field_monitor_update": (c_int32, [c_void_p, c_int32, c_int64, c_void_p, c_int32,
                                          c_int32, c_void_p, c_void_p]),
}

_lib = None


def lib():
    """Load libwakis_fit.so once; raise ImportError if it is not built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m wakis_amd.build` "
                "(the HIP engine has no CPU fallback)")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in EXPORTS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc, what=""):
    """Map a WF_* status to the reference's exception classes (solverFIT3D.py, field.py)."""
    if rc == WF_OK:
        return
    msg = lib().wf_last_error().decode(errors="replace")
    if what:
        msg = f"{what}: {msg}"
    if rc == WF_EINVAL:
        raise ValueError(msg)
    if rc == WF_ERANGE:
        raise IndexError(msg)
    raise RuntimeError(msg)


def ptr(t):
    """Device pointer of a torch tensor (or None)."""
    return None if t is None else c_void_p(t.data_ptr())


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return c_void_p(s.cuda_stream)
