"""FieldMonitor: running DFTs of the electric field at chosen frequencies.

Mirrors ``wakis.field_monitors.FieldMonitor`` (field_monitors.py:14-118): same
constructor, attributes (frequencies, time_index, dt, Ex_acc, Ey_acc, Ez_acc,
shape, xp) and methods. The accumulators are complex128 device tensors of
shape (n_freq, Nx, Ny, Nz) (views of an x-fastest buffer); each ``update``
adds E * exp(-2j*pi*f*t) on the device (``wf_field_monitor_update``), with the
phase evaluated by numpy as the reference's scalar expression is (complex64 for
an FP32 dt, complex128 otherwise) and the product formed in numpy's promoted
type, t = time_index*dt.
Inside ``wakesolve`` the step call itself accumulates after every step from a
device table of those phases (``wf_step_args.mon_*``), so a monitored run needs
no per-step host work.
"""

from typing import Sequence

import numpy as np
import torch

from . import _lib


class FieldMonitor:
    def __init__(self, frequencies: Sequence[float]):
        self.frequencies = np.array(frequencies)
        self.time_index = 0
        self.dt = None
        self.Ex_acc = None
        self.Ey_acc = None
        self.Ez_acc = None
        self.shape = None
        self.xp = None
        self._acc = None

    def _start(self, E, dt):
        """First-update set-up (field_monitors.py:78-89): dt, shape and the
        complex128 accumulators."""
        if self.dt is None:
            self.dt = dt
            self.shape = (E.Nx, E.Ny, E.Nz)
            self.xp = E.xp
            nf = len(self.frequencies)
            # [3][nf][Nz][Ny][Nx] complex128; each component viewed as (nf, Nx, Ny, Nz)
            self._acc = torch.zeros((3, nf, E.Nz, E.Ny, E.Nx), dtype=torch.complex128,
                                    device=E.array.device)
            self.Ex_acc, self.Ey_acc, self.Ez_acc = (self._acc[c].permute(0, 3, 2, 1)
                                                     for c in range(3))

    def _phase_values(self, n):
        """exp(-2j*pi*f*t) of the next n updates as numpy evaluates the
        per-update expression (field_monitors.py:90-97), t = time_index*dt:
        (-2j*pi*f) is a Python complex, and its product with an FP32 t is
        complex64 (NEP 50), with an FP64 t complex128. Returns (n, n_freq)."""
        from .sources import step_times
        t = step_times(self.time_index + n, self.dt)[self.time_index:]
        t0 = self.time_index * self.dt
        cols = []
        for fr in self.frequencies:
            x = -2j * np.pi * fr
            rt = np.asarray(x * t0).dtype
            tr = t.astype(np.zeros(1, dtype=rt).real.dtype)  # t as the product converts it
            cols.append(np.exp(np.asarray(x, dtype=rt) * tr))
        return np.stack(cols, axis=1)

    def _phases(self, n, device):
        """The phases of the next n updates as an (n, n_freq, 2) device table
        of (re, im) in the phase's component type, and whether that is FP32."""
        ph = self._phase_values(n)
        f32 = ph.dtype == np.complex64
        tab = np.ascontiguousarray(np.stack([ph.real, ph.imag], axis=2))
        return torch.as_tensor(tab, device=device), f32

    def update(self, E, dt: float):
        """Accumulate one snapshot of the Field E (field_monitors.py:59-102)."""
        self._start(E, dt)
        ph, f32 = self._phases(1, E.array.device)
        code = _lib.WF_F64 if E.array.dtype == torch.float64 else _lib.WF_F32
        _lib.check(_lib.lib().wf_field_monitor_update(
            _lib.ptr(E.array), code, E.N, _lib.ptr(ph), int(f32), len(self.frequencies),
            _lib.ptr(self._acc), _lib.stream_ptr()), "wf_field_monitor_update")
        self.time_index += 1

    def _step_args(self, args, E, dt, nsteps):
        """Let one wf_step call of nsteps steps accumulate after every step
        (wf_step_args.mon_*), as nsteps update() calls would; returns the phase
        table, which the caller keeps until the call has been issued."""
        self._start(E, dt)
        ph, f32 = self._phases(nsteps, E.array.device)
        args.mon_acc, args.mon_phase = _lib.ptr(self._acc), _lib.ptr(ph)
        args.mon_row0, args.mon_nf = 0, len(self.frequencies)
        args.mon_phase_f32 = int(f32)
        return ph

    def get_components(self):
        """{'Ex', 'Ey', 'Ez'}: complex128 accumulators (n_freq, Nx, Ny, Nz)."""
        return {"Ex": self.Ex_acc, "Ey": self.Ey_acc, "Ez": self.Ez_acc}

