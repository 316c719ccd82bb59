"""FieldMonitor: running DFTs of the electric field at chosen frequencies.

Mirrors ``wakis.field_monitors.FieldMonitor`` (field_monitors.py:14-118): same
constructor, attributes (frequencies, time_index, dt, Ex_acc, Ey_acc, Ez_acc,
shape, xp) and methods. The accumulators are complex128 device tensors of
shape (n_freq, Nx, Ny, Nz) (views of an x-fastest buffer); each ``update``
adds E * exp(-2j*pi*f*t) on the device (``wf_field_monitor_update``), with the
phase evaluated by numpy in complex128 as the reference does, t = time_index*dt.
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

    def update(self, E, dt: float):
        """Accumulate one snapshot of the Field E (field_monitors.py:59-102)."""
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
        t = self.time_index * self.dt
        phase = np.exp(-2j * np.pi * self.frequencies * t).astype(np.complex128)
        ph = torch.as_tensor(np.ascontiguousarray(np.stack([phase.real, phase.imag], axis=1)),
                             device=E.array.device)
        code = _lib.WF_F64 if E.array.dtype == torch.float64 else _lib.WF_F32
        _lib.check(_lib.lib().wf_field_monitor_update(
            _lib.ptr(E.array), code, E.N, _lib.ptr(ph), len(self.frequencies),
            _lib.ptr(self._acc), _lib.stream_ptr()), "wf_field_monitor_update")
        self.time_index += 1

    def get_components(self):
        """{'Ex', 'Ey', 'Ez'}: complex128 accumulators (n_freq, Nx, Ny, Nz)."""
        return {"Ex": self.Ex_acc, "Ey": self.Ey_acc, "Ez": self.Ez_acc}

