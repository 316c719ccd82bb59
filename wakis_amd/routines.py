"""Time loops ``emsolve`` and ``wakesolve`` (wakis/routines.py:17-482).

``wakesolve`` keeps the whole step loop on the device: the beam current is
uploaded once as an (Nt, Nz) table, the step kernel writes the beam line and
captures the Ez window after every E update, and the wake potential/impedance
are computed from that capture by device kernels. There is no per-step host
traffic unless a callback or field monitor asks for it.
"""

import time

import numpy as np
import torch

from . import _lib
from .field import torch_dtype


class RoutinesMixin:
    def emsolve(self, Nt, source=None, callback=None, save=False, fields=["E"],
                components=["Abs"], save_every=1, subdomain=None, plot=False, plot_every=1,
                use_etd=False, plot3d=False, **kwargs):
        """Time loop with host sources and the CPU-semantics ABC (routines.py:17-229)."""
        if save or plot or plot3d:
            raise NotImplementedError("emsolve: HDF5 saving and plotting are not part of "
                                      "this engine")
        self.Nt = Nt
        if source is not None:
            self.source = source
        low_abc = sum(1 << a for a in range(3) if self.bc_low[a].lower() == "abc")
        for n in range(Nt):
            if source is not None:
                source.update(self, n * self.dt)
            self.one_step()
            if self.activate_abc and low_abc:
                _lib.check(_lib.lib().wf_abc_low(self._handle, _lib.ptr(self.E.array),
                                                 _lib.ptr(self.H.array), low_abc,
                                                 _lib.stream_ptr()), "wf_abc_low")
            if callback is not None:
                callback(self, n * self.dt)

    def wakesolve(self, wakelength, wake=None, callback=None, compute_plane="both", plot=False,
                  plot_func=None, plot_from=None, plot_every=1, plot_until=None, save_J=False,
                  add_space=None, use_field_monitor=False, field_monitor=None, use_edt=None,
                  chunk=None, **kwargs):
        """Wakefield run and WakeSolver post-processing (routines.py:231-482)."""
        if wake is not None:
            self.wake = wake
        if add_space is not None and self.wake is not None:
            self.wake.skip_cells = add_space
        if plot or save_J:
            raise NotImplementedError("wakesolve: plotting and Jz.h5 saving are not part of "
                                      "this engine")
        ctx = self._wake_setup(wakelength, compute_plane)
        Nt = ctx["Nt"]
        print("Running electromagnetic time-domain simulation...")
        t0 = time.time()
        per_step = callback is not None or (use_field_monitor and field_monitor is not None)
        step = 1 if per_step else (Nt if chunk is None else int(chunk))
        n = 0
        while n < Nt:
            k = min(step, Nt - n)
            self._run(self._wake_args(ctx, n, k))
            self.nstep += k
            if per_step:
                if use_field_monitor and field_monitor is not None:
                    field_monitor.update(self.E, self.dt)
                if callback is not None:
                    callback(self, n * self.dt)
            n += k
        torch.cuda.synchronize(self.device)
        self.simulationTime = time.time() - t0
        self._wake_finish(ctx, compute_plane)

    def _wake_setup(self, wakelength, compute_plane):
        """Set-up part of wakesolve (routines.py:319-440): time axis, beam table,
        Ez capture buffer with sample 0, first-step ghost zeroing."""
        from .sources import Beam

        if self.wake is None:
            raise AttributeError("Wake solver information not passed to the solver instantiation")
        self.xtest, self.ytest = self.wake.xtest, self.wake.ytest
        self.ixt = np.abs(self.x - self.xtest).argmin()
        self.iyt = np.abs(self.y - self.ytest).argmin()
        if compute_plane.lower() == "longitudinal":
            ix0, iy0, nx, ny = self.ixt, self.iyt, 1, 1
        else:
            # the reference slices E[ixt-1:ixt+2, iyt-1:iyt+2] (routines.py:347-352):
            # with numpy's slice rules a window that leaves the grid is empty or
            # clipped; nothing is captured then and the wake post-processing fails
            xs = np.arange(self.Nx)[self.ixt - 1:self.ixt + 2]
            ys = np.arange(self.Ny)[self.iyt - 1:self.iyt + 2]
            ix0, iy0, nx, ny = self.ixt - 1, self.iyt - 1, 3, 3
            if len(xs) != 3 or len(ys) != 3:
                ix0, iy0, nx, ny = 0, 0, 0, 0

        self.wake.wakelength = wakelength
        self.ti = self.wake.ti
        self.v = self.wake.v
        z = self.z
        tmax = (wakelength + self.ti * self.v + (z.max() - z.min())) / self.v
        Nt = int(tmax / self.dt)
        self.tmax, self.Nt = tmax, Nt

        beam = Beam(q=self.wake.q, sigmaz=self.wake.sigmaz, beta=self.wake.beta,
                    xsource=self.wake.xsource, ysource=self.wake.ysource)
        dev, tdt = self.device, torch_dtype(self.dtype)
        table = torch.as_tensor(beam.table(self, Nt), device=dev).to(tdt).contiguous()
        self._wake_beam = beam
        cap = torch.zeros((nx * ny, max(Nt, 1), self.Nz), dtype=tdt, device=dev)
        if Nt > 0 and nx * ny > 0:  # sample 0: E before the first step (routines.py:437-440)
            _lib.check(_lib.lib().wf_capture(self._handle, _lib.ptr(self.E.array), _lib.ptr(cap),
                                             ix0, iy0, nx, ny, Nt, 0, _lib.stream_ptr()),
                       "wf_capture")
        if self.step_0:
            self._set_ghosts_to_0()
            self.step_0 = False
        return dict(Nt=Nt, beam=beam, table=table, cap=cap, win=(ix0, iy0, nx, ny),
                    t_axis=np.arange(0, Nt * self.dt, self.dt))

    def _wake_args(self, ctx, n, k):
        """wf_step arguments for steps n .. n+k-1 of the wakesolve run in ctx."""
        beam, (ix0, iy0, nx, ny), Nt = ctx["beam"], ctx["win"], ctx["Nt"]
        # the beam table rows are what Beam.update writes: J stays "beam line only"
        self._beam_line = (int(beam.ixs), int(beam.iys)) \
            if self.J.version == self._J_clean_version else self._beam_line
        a = self._step_args(k)
        a.ixs, a.iys = int(beam.ixs), int(beam.iys)
        if a.jmode == _lib.WF_J_NONE:
            a.jmode = _lib.WF_J_LINE
        a.beam_table, a.beam_row0, a.beam_rows = _lib.ptr(ctx["table"]), n, Nt
        a.capture = _lib.ptr(ctx["cap"]) if nx * ny > 0 else None
        a.cap_ix0, a.cap_iy0, a.cap_nx, a.cap_ny = ix0, iy0, nx, ny
        a.cap_samples, a.cap_sample0 = Nt, n + 1
        return a

    def _wake_finish(self, ctx, compute_plane):
        """Hand the capture to the WakeSolver and run its post-processing."""
        ix0, iy0, nx, ny = ctx["win"]
        self.wake._set_capture(cap=ctx["cap"], t=ctx["t_axis"], xf=self.x[ix0:ix0 + nx],
                               yf=self.y[iy0:iy0 + ny], zf=self.z, dx=self.grid.dx,
                               dy=self.grid.dy, dz=self.grid.dz, dtype=self.dtype,
                               nsamples=ctx["Nt"])
        self.wake.solve(compute_plane=compute_plane)
