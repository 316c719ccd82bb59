"""Time loops ``emsolve`` and ``wakesolve`` (wakis/routines.py:17-482).

``wakesolve`` keeps the whole step loop on the device: the beam current is
uploaded once as an (Nt, Nz) table, the step kernel writes the beam line and
captures the Ez window after every E update, and the wake potential/impedance
are computed from that capture by device kernels; a field monitor accumulates
inside the same step calls. There is no per-step host traffic unless a
callback asks for it.
"""

import os
import time

import numpy as np
import torch

from . import _lib
from .field import torch_dtype


class RoutinesMixin:
    def emsolve(self, Nt, source=None, callback=None, save=False, fields=["E"],
                components=["Abs"], save_every=1, subdomain=None, plot=False, plot_every=1,
                use_etd=False, plot3d=False, **kwargs):
        """Time loop with host sources and the CPU-semantics ABC (routines.py:17-229).

        ``save=True`` writes one HDF5 file per saved field, named field +
        component + '.h5' ('Ex.h5', 'EAbs.h5', ...; a name of two or more
        letters in ``fields`` carries its own component), with the axes x, y,
        z, dx, dy, dz, the time axis of the saved steps and, every
        ``save_every`` steps, the dataset '#nnnnn' of F[xx, yy, zz, component]
        over ``subdomain`` after the source update of step n (routines.py:
        132-156, 193-204). (The reference's own set-up loop iterates over the
        file names instead of the files and fails before the first step; this
        is the layout its loop body writes.)"""
        if plot or plot3d:
            raise NotImplementedError("emsolve: plotting is not part of this engine")
        self.Nt = Nt
        if source is not None:
            self.source = source
        hfs = self._emsolve_files(Nt, fields, components, save_every) if save else {}
        if subdomain is not None:
            xx, yy, zz = subdomain
        else:
            xx, yy, zz = slice(0, self.Nx), slice(0, self.Ny), slice(0, self.Nz)
        low_abc = sum(1 << a for a in range(3) if self.bc_low[a].lower() == "abc")
        try:
            self._emsolve_loop(Nt, source, callback, hfs, save_every, (xx, yy, zz), low_abc)
        finally:
            for hf in hfs.values():
                hf.close()

    def _emsolve_files(self, Nt, fields, components, save_every):
        from . import h5io
        names = []
        for field in fields:
            if len(field) == 1:
                names += [field + component for component in components]
            else:
                names.append(field)
        hfs = {}
        for name in names:
            if name[0] not in ("E", "H", "J") or name[1:] not in ("x", "y", "z", "Abs"):
                raise AttributeError(f"Component {name} not valid. Input must have a field "
                                     "[\"E\", \"H\", \"J\"] and a component [\"x\", \"y\", "
                                     "\"z\", \"Abs\"]")
            hf = h5io.File(name + ".h5", "w")
            hfs[name] = hf
            hf["x"], hf["y"], hf["z"] = self.x, self.y, self.z
            hf["dx"], hf["dy"], hf["dz"] = self.grid.dx, self.grid.dy, self.grid.dz
            hf["t"] = np.arange(0, Nt * self.dt, save_every * self.dt)
        return hfs

    def _emsolve_loop(self, Nt, source, callback, hfs, save_every, sub, low_abc):
        xx, yy, zz = sub
        for n in range(Nt):
            if source is not None:
                source.update(self, n * self.dt)
            if hfs and n % save_every == 0:
                for name, hf in hfs.items():
                    hf["#" + str(n).zfill(5)] = getattr(self, name[0])[xx, yy, zz, name[1:]]
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
                  chunk=None, checkpoint=None, checkpoint_every=None, resume=None, **kwargs):
        """Wakefield run and WakeSolver post-processing (routines.py:231-482).

        Beyond the reference: ``checkpoint_every=m`` with ``checkpoint=path``
        writes the run state (E, H, J, the step index and the Ez samples
        captured so far) to ``path`` every m steps, replacing the file
        atomically; ``resume=path`` continues such a run in a new solver built
        the same way, and the result is bitwise that of the uninterrupted run.
        The fields are checked for Inf/NaN on the device at the end
        (``fields_finite``). Runs of 2,000 steps or more on grids of 2^22
        cells or more first time the fused kernel's workgroup shapes
        (``tune_fused``); the fields are bitwise the same for every shape.
        """
        if wake is not None:
            self.wake = wake
        if add_space is not None and self.wake is not None:
            self.wake.skip_cells = add_space
        if plot:
            raise NotImplementedError("wakesolve: plotting is not part of this engine")
        every = None
        if checkpoint_every is not None:
            every = int(checkpoint_every)
            if every < 1 or checkpoint is None:
                raise ValueError("wakesolve: checkpoint_every needs a positive step count "
                                 "and a checkpoint file name")
        ctx = self._wake_setup(wakelength, compute_plane)
        Nt = ctx["Nt"]
        if Nt >= 2000 and self.N >= (1 << 22):  # long run: measure the kernel shapes first
            self.tune_fused(lambda k: self._wake_args(ctx, 0, k))
        if save_J:  # J_z on the window before the run, for Jz.h5
            ix0, iy0, nx, ny = ctx["win"]
            ctx["J0"] = self.J.to_matrix("z")[ix0:ix0 + nx, iy0:iy0 + ny, :].reshape(
                nx * ny, self.Nz).cpu().numpy()
        n = 0 if resume is None else self._wake_resume(ctx, resume)
        print("Running electromagnetic time-domain simulation...")
        t0 = time.time()
        monitor = field_monitor if (use_field_monitor and field_monitor is not None) else None
        per_step = callback is not None
        step = 1 if per_step else (Nt if chunk is None else int(chunk))
        while n < Nt:
            k = min(step, Nt - n)
            if every is not None:
                k = min(k, every - n % every)
            args = self._wake_args(ctx, n, k)
            if monitor is not None:  # accumulated after every step of the call
                phases = monitor._step_args(args, self.E, self.dt, k)
            self._run(args)
            self.nstep += k
            if monitor is not None:
                monitor.time_index += k
                del phases
            if callback is not None:
                callback(self, n * self.dt)
            n += k
            if every is not None and n % every == 0 and n < Nt:
                self._wake_checkpoint(ctx, checkpoint, n)
        torch.cuda.synchronize(self.device)
        self.simulationTime = time.time() - t0
        self.fields_finite = self.check_finite(("E", "H"))
        if not self.fields_finite:
            print("[!] Warning: the fields hold Inf/NaN values after the time stepping")
        if save_J:
            self._write_Jz_file(ctx, compute_plane)
        self._wake_finish(ctx, compute_plane)
        # run log (routines.py:478-482); wakis.log is written with the saved results
        self.logger.wakeSolver = self.wake.logger.wakeSolver
        self.logger.wakeSolver["wakelength"] = wakelength
        self.logger.wakeSolver["simulationTime"] = time.time() - t0
        if self.wake.save:
            self.logger.save_logs()

    def _wake_state_key(self, ctx):
        return dict(wake_Nt=np.int64(ctx["Nt"]), wake_dt=np.float64(self.dt),
                    wake_dims=np.asarray(self.dims, dtype=np.int64),
                    wake_dtype=np.str_(np.dtype(self.dtype).name),
                    wake_win=np.asarray(ctx["win"], dtype=np.int64))

    def _wake_checkpoint(self, ctx, path, n):
        """Write the state after n steps of the run in ctx: E, H, J (the keys of
        save_state, so load_state reads the file too), the step index, the Ez
        samples 0..n and the fast-path facts of J. Written to a temporary file
        and renamed, so an interrupted write keeps the previous checkpoint."""
        import os
        bl = self._beam_line if self._beam_line is not None else (-1, -1)
        state = dict(E=self.E.to_numpy(), H=self.H.to_numpy(), J=self.J.to_numpy(),
                     wake_n=np.int64(n), wake_cap=ctx["cap"][:, :n + 1].cpu().numpy(),
                     nstep=np.int64(self.nstep),
                     j_beam_only=np.bool_(self.J.version == self._J_clean_version),
                     j_sigma_e=np.bool_(self.use_conductivity and self._sigmaE_holds()),
                     beam_line=np.asarray(bl, dtype=np.int64), **self._wake_state_key(ctx))
        if "J0" in ctx:
            state["wake_J0"] = ctx["J0"]
        tmp = f"{path}.part"
        with open(tmp, "wb") as f:
            np.savez(f, **state)
        os.replace(tmp, path)

    def _wake_resume(self, ctx, path):
        """Load a _wake_checkpoint file into this run; returns the next step."""
        with np.load(path, allow_pickle=False) as st:
            want = self._wake_state_key(ctx)
            for key, val in want.items():
                if key not in st or not np.array_equal(st[key], val):
                    raise ValueError(f"wakesolve: {path} was written by a different run "
                                     f"({key}: {st[key] if key in st else None} != {val})")
            n = int(st["wake_n"])
            if not 0 < n <= ctx["Nt"]:
                raise ValueError(f"wakesolve: {path} holds step {n} of {ctx['Nt']}")
            for name in ("E", "H", "J"):
                getattr(self, name).fromarray(st[name])
            cap = st["wake_cap"]
            if cap.shape != tuple(ctx["cap"][:, :n + 1].shape):
                raise ValueError(f"wakesolve: {path} holds an Ez capture of shape {cap.shape}")
            ctx["cap"][:, :n + 1].copy_(torch.as_tensor(cap))
            self.nstep = int(st["nstep"])
            if bool(st["j_beam_only"]):  # J is zero off the beam line, as in the saved run
                self._J_clean_version = self.J.version
                bl = tuple(int(v) for v in st["beam_line"])
                self._beam_line = None if bl[0] < 0 else bl
            if bool(st["j_sigma_e"]):
                self._sigmaE_versions = (self.E.version, self.J.version)
            if "J0" in ctx:
                if "wake_J0" not in st:
                    raise ValueError(f"wakesolve: {path} holds no Jz window for save_J")
                ctx["J0"] = st["wake_J0"]
        return n

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

    def _write_Ez_file(self, ctx, compute_plane):
        """Ez.h5 of the run (routines.py:376-389, 437-440): the test window's
        axes x, y, z (scalars x[ixt], y[iyt] for a longitudinal run), dx, dy,
        dz, the time axis t, and for every step n < Nt the dataset '#%05d' of
        Ez on the window before step n (shape (Nz,) or (3, 3, Nz)) in the
        field dtype, written from the device capture after the run."""
        from . import h5io
        ix0, iy0, nx, ny = ctx["win"]
        longit = compute_plane.lower() == "longitudinal"
        path = self.wake.Ez_file
        folder = os.path.dirname(path)
        if folder:
            os.makedirs(folder, exist_ok=True)
        with h5io.File(path, "w") as hf:
            if longit:
                hf["x"], hf["y"] = self.x[self.ixt], self.y[self.iyt]
            else:
                hf["x"] = self.x[self.ixt - 1:self.ixt + 2]
                hf["y"] = self.y[self.iyt - 1:self.iyt + 2]
            hf["z"] = self.z
            hf["dx"], hf["dy"], hf["dz"] = self.grid.dx, self.grid.dy, self.grid.dz
            hf["t"] = ctx["t_axis"]
            if nx * ny == 0:  # the window leaves the grid: nothing was captured
                print("[!] Warning: the Ez test window leaves the grid; Ez.h5 holds no samples")
                return
            cap = ctx["cap"].cpu().numpy()  # [lines][Nt][Nz], lines x-major
            for n in range(ctx["Nt"]):
                a = cap[:, n, :]
                hf["#" + str(n).zfill(5)] = a[0] if longit else a.reshape(nx, ny, self.Nz)

    def _write_Jz_file(self, ctx, compute_plane):
        """Jz.h5 (routines.py:391-401, 441-442): like Ez.h5, with J_z on the
        window before each step n. That J is the beam's current of step n on
        the beam line (Beam.update runs before the save) and, elsewhere, what
        the previous step left: sigma * E (the step's J update) when the run is
        conductive, else J as it was before the run; sample 0 is the J before
        the run. It is formed here from the Ez capture, the beam table and
        the sigma array, with the step's single rounding of sigma * E."""
        from . import h5io
        ix0, iy0, nx, ny = ctx["win"]
        longit = compute_plane.lower() == "longitudinal"
        with h5io.File("Jz.h5", "w") as hf:
            if longit:
                hf["x"], hf["y"] = self.x[self.ixt], self.y[self.iyt]
            else:
                hf["x"] = self.x[self.ixt - 1:self.ixt + 2]
                hf["y"] = self.y[self.iyt - 1:self.iyt + 2]
            hf["z"] = self.z
            hf["dx"], hf["dy"], hf["dz"] = self.grid.dx, self.grid.dy, self.grid.dz
            hf["t"] = ctx["t_axis"]
            if nx * ny == 0:
                return
            cap = ctx["cap"].cpu().numpy()
            J0 = ctx["J0"]
            sig = self.sigma.to_matrix("z")[ix0:ix0 + nx, iy0:iy0 + ny, :].reshape(
                nx * ny, self.Nz).cpu().numpy().astype(cap.dtype)
            beam = ctx["beam"]
            bi, bj = int(beam.ixs) - ix0, int(beam.iys) - iy0
            bline = bi * ny + bj if (0 <= bi < nx and 0 <= bj < ny) else None
            table = ctx["table"].cpu().numpy() if bline is not None else None
            for n in range(ctx["Nt"]):
                a = (sig * cap[:, n, :]) if (n > 0 and self.use_conductivity) else J0.copy()
                if bline is not None:
                    a[bline] = table[n]
                hf["#" + str(n).zfill(5)] = a[0] if longit else a.reshape(nx, ny, self.Nz)

    def _wake_finish(self, ctx, compute_plane):
        """Hand the capture to the WakeSolver (and to Ez.h5 when the wake
        results are saved or an Ez_file was named), then post-process."""
        if self.wake.save or getattr(self.wake, "_Ez_file_given", False):
            self._write_Ez_file(ctx, compute_plane)
        ix0, iy0, nx, ny = ctx["win"]
        self.wake._set_capture(cap=ctx["cap"], t=ctx["t_axis"], xf=self.x[ix0:ix0 + nx],
                               yf=self.y[iy0:iy0 + ny], zf=self.z, dx=self.grid.dx,
                               dy=self.grid.dy, dz=self.grid.dz, dtype=self.dtype,
                               nsamples=ctx["Nt"])
        self.wake.solve(compute_plane=compute_plane)
