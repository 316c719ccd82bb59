"""WakeSolver: wake potential and impedance from the device Ez capture.

Mirrors ``wakis.wakeSolver.WakeSolver`` (wakeSolver.py:22-254, 310-870):
same constructor, same ``solve`` dispatch and result attributes (s, WP, WP_3d,
WPx, WPy, f, Z, Zx, Zy, lambdas, lambdaf). The Ez samples come from the
device capture written during ``wakesolve`` instead of ``Ez.h5``; the line
integrals (``wf_wake_wp``), the transverse prefix sums (``wf_wake_trans``)
and the DFT bins (``wf_dft``) run as HIP kernels in FP64.
"""

import ctypes
import os
import time

import numpy as np
import torch
from scipy.constants import c as c_light

from . import _lib
from .logger import Logger


class WakeSolver:
    def __init__(self, wakelength=None, q=1e-9, sigmaz=1e-3, beta=1.0, xsource=0.0,
                 ysource=0.0, xtest=0.0, ytest=0.0, chargedist=None, ti=None, fmax=None,
                 compute_plane="both", skip_cells=0, add_space=None, Ez_file=None, save=True,
                 results_folder="results/", verbose=1, counter_moving=False):
        self.verbose = verbose
        self.q, self.sigmaz, self.beta = q, sigmaz, beta
        self.v = self.beta * c_light
        self.xsource, self.ysource = xsource, ysource
        self.xtest, self.ytest = xtest, ytest
        self.chargedist = chargedist
        self.fmax = self.v / (3.0 * self.sigmaz) if fmax is None else fmax
        self.skip_cells = skip_cells if add_space is None else add_space
        self.compute_plane = compute_plane
        self.counter_moving = counter_moving
        if ti is not None:
            self.ti = ti
        else:  # wakeSolver.py:164-172 (CST convention for beta <= 1)
            self.ti = 8.548921333333334 * self.sigmaz / (np.sqrt(self.beta) * self.v)
        # Ez.h5 is written by wakesolve when results are saved or a file name is
        # given; the default name is the reference's (wakeSolver.py:202-203)
        self._Ez_file_given = Ez_file is not None
        self.Ez_file = Ez_file
        self.Ez_hf = None
        self.Ezt = None  # Ez(x_test, y_test, z, t) given to calc_long_WP
        self.t = None
        self.xf = self.yf = self.zf = None  # field subdomain
        self.x = self.y = self.z = None  # full simulation domain
        self.wakelength = wakelength
        self.s = self.lambdas = self.WP = self.WP_3d = None
        self.n_transverse_cells = 1
        self.WPx = self.WPy = self.f = self.Z = self.Zx = self.Zy = self.lambdaf = None
        self.loss_factor = None
        self.save = save
        self.folder = results_folder if results_folder.endswith("/") else results_folder + "/"
        if self.save:
            os.makedirs(self.folder, exist_ok=True)
        if self.Ez_file is None:
            self.Ez_file = self.folder + "Ez.h5"
        self._cap = None
        self.logger = Logger()
        self.assign_logs()

    def assign_logs(self):
        """The wake parameters into ``logger.wakeSolver`` (wakeSolver.py:1709-1723)."""
        for key in ("ti", "q", "sigmaz", "beta", "xsource", "ysource", "xtest", "ytest",
                    "chargedist", "skip_cells"):
            self.logger.wakeSolver[key] = getattr(self, key)
        self.logger.wakeSolver["results_folder"] = self.folder

    def copy(self):
        """A shallow copy: a new WakeSolver sharing this one's attribute values
        (wakeSolver.py:1509-1520)."""
        obj = type(self).__new__(type(self))
        obj.__dict__.update(self.__dict__)
        return obj

    def log(self, msg):
        if self.verbose:
            print(msg)

    # capture handed over by SolverFIT3D.wakesolve
    def _set_capture(self, cap, t, xf, yf, zf, dx, dy, dz, dtype, nsamples):
        self._cap = cap
        self._cap_from_file = False
        self._cap_dtype = np.dtype(dtype)
        self._cap_samples = nsamples
        self.t, self.xf, self.yf, self.zf = t, xf, yf, zf
        self.dx, self.dy, self.dz = dx, dy, dz

    def solve(self, compute_plane=None, **kwargs):
        if compute_plane is None:
            compute_plane = self.compute_plane
        for key, val in kwargs.items():
            setattr(self, key, val)
        t0 = time.time()
        if compute_plane.lower() in ["both", "transverse"]:
            self.calc_long_WP_3d()
            self.calc_trans_WP()
            self.calc_long_Z()
            self.calc_trans_Z()
            self.calc_loss_factor()
        elif compute_plane.lower() == "longitudinal":
            self.calc_long_WP()
            self.calc_long_Z()
            self.calc_loss_factor()
        self.log("Calculation terminated in %ds" % (time.time() - t0))

    # ---------------------------------------------------------------- WP
    def _s_axis(self):
        nt = len(self.t)
        dt = self.t[2] - self.t[1]
        zmax, zmin = np.max(self.zf), np.min(self.zf)
        if self.wakelength is not None:
            wakelength = self.wakelength
        else:
            wakelength = nt * dt * self.v - (zmax - zmin) - self.ti * self.v
            self.wakelength = wakelength
        return np.arange(-self.ti * self.v, wakelength, dt * self.v), dt, zmin

    def _zz(self):
        """z rows kept by skip_cells: (first row, number of rows)."""
        z = self.zf[slice(self.skip_cells, -self.skip_cells) if self.skip_cells != 0
                    else np.s_[:]]
        return (self.skip_cells if self.skip_cells != 0 else 0), len(z)

    def _device_wp(self, cap, nlines, cap_samples, nt, ncols, counter_moving=False, row0=None):
        """wf_wake_wp over `nlines` lines of the device array `cap`
        ([nlines][cap_samples][Nz]) whose z index row0 holds the first z row kept
        by skip_cells (default: the skip_cells offset of a full-z capture);
        returns s and WP3d [nlines][len(s)] (FP64, on the device)."""
        s, dt, zmin = self._s_axis()
        k0, nz = self._zz()
        z = np.ascontiguousarray(self.zf[k0:k0 + nz], dtype=np.float64)
        if row0 is not None:
            k0 = row0
        Nz = cap.shape[-1]
        dev = cap.device
        zt = torch.as_tensor(z, device=dev)
        dzt = torch.as_tensor(np.ascontiguousarray(self.dz[:max(nz, 1)], dtype=np.float64),
                              device=dev)
        st_ = torch.as_tensor(np.ascontiguousarray(s, dtype=np.float64), device=dev)
        WP3d = torch.zeros((nlines, len(s)), dtype=torch.float64, device=dev)
        if len(s) == 0:  # wakelength below -ti*v: an empty s axis, nothing to integrate
            return s, WP3d
        if counter_moving:  # wakeSolver.py:505-511: zmax/(-1*v)
            zref_v = float(np.max(self.zf) / (-1 * self.v))
        else:  # wakeSolver.py:524-528: zmin/v
            zref_v = float(zmin / self.v)
        dtype_code = _lib.WF_F64 if cap.dtype == torch.float64 else _lib.WF_F32
        _lib.check(_lib.lib().wf_wake_wp(
            _lib.ptr(cap), dtype_code, nlines, cap_samples, Nz, k0, nz,
            _lib.ptr(zt), _lib.ptr(dzt), _lib.ptr(st_), len(s), float(self.v), zref_v,
            float(self.t[0]), float(self.ti), float(dt), nt, ncols, int(bool(counter_moving)),
            float(self.q * 1e12), _lib.ptr(WP3d), _lib.stream_ptr()), "wf_wake_wp")
        return s, WP3d

    # ------------------------------------------------------------- Ez.h5
    def read_Ez(self, filename=None, return_value=False):
        """Open an Ez.h5 file and take its axes (wakeSolver.py:1313-1360):
        xf, yf, zf, dx, dy, dz and t from the datasets of those names. With
        ``return_value`` the open file is returned (and kept as ``Ez_hf``);
        otherwise it is closed again."""
        from . import h5io
        if filename is None:
            filename = self.Ez_file
        hf = h5io.File(filename, "r")
        self.log(f"Reading h5 file {filename}")
        self.log("Size of the h5 file: " + str(round((os.path.getsize(filename) / 10**9), 2))
                 + " Gb")
        self.Ez_hf = hf
        self.Ez_file = filename
        keys = set(hf.keys())
        for key, attr in (("x", "xf"), ("y", "yf"), ("z", "zf"), ("dx", "dx"), ("dy", "dy"),
                          ("dz", "dz"), ("t", "t")):
            if key in keys:
                setattr(self, attr, np.array(hf[key]))
        if return_value:
            return hf
        hf.close()
        self.Ez_hf = None

    def _capture_from_file(self):
        """The Ez samples of an Ez.h5 file as a device capture [lines][samples][Nz].

        The samples are the datasets whose names start with '#', in sorted name
        order (wakeSolver.py:486-491). A 3-D sample (x, y, z) gives the
        (2n+1)^2 lines around its centre [shape//2 + i, shape//2 + j]
        (:496-497); a 1-D sample is one line (:384-393)."""
        self.read_Ez(return_value=True)
        hf = self.Ez_hf
        try:
            keys = sorted(k for k in hf.keys() if isinstance(k, str) and k.startswith("#"))
            if not keys:
                raise IndexError(f"no Ez samples ('#' datasets) in {self.Ez_file}")
            first = hf[keys[0]]
            i0 = self.n_transverse_cells
            if first.ndim == 3:
                cx, cy = first.shape[0] // 2, first.shape[1] // 2
                if (cx - i0 < 0 or cy - i0 < 0 or cx + i0 >= first.shape[0]
                        or cy + i0 >= first.shape[1]):
                    raise IndexError(f"Ez samples of shape {first.shape} hold no "
                                     f"{2 * i0 + 1}x{2 * i0 + 1} window")
                sel = (slice(cx - i0, cx + i0 + 1), slice(cy - i0, cy + i0 + 1))
                lines = (2 * i0 + 1) ** 2
            elif first.ndim == 1:
                sel, lines = None, 1
            else:
                raise IndexError(f"Ez samples of shape {first.shape}: 1-D or 3-D expected")
            nz = first.shape[-1]
            host = np.empty((lines, len(keys), nz), dtype=first.dtype)
            for n, key in enumerate(keys):
                a = hf[key][:]
                host[:, n, :] = a.reshape(1, nz) if sel is None else a[sel].reshape(lines, nz)
        finally:
            hf.close()
            self.Ez_hf = None
        if self.zf is None:
            self.zf = self.z
        self._cap = torch.as_tensor(host, device=torch.device("cuda"))
        self._cap_dtype = host.dtype
        self._cap_samples = len(keys)
        self._cap_from_file = True
        return self._cap

    def _captured(self):
        if self._cap is None:
            if self.Ez_file is not None and os.path.exists(self.Ez_file):
                return self._capture_from_file()
            raise RuntimeError("no Ez samples: run SolverFIT3D.wakesolve first, give an Ez.h5 "
                               f"file (Ez_file, not found: {self.Ez_file}), or pass Ezt to "
                               "calc_long_WP")
        if self._cap.shape[0] < 1 or self._cap_samples < 1:
            raise IndexError("no Ez captured on the test window (it leaves the grid)")
        return self._cap

    def calc_long_WP_3d(self, **kwargs):
        """wakeSolver.py:422-554 (3x3 lines, carry between lines; the
        counter-moving integral of :503-517 when counter_moving)."""
        for key, val in kwargs.items():
            setattr(self, key, val)
        i0 = self.n_transverse_cells
        nl = (2 * i0 + 1) ** 2
        cap = self._captured()
        if cap.shape[0] < nl:
            raise IndexError("no Ez captured on the test window (it leaves the grid)")
        # Ezt has len(t) columns (allocated before nt = min(nt, len(keys)), :487-491)
        nt = min(len(self.t), self._cap_samples)
        s, WP3d = self._device_wp(cap, nl, cap.shape[1], nt, len(self.t),
                                  counter_moving=self.counter_moving)
        self._WP3d_dev = WP3d
        self.s = s
        self.WP_3d = WP3d.cpu().numpy().reshape(2 * i0 + 1, 2 * i0 + 1, len(s))
        self.WP = self.WP_3d[i0, i0, :].copy()
        self._save_txt("WP.txt", self.s, self.WP, "WP [V/pC]")

    def calc_long_WP(self, Ezt=None, **kwargs):
        """wakeSolver.py:310-420: the wake potential of one line of Ez samples.

        Without ``Ezt`` the line comes from the device capture: a single
        captured line (compute_plane='longitudinal'), or, on a 3x3 window, the
        line [shape//2 + 1, shape//2 + 1] that the reference reads there
        (wakeSolver.py:385-388: the corner line at (+1, +1), not the centre).
        ``Ezt`` is an (nz, nt) array of Ez(x_test, y_test, z, t) samples, nz the
        number of z rows kept by skip_cells, integrated over min(len(t), nt)
        samples (the documented input, wakeSolver.py:326-327). The reference
        itself reads the samples from Ez.h5 even when Ezt is given (an open
        file overwrites Ezt, :375-393; without one it fails at :376)."""
        for key, val in kwargs.items():
            setattr(self, key, val)
        if self.zf is None:
            self.zf = self.z
        if Ezt is not None:
            Ezt = np.asarray(Ezt, dtype=np.float64)
            k0, nz = self._zz()
            if Ezt.ndim != 2 or Ezt.shape[0] != nz or Ezt.shape[1] < 1:
                raise ValueError(f"calc_long_WP: Ezt must have shape (nz, nt) with nz={nz} "
                                 f"z rows, got {Ezt.shape}")
            self.Ezt = Ezt
            # as a one-line capture [1][ncols][nz] whose first z row is row 0
            cap = torch.as_tensor(np.ascontiguousarray(Ezt.T)[None], device=torch.device("cuda"))
            ncols = Ezt.shape[1]
            s, WP3d = self._device_wp(cap, 1, ncols, min(len(self.t), ncols), ncols, row0=0)
            WP = WP3d[0]
        else:
            cap = self._captured()
            nl = cap.shape[0]
            if nl == 1:
                if self.skip_cells != 0:  # :390-392 assigns the full line: shape error
                    raise ValueError("could not broadcast the captured line into Ezt with "
                                     "skip_cells")
                line = 0
            else:
                side = int(round(np.sqrt(nl)))
                line = (side // 2 + 1) * side + (side // 2 + 1)
            # Ezt has min(len(t), len(keys)) columns here (:381-382)
            nt = min(len(self.t), self._cap_samples)
            s, WP3d = self._device_wp(cap[line:line + 1], 1, cap.shape[1], nt, nt)
            WP = WP3d[0]
        self.s = s
        self.WP = WP.cpu().numpy()
        self._save_txt("WP.txt", self.s, self.WP, "WP [V/pC]")

    def calc_trans_WP(self, **kwargs):
        """wakeSolver.py:556-629 on the device (numpy-exact prefix sums)."""
        for key, val in kwargs.items():
            setattr(self, key, val)
        ds = self.s[2] - self.s[1]
        dxf, dyf = np.diff(self.xf), np.diff(self.yf)
        i0 = self.n_transverse_cells
        ns = len(self.s)
        WP3d = getattr(self, "_WP3d_dev", None)
        if WP3d is None:
            WP3d = torch.as_tensor(np.ascontiguousarray(self.WP_3d.reshape(-1, ns)),
                                   device=torch.device("cuda"))
        WPx = torch.zeros(ns, dtype=torch.float64, device=WP3d.device)
        WPy = torch.zeros_like(WPx)
        _lib.check(_lib.lib().wf_wake_trans(
            _lib.ptr(WP3d), ns, float(ds), float(dxf[i0 - 1] + dxf[i0]),
            float(dyf[i0 - 1] + dyf[i0]), _lib.ptr(WPx), _lib.ptr(WPy), _lib.stream_ptr()),
            "wf_wake_trans")
        self.WPx, self.WPy = WPx.cpu().numpy(), WPy.cpu().numpy()
        self._save_txt("WPx.txt", self.s, self.WPx, "WP [V/pC]")
        self._save_txt("WPy.txt", self.s, self.WPy, "WP [V/pC]")

    # ----------------------------------------------------------------- Z
    def calc_lambdas_analytic(self, **kwargs):
        """wakeSolver.py:838-870."""
        for key, val in kwargs.items():
            setattr(self, key, val)
        self.lambdas = (1 / (self.sigmaz * np.sqrt(2 * np.pi))
                        * np.exp(-(self.s ** 2) / (2 * self.sigmaz ** 2)))
        self._save_txt("lambda.txt", self.s, self.lambdas, "Charge distribution [C/m]",
                       x_label="s [Hz]")

    def calc_lambdas(self, **kwargs):
        """wakeSolver.py:783-836 (interpolated user charge distribution)."""
        for key, val in kwargs.items():
            setattr(self, key, val)
        if isinstance(self.chargedist, dict):
            keys = list(self.chargedist.keys())
            z, chargedist = self.chargedist[keys[0]], self.chargedist[keys[1]]
        else:
            chargedist = self.chargedist
            z = self.zf if self.zf is not None and len(self.zf) == len(chargedist) else None
            if z is None:
                raise ValueError("Dimension error: check input dimensions")
        self.lambdas = np.interp(self.s, z, chargedist / self.q)
        self._save_txt("lambda.txt", self.s, self.lambdas, "Charge distribution [C/m]",
                       x_label="s [Hz]")

    def _dft_bins(self, signals, samples, fmax):
        ds = np.mean(self.s[1:] - self.s[:-1])
        N = int((self.v / ds) // fmax * samples)
        ffft = np.fft.fftfreq(N, ds / self.v)
        mask = np.logical_and(ffft >= 0, ffft < fmax)
        bins = np.nonzero(mask)[0]
        nb = len(bins)
        if nb and not np.array_equal(bins, np.arange(nb)):
            raise RuntimeError("unexpected DFT bin layout")
        L = min(len(self.s), N)
        x = np.ascontiguousarray(np.stack([np.asarray(sig, dtype=np.float64)[:L]
                                           for sig in signals]))
        xt = torch.as_tensor(x, device=torch.device("cuda"))
        out = torch.zeros((len(signals), max(nb, 1), 2), dtype=torch.float64, device=xt.device)
        _lib.check(_lib.lib().wf_dft(_lib.ptr(xt), len(signals), L, N, nb, _lib.ptr(out),
                                     _lib.stream_ptr()), "wf_dft")
        o = out.cpu().numpy()[:, :nb]
        return ffft[mask], ds, o[..., 0] + 1j * o[..., 1]

    def calc_long_Z(self, samples=1001, fmax=None, **kwargs):
        """wakeSolver.py:631-719 with the kept DFT bins evaluated on the device."""
        for key, val in kwargs.items():
            setattr(self, key, val)
        if self.lambdas is None and self.chargedist is not None:
            self.calc_lambdas()
        elif self.lambdas is None or len(self.lambdas) != len(self.s):
            self.calc_lambdas_analytic()
        fmax = self.fmax if fmax is None else fmax
        f, ds, X = self._dft_bins([self.lambdas * self.v, self.WP * 1e12], samples, fmax)
        lambdaf = X[0] * ds
        WPf = X[1] * ds
        self.f = f
        self.Z = -WPf / lambdaf
        self.lambdaf = lambdaf
        self._save_txt("Z.txt", self.f, self.Z, "Z [Ohm]", x_label="f [Hz]")
        self._save_txt("spectrum.txt", self.f, self.lambdaf,
                       "Charge distribution spectrum [C/s]", x_label="f [Hz]")

    def calc_trans_Z(self, samples=1001, fmax=None):
        """wakeSolver.py:721-781."""
        fmax = self.fmax if fmax is None else fmax
        f, ds, X = self._dft_bins([self.lambdas * self.v, self.WPx * 1e12, self.WPy * 1e12],
                                  samples, fmax)
        lambdaf = X[0] * ds
        self.Zx = 1j * (X[1] * ds) / lambdaf
        self.Zy = 1j * (X[2] * ds) / lambdaf
        self.fx = self.fy = f
        self._save_txt("Zx.txt", self.fx, self.Zx, "Zx [Ohm]", x_label="f [Hz]")
        self._save_txt("Zy.txt", self.fy, self.Zy, "Zy [Ohm]", x_label="f [Hz]")

    def calc_loss_factor(self):
        """Loss factor k = -integral WP(s) lambda(s) ds in V/pC (no reference
        counterpart): a compensated FP64 sum over the s axis on the device,
        ds = s[2] - s[1] as for the transverse integral."""
        if self.lambdas is None or len(self.lambdas) != len(self.s):
            self.calc_lambdas_analytic()
        dev = torch.device("cuda")
        wp = torch.as_tensor(np.ascontiguousarray(self.WP, dtype=np.float64), device=dev)
        lam = torch.as_tensor(np.ascontiguousarray(self.lambdas, dtype=np.float64), device=dev)
        out = torch.zeros(1, dtype=torch.float64, device=dev)
        _lib.check(_lib.lib().wf_loss_factor(_lib.ptr(wp), _lib.ptr(lam), len(self.s),
                                             float(self.s[2] - self.s[1]), _lib.ptr(out),
                                             _lib.stream_ptr()), "wf_loss_factor")
        self.loss_factor = float(out.item())
        return self.loss_factor

    def _save_txt(self, name, a, b, label, x_label="s [m]"):
        """The result files of wakeSolver.py:411-416, 705-718, 771-781, 827-835
        (same names, columns and two header lines), written when save=True."""
        if self.save:
            np.savetxt(self.folder + name, np.c_[a, b],
                       header="   " + x_label + " " * 20 + label + "\n" + "-" * 48)

    def save_txt(self, f_name, x_data=None, y_data=None, x_name="X [-]", y_name="Y [-]"):
        """Two-column text file with a two-line header (wakeSolver.py:1418-1471)."""
        if f_name.endswith(".txt"):
            f_name = f_name[:-4]
        if x_data is not None and y_data is not None:
            np.savetxt(f_name + ".txt", np.c_[x_data, y_data],
                       header="   " + x_name + " " * 20 + y_name + "\n" + "-" * 48)
        else:
            self.log("[!] txt not saved, please provide x_data and y_data")

    def read_txt(self, txt, skiprows=2, delimiter=None, usecols=None):
        """Columns of a text file as a dict keyed by the header names ("s[m]",
        ...; spaces and '#' dropped, split after each ']'), or by column index
        when the header does not name them; complex columns are read as
        complex128 (wakeSolver.py:1362-1416)."""
        try:
            load = np.loadtxt(txt, skiprows=skiprows, delimiter=delimiter, usecols=usecols)
        except ValueError:
            self.log(f"[!] Using dtype=np.complex128 to read {txt}")
            load = np.loadtxt(txt, skiprows=skiprows, delimiter=delimiter, usecols=usecols,
                              dtype=np.complex128)
        load = np.atleast_2d(load)
        try:
            with open(txt) as f:
                header = f.readline()
            names = header.replace(" ", "").replace("#", "").replace("\n", "").split("]")
            d = {names[i] + "]": load[:, i] for i in range(load.shape[1])}
        except IndexError:
            self.log("[!] Using integer keys since no header was found")
            d = {i: load[:, i] for i in range(load.shape[1])}
        return d

    def load_results(self, folder=None):
        """Read the result files of a run saved with save=True (wakeSolver.py:1473-1507):
        lambdas, WPx, WPy, s, WP, lambdaf, Zx, Zy, f, Z."""
        if folder is None and self.folder is None:
            raise Exception("[!] Please provide a folder path")
        elif folder is None:
            folder = self.folder
        if not folder.endswith("/"):
            folder = folder + "/"
        _, self.lambdas = self.read_txt(folder + "lambda.txt").values()
        _, self.WPx = self.read_txt(folder + "WPx.txt").values()
        _, self.WPy = self.read_txt(folder + "WPy.txt").values()
        self.s, self.WP = self.read_txt(folder + "WP.txt").values()
        _, self.lambdaf = self.read_txt(folder + "spectrum.txt").values()
        _, self.Zx = self.read_txt(folder + "Zx.txt").values()
        _, self.Zy = self.read_txt(folder + "Zy.txt").values()
        self.f, self.Z = self.read_txt(folder + "Z.txt").values()
        self.f = np.abs(self.f)
        self.wakelength = self.s[-1]
        self.folder = folder
        self.Ez_file = folder + "Ez.h5"
