"""Time-dependent sources (wakis/sources.py).

``Beam`` (sources.py:22-110) writes the Gaussian beam current on the z line
through the source. ``Beam.table`` evaluates the same expression for every
step of a run at once (same per-element IEEE operations as the per-step
update, including NumPy's scalar promotion of ``n*dt``), so ``wakesolve`` can
upload it once and let the step kernel read one row per step.
``PlaneWave`` (sources.py:142-282), ``WavePacket`` (:285-455), ``Dipole``
(:458-566) and ``Pulse`` (:569-762) write E, H or J between device steps
through the Field indexing (device-side fills; the wave packet's transverse
Gaussian is kept on the device in FP64 and scaled there with the same
rounding order as the reference's NumPy expression).
"""

import numpy as np
import torch
from scipy.constants import c as c_light
from scipy.constants import mu_0


def step_times(nsteps, dt):
    """``n * dt`` for n = 0..nsteps-1 with the scalar promotion of ``n * dt``.

    A Python int times a NumPy floating scalar keeps the scalar's precision;
    a Python int times a Python float is a double.
    """
    n = np.arange(nsteps)
    if isinstance(dt, np.floating):
        return n.astype(type(dt)) * dt
    return n.astype(np.float64) * float(dt)


class Beam:
    def __init__(self, xsource=0.0, ysource=0.0, beta=1.0, q=1e-9, sigmaz=None, ti=None):
        self.xsource, self.ysource = xsource, ysource
        self.sigmaz = sigmaz
        self.q = q
        self.beta = beta
        self.v = c_light * beta
        if ti is not None:
            self.ti = ti
        else:
            self.ti = 8.548921333333334 * self.sigmaz / self.v
        self.is_first_update = True

    def _locate(self, solver):
        if self.is_first_update:
            self.ixs = np.abs(solver.x - self.xsource).argmin()
            self.iys = np.abs(solver.y - self.ysource).argmin()
            self.is_first_update = False
            self.zmin = solver.z.min()

    def _line(self, solver, t):
        # t: scalar or (nt,1) column of times
        s0 = self.zmin - self.v * self.ti
        s = solver.z - self.v * t
        profile = (1 / np.sqrt(2 * np.pi * self.sigmaz ** 2)
                   * np.exp(-((s - s0) ** 2) / (2 * self.sigmaz ** 2)))
        return self.q * self.v * profile / solver.dx[self.ixs] / solver.dy[self.iys]

    def update(self, solver, t):
        """J[ixs, iys, :, 'z'] at time t (sources.py:76-110)."""
        self._locate(solver)
        solver._write_beam_line(self.ixs, self.iys, self._line(solver, t))

    def table(self, solver, nsteps):
        """Rows n = 0..nsteps-1 of the line written by ``update(solver, n*dt)``."""
        self._locate(solver)
        t = step_times(nsteps, solver.dt)
        return self._line(solver, t[:, None]).astype(solver.dtype)


class PlaneWave:
    """z-travelling plane wave written on a z plane (sources.py:142-282)."""

    def __init__(self, xs=None, ys=None, zs=0, nodes=None, f=None, amplitude=1.0, beta=1.0,
                 phase=0):
        self.nodes, self.beta = nodes, beta
        self.xs, self.ys, self.zs = xs, ys, zs
        self.f, self.amplitude = f, amplitude
        self.is_first_update = True
        self.phase = phase
        self.vp = self.beta * c_light
        self.w = 2 * np.pi * self.f
        self.kz = self.w / c_light
        self.tmax = np.inf
        if self.nodes is not None:
            self.tmax = self.nodes / self.f

    def update(self, solver, t):
        if self.is_first_update:
            if self.xs is None:
                self.xs = slice(0, solver.Nx)
            if self.ys is None:
                self.ys = slice(0, solver.Ny)
            self.is_first_update = False
        if t <= self.tmax:
            solver.H[self.xs, self.ys, self.zs, "y"] = -self.amplitude * np.cos(
                self.w * t + self.phase)
            solver.E[self.xs, self.ys, self.zs, "x"] = (
                self.amplitude * np.cos(self.w * t + self.phase) / (self.kz / (mu_0 * self.vp)))
        else:
            solver.H[self.xs, self.ys, self.zs, "y"] = 0.0
            solver.E[self.xs, self.ys, self.zs, "x"] = 0.0


class WavePacket:
    """Gaussian wave packet travelling in z, written on a z plane
    (sources.py:285-455).

    H_y = -A cos(w t + phi) g_xy g_t and E_x = A mu0 c cos(w t + phi) g_xy g_t,
    g_xy = exp(-(X^2 + Y^2) / (2 sigma_xy^2)) over the plane (NumPy
    ``meshgrid`` of x[xs], y[ys], default 'xy' indexing, as the reference),
    g_t = exp(-(s - s0)^2 / (2 sigma_z^2)) with s = z_min - beta c t and
    s0 = z_min - t_inj. Defaults: sigma_z = 10 mean(dz), sigma_xy = 5 mean of
    mean(dx), mean(dy), t_inj = 6 sigma_z; f from the wavelength when f is
    None (sources.py:380-395).
    """

    def __init__(self, xs=None, ys=None, zs=0, sigmaz=None, sigmaxy=None, tinj=None,
                 wavelength=None, f=None, amplitude=1.0, beta=1.0, phase=0):
        self.beta = beta
        self.xs, self.ys, self.zs = xs, ys, zs
        self.f, self.wavelength = f, wavelength
        self.sigmaxy, self.sigmaz, self.tinj = sigmaxy, sigmaz, tinj
        self.amplitude, self.phase = amplitude, phase
        if self.f is None:
            self.f = c_light / self.wavelength
        self.w = 2 * np.pi * self.f
        if self.tinj is None:
            self.tinj = 6 * self.sigmaz
        self.is_first_update = True
        self._gxy = None

    def update(self, solver, t):
        if self.is_first_update:
            if self.xs is None:
                self.xs = slice(0, solver.Nx)
            if self.ys is None:
                self.ys = slice(0, solver.Ny)
            if self.sigmaz is None:
                self.sigmaz = 10 * np.mean(solver.dz)
            if self.sigmaxy is None:
                self.sigmaxy = 5 * np.mean([np.mean(solver.dx), np.mean(solver.dy)])
            if self.tinj is None:
                self.tinj = 6 * self.sigmaz
            X, Y = np.meshgrid(solver.x[self.xs], solver.y[self.ys])
            gxy = np.exp(-(X ** 2 + Y ** 2) / (2 * self.sigmaxy ** 2))
            self._gxy = torch.as_tensor(gxy, dtype=torch.float64, device=solver.device)
            self.is_first_update = False
        s0 = solver.z.min() - self.tinj
        s = solver.z.min() - self.beta * c_light * t
        gausst = np.exp(-((s - s0) ** 2) / (2 * self.sigmaz ** 2))
        cw = np.cos(self.w * t + self.phase)
        # (-A cos) g_xy g_t and (A mu0 c cos) g_xy g_t, left to right as in NumPy
        solver.H[self.xs, self.ys, self.zs, "y"] = (self._gxy * float(-self.amplitude * cw)
                                                    * float(gausst))
        solver.E[self.xs, self.ys, self.zs, "x"] = (
            self._gxy * float(self.amplitude * mu_0 * c_light * cw) * float(gausst))


def _field_write(solver, field, key, value):
    if field == "E":
        solver.E[key] = value
    elif field == "H":
        solver.H[key] = value
    elif field == "J":
        solver.J[key] = value
    else:
        print(f'Field "{field}" not valid, should be "E", "H" or "J"]')


class Dipole:
    """Sinusoidal point (or slice) source A sin(w t + phi) on one component of
    E, H or J (sources.py:458-566). Defaults: the grid centre; f = nodes / T
    with T = (z_max - z_min) / c."""

    def __init__(self, field="E", component="z", xs=None, ys=None, zs=None, nodes=10, f=None,
                 amplitude=1.0, phase=0):
        self.nodes = nodes
        self.xs, self.ys, self.zs = xs, ys, zs
        self.f, self.field, self.component = f, field, component
        self.amplitude, self.phase = amplitude, phase
        if len(field) == 2:  # 'Ex' style
            self.component = field[1]
            self.field = field[0]
        self.is_first_update = True

    def update(self, solver, t):
        if self.is_first_update:
            if self.xs is None:
                self.xs = int(solver.Nx / 2)
            if self.ys is None:
                self.ys = int(solver.Ny / 2)
            if self.zs is None:
                self.zs = int(solver.Nz / 2)
            if self.f is None:
                T = (solver.z.max() - solver.z.min()) / c_light
                self.f = self.nodes / T
            self.w = 2 * np.pi * self.f
            self.is_first_update = False
        _field_write(solver, self.field, (self.xs, self.ys, self.zs, self.component),
                     self.amplitude * np.sin(self.w * t + self.phase))


class Pulse:
    """Pulse of the given time profile on one component of E, H or J
    (sources.py:569-762). Profiles of tau = c t - delay: "Harris"
    (10 - 15 cos(2 pi tau/L) + 6 cos(4 pi tau/L) - cos(6 pi tau/L)) / 32 for
    tau < L, else 0; "Gaussian" exp(-(tau - 5 L/10)^2 / (2 (L/10)^2));
    "Rectangular" 1 for 0 < tau < L. Defaults: the grid centre; L = 50 dt."""

    def __init__(self, field="E", component="z", xs=None, ys=None, zs=None, shape="Harris",
                 L=None, amplitude=1.0, delay=0.0):
        self.xs, self.ys, self.zs = xs, ys, zs
        self.field, self.component = field, component
        self.amplitude, self.shape, self.L, self.delay = amplitude, shape, L, delay
        if len(field) == 2:
            self.component = field[1]
            self.field = field[0]
        profiles = {"harris": self.harris_pulse, "gaussian": self.gaussian_pulse,
                    "rectangular": self.rectangular_pulse}
        if shape.lower() in profiles:
            self.tprofile = profiles[shape.lower()]
        else:
            print('** shape does not, match available types: "Harris", "Gaussian", '
                  '"Rectangular"')
        self.is_first_update = True

    def harris_pulse(self, t):
        t = t * c_light - self.delay
        prof = lambda u: (10 - 15 * np.cos(2 * np.pi / self.L * u)
                          + 6 * np.cos(4 * np.pi / self.L * u)
                          - np.cos(6 * np.pi / self.L * u)) / 32
        if np.ndim(t) == 0:
            return prof(t) if t < self.L else 0.0
        return prof(t)  # time arrays: the profile without the cut (as the reference)

    def gaussian_pulse(self, t):
        t = t * c_light - self.delay
        return np.exp(-((t - 5 * (self.L / 10)) ** 2) / (2 * (self.L / 10) ** 2))

    def rectangular_pulse(self, t):
        t = t * c_light - self.delay
        return 1.0 if (t < self.L and t > 0.0) else 0.0

    def update(self, solver, t):
        if self.is_first_update:
            if self.xs is None:
                self.xs = int(solver.Nx / 2)
            if self.ys is None:
                self.ys = int(solver.Ny / 2)
            if self.zs is None:
                self.zs = int(solver.Nz / 2)
            if self.L is None:
                self.L = 50 * solver.dt
            self.is_first_update = False
        _field_write(solver, self.field, (self.xs, self.ys, self.zs, self.component),
                     self.amplitude * self.tprofile(t))
