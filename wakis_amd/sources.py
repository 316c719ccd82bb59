"""Time-dependent sources (wakis/sources.py).

``Beam`` (sources.py:22-110) writes the Gaussian beam current on the z line
through the source. ``Beam.table`` evaluates the same expression for every
step of a run at once (same per-element IEEE operations as the per-step
update, including NumPy's scalar promotion of ``n*dt``), so ``wakesolve`` can
upload it once and let the step kernel read one row per step.
``PlaneWave`` (sources.py:142-282) writes E/H on a z plane from the host
between device steps.
"""

import numpy as np
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
