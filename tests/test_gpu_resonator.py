"""GPU: the reference's analytic resonator test (tests/test_002_analytic_resonator.py)
through the device engine's user-loop API: a 1 m PEC cube on 30^3 cells, H set to
the TE101 mode at t = -dt/2, 5000 one_step calls with Ey read at the centre
after each; the spectrum peaks at the analytic c/2*sqrt(2) = 2.121e8 Hz within 5%."""

import math

import numpy as np
import pytest
from scipy.constants import c as c_light

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_te101_resonance(dtype):
    import wakis_amd as wa
    print("te101 resonator", np.dtype(dtype).name, flush=True)
    L, N = 1.0, 30
    d = L / N
    grid = wa.GridFIT3D(-L / 2 + d / 2, L / 2 + d / 2, -L / 2 + d / 2, L / 2 + d / 2,
                        -L / 2 + d / 2, L / 2 + d / 2, N, N, N, verbose=0)
    s = wa.SolverFIT3D(grid, bc_low=["pec"] * 3, bc_high=["pec"] * 3, dtype=dtype, verbose=0)
    m, n, p = 1, 0, 1
    ii = np.arange(N)
    X, Y, Z = np.meshgrid(ii * d + grid.x[0], ii * d + grid.y[0], ii * d + grid.z[0],
                          indexing="ij")
    t = -0.5 * float(s.dt)
    h2 = (m * np.pi / L) ** 2 + (n * np.pi / L) ** 2 + (p * np.pi / L) ** 2
    wt = np.cos(np.sqrt(2) * np.pi / L * c_light * t)
    Hz = (np.cos(m * np.pi / L * (X - L / 2)) * np.cos(n * np.pi / L * (Y - L / 2))
          * np.sin(p * np.pi / L * (Z - L / 2)) * wt)
    Hx = (-2 / h2 * (m * np.pi / L) * (p * np.pi / L) * np.sin(m * np.pi / L * (X - L / 2))
          * np.cos(n * np.pi / L * (Y - L / 2)) * np.cos(p * np.pi / L * (Z - L / 2)) * wt)
    s.H[:, :, :, "x"] = Hx.astype(dtype)
    s.H[:, :, :, "z"] = Hz.astype(dtype)
    Ey = []
    for _ in range(5000):
        Ey.append(float(s.E[N // 2, N // 2, N // 2, "y"]))
        s.one_step()
    Ey = np.array(Ey)
    F = np.fft.fftshift(np.fft.fft(Ey))[len(Ey) // 2:]
    f = np.fft.fftshift(np.fft.fftfreq(len(Ey), d=float(s.dt)))[len(Ey) // 2:]
    fr = f[np.argmax(np.abs(F))]
    assert fr == pytest.approx(2.121e8, 0.05)
    assert fr == pytest.approx(c_light / 2 * math.sqrt(2), 0.05)
