"""CPU pins of the oracle's source restatements (sources.py:285-762) by closed
forms: the Harris pulse is 0 at tau = 0, 1 at tau = L/2, cut at L, and its
mean over [0, L] is 10/32; the Gaussian pulse peaks at 1 at tau = L/2 with
exp(-1/2) one sigma = L/10 away; the rectangular pulse is the open window
(0, L); the dipole's default frequency is nodes * c / (z_max - z_min); the wave
packet's plane is -A cos(w t + phi) exp(-r^2 / (2 sigma_xy^2)) at the
injection time (g_t = 1)."""

import numpy as np
import pytest
from scipy.constants import c as c_light
from scipy.constants import mu_0

from oracle import fit_oracle as fo


def _solver(dtype=np.float64):
    x, y, z = fo.uniform_axes(-1.0, 1.0, -1.0, 1.0, 0.0, 1.0, 9, 9, 12)
    return fo.OracleSolver(x, y, z, ["pec"] * 3, ["pec"] * 3, dtype=dtype)


def test_harris_pulse_closed_forms():
    L = 2.0
    p = fo.OraclePulse(shape="Harris", L=L)
    t = lambda tau: tau / c_light
    assert p.profile(t(0.0), L) == 0.0
    assert p.profile(t(L / 2), L) == pytest.approx(1.0, abs=1e-12)
    assert p.profile(t(L), L) == 0.0 and p.profile(t(3 * L), L) == 0.0
    tau = np.linspace(0.0, L, 200001)
    vals = np.array([p.profile(t(u), L) for u in tau[:-1]] + [0.0])
    assert np.trapezoid(vals, tau) / L == pytest.approx(10 / 32, rel=1e-6)


def test_gaussian_and_rectangular_pulses():
    L = 1.0
    g = fo.OraclePulse(shape="Gaussian", L=L)
    t = lambda tau: tau / c_light
    assert g.profile(t(L / 2), L) == pytest.approx(1.0, abs=1e-15)
    assert g.profile(t(L / 2 + L / 10), L) == pytest.approx(np.exp(-0.5), rel=1e-12)
    r = fo.OraclePulse(shape="Rectangular", L=L, delay=0.25)
    assert r.profile(t(0.25), L) == 0.0  # tau = 0: outside the open window
    assert r.profile(t(0.5), L) == 1.0
    assert r.profile(t(1.25), L) == 0.0


def test_dipole_default_frequency_and_phase():
    s = _solver()
    d = fo.OracleDipole(field="Ez", nodes=4, amplitude=3.0, phase=np.pi / 2)
    d.update(s, 0.0)
    Ez = fo.comp(s.E, s.dims, "z")
    assert Ez[4, 4, 6] == pytest.approx(3.0)  # sin(pi/2) at the grid centre
    f = 4 * c_light / (s.z.max() - s.z.min())
    d.update(s, 0.25 / f)  # a quarter period later: sin(pi) = 0
    assert abs(Ez[4, 4, 6]) < 1e-12


def test_wave_packet_plane_at_injection():
    s = _solver()
    sig = 0.5
    wp = fo.OracleWavePacket(zs=2, sigmaz=0.1, sigmaxy=sig, f=1e8, amplitude=2.0)
    t = wp.tinj / c_light  # s = s0: the temporal Gaussian is 1
    wp.update(s, t)
    X, Y = np.meshgrid(s.x, s.y)
    cw = np.cos(2 * np.pi * 1e8 * t)
    Hy = fo.comp(s.H, s.dims, "y")[:, :, 2]
    Ex = fo.comp(s.E, s.dims, "x")[:, :, 2]
    np.testing.assert_allclose(Hy, -2.0 * cw * np.exp(-(X ** 2 + Y ** 2) / (2 * sig ** 2)),
                               rtol=1e-12)
    np.testing.assert_allclose(Ex, -Hy * mu_0 * c_light, rtol=1e-12)
