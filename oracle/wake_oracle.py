"""CPU oracle for the WakeSolver post-processing (wakeSolver.py).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``). Works on the in-memory
Ez capture returned by ``fit_oracle.wakesolve_fields`` (what ``Ez.h5`` holds in
the reference). The double loop over (s, z) of the reference is vectorised
over s only, keeping the per-s accumulation order over z, so each WP value is
the same sequence of IEEE operations as the reference loop.
"""

import numpy as np
from scipy.constants import c as c_light


def wake_ti(sigmaz, beta, ti=None):
    """WakeSolver injection time (wakeSolver.py:164-172)."""
    v = beta * c_light
    if ti is not None:
        return ti
    return 8.548921333333334 * sigmaz / (np.sqrt(beta) * v)


def s_axis(ti, v, wakelength, t):
    """s = arange(-ti*v, wakelength, dt*v) with dt = t[2]-t[1] (wakeSolver.py:454-477)."""
    dt = t[2] - t[1]
    return np.arange(-ti * v, wakelength, dt * v)


def _line_integral(WP, Ezt, z, s, v, zmin, t0, ti, dt, dz, nt):
    """In-place ``WP[n] += Ezt[k, it]*dz[k]`` over k (wakeSolver.py:520-533)."""
    for k in range(len(z)):
        ts = (z[k] + s) / v - zmin / v - t0 + ti
        it = (ts / dt).astype(np.int64)  # int() truncates toward zero
        valid = it < nt
        idx = it[valid]
        idx = np.where(idx < 0, idx + Ezt.shape[1], idx)  # python negative index
        WP[valid] = WP[valid] + Ezt[k, idx] * dz[k]
    return WP


def _line_integral_counter_moving(WP, Ezt, z, s, v, zmax, t0, ti, dt, dz, nt):
    """In-place ``WP[n] += Ezt[-k-1, it]*(-1*dz[k])`` over k: the counter-moving
    branch of wakeSolver.py:503-517, with
    ``ts = (z[-k-1] - s)/(-1*v) - zmax/(-1*v) - t0 + ti``."""
    nz = len(z)
    for k in range(nz):
        ts = (z[-k - 1] - s) / (-1 * v) - zmax / (-1 * v) - t0 + ti
        it = (ts / dt).astype(np.int64)
        valid = it < nt
        idx = it[valid]
        idx = np.where(idx < 0, idx + Ezt.shape[1], idx)
        WP[valid] = WP[valid] + Ezt[-k - 1, idx] * (-1 * dz[k])
    return WP


def long_WP_3d(cap, q, v, ti, wakelength, skip_cells=0, n_transverse_cells=1,
               counter_moving=False):
    """``calc_long_WP_3d`` (wakeSolver.py:422-554), both branches.

    ``cap['Ez']`` has shape (nsamples, 3, 3, Nz). The running WP is zeroed
    once and carried from one transverse line to the next (wakeSolver.py:485,
    536-537), exactly as the reference does. ``counter_moving`` integrates
    along -z from zmax (wakeSolver.py:503-517).
    """
    t = cap["t"]
    Ez = cap["Ez"]
    nt = len(t)
    dt = t[2] - t[1]
    zf = cap["zf"]
    zmax = np.max(zf)
    zmin = np.min(zf)
    zz = slice(skip_cells, -skip_cells) if skip_cells != 0 else np.s_[:]
    z = zf[zz]
    nz = len(z)
    s = np.arange(-ti * v, wakelength, dt * v)
    i0 = j0 = n_transverse_cells
    WP = np.zeros_like(s)
    WP_3d = np.zeros((2 * i0 + 1, 2 * j0 + 1, len(s)))
    Ezt = np.zeros((nz, nt))  # allocated with len(t) columns, as the reference
    nt = min(nt, Ez.shape[0])
    for i in range(-i0, i0 + 1):
        for j in range(-j0, j0 + 1):
            Ezt[:, :nt] = np.asarray(Ez[:nt, Ez.shape[1] // 2 + i, Ez.shape[2] // 2 + j, zz],
                                     dtype=np.float64).T
            if counter_moving:
                WP = _line_integral_counter_moving(WP, Ezt, z, s, v, zmax, t[0], ti, dt,
                                                   cap["dz"], nt)
            else:
                WP = _line_integral(WP, Ezt, z, s, v, zmin, t[0], ti, dt, cap["dz"], nt)
            WP = WP / (q * 1e12)
            WP_3d[i0 + i, j0 + j, :] = WP
    return s, WP_3d[i0, j0, :].copy(), WP_3d


def long_WP_1d(cap, q, v, ti, wakelength, skip_cells=0):
    """``calc_long_WP`` (wakeSolver.py:310-420) on a capture of Ez samples.

    A line capture (nsamples, Nz) is assigned whole (lines 390-392, so
    ``skip_cells`` gives numpy's broadcast error there); a window capture
    (nsamples, nx, ny, Nz) reads the line [nx//2 + 1, ny//2 + 1, zz]
    (lines 385-388: for a 3x3 window the corner line at (+1, +1), not the
    centre). Ezt has min(len(t), nsamples) columns (allocated after that min,
    line 382), which is where a negative time index wraps.
    """
    t = cap["t"]
    Ez = cap["Ez"]
    nt = len(t)
    zf = cap["zf"]
    zz = slice(skip_cells, -skip_cells) if skip_cells != 0 else np.s_[:]
    z = zf[zz]
    nt = min(nt, Ez.shape[0])
    Ezt = np.zeros((len(z), nt))
    for n in range(nt):
        if Ez.ndim == 4:
            Ezt[:, n] = Ez[n, Ez.shape[1] // 2 + 1, Ez.shape[2] // 2 + 1, zz]
        else:
            Ezt[:, n] = Ez[n]  # the reference assigns the full line (no zz slicing)
    return long_WP_from_Ezt(Ezt, t, zf, cap["dz"], q, v, ti, wakelength, skip_cells, nt)


def long_WP_from_Ezt(Ezt, t, zf, dz, q, v, ti, wakelength, skip_cells=0, nt=None):
    """The integral of ``calc_long_WP`` (wakeSolver.py:395-408) over an
    (nz, ncols) array Ezt of Ez(x_test, y_test, z[zz], t) samples: terms with
    it < nt, nt = len(t) unless given (a negative it wraps over the ncols
    columns), then WP / (q*1e12)."""
    nt = len(t) if nt is None else nt
    dt = t[2] - t[1]
    zmin = np.min(zf)
    zz = slice(skip_cells, -skip_cells) if skip_cells != 0 else np.s_[:]
    z = zf[zz]
    s = np.arange(-ti * v, wakelength, dt * v)
    WP = np.zeros_like(s)
    WP = _line_integral(WP, np.asarray(Ezt, dtype=np.float64), z, s, v, zmin, t[0], ti, dt,
                        dz, nt)
    return s, WP / (q * 1e12)


def trans_WP(s, WP_3d, xf, yf, n_transverse_cells=1):
    """``calc_trans_WP`` (wakeSolver.py:556-629): prefix sums by np.sum, central diff."""
    ds = s[2] - s[1]
    i0 = j0 = n_transverse_cells
    WPx = np.zeros_like(s)
    WPy = np.zeros_like(s)
    int_WP = np.zeros_like(WP_3d)
    dxf = np.diff(xf)
    dyf = np.diff(yf)
    for n in range(len(s)):
        for i in range(-i0, i0 + 1):
            for j in range(-j0, j0 + 1):
                int_WP[i0 + i, j0 + j, n] = np.sum(WP_3d[i0 + i, j0 + j, 0:n]) * ds
        WPx[n] = -(int_WP[i0 + 1, j0, n] - int_WP[i0 - 1, j0, n]) / (dxf[i0 - 1] + dxf[i0])
        WPy[n] = -(int_WP[i0, j0 + 1, n] - int_WP[i0, j0 - 1, n]) / (dyf[j0 - 1] + dyf[j0])
    return WPx, WPy


def lambdas_analytic(s, sigmaz):
    """``calc_lambdas_analytic`` (wakeSolver.py:838-870)."""
    return 1 / (sigmaz * np.sqrt(2 * np.pi)) * np.exp(-(s ** 2) / (2 * sigmaz ** 2))


def dft_setup(s, v, fmax, samples=1001):
    """ds, zero-padded length N (wakeSolver.py:679-683)."""
    ds = np.mean(s[1:] - s[:-1])
    N = int((v / ds) // fmax * samples)
    return ds, N


def long_Z(s, WP, lambdas, v, fmax, samples=1001):
    """``calc_long_Z`` (wakeSolver.py:631-719)."""
    ds, N = dft_setup(s, v, fmax, samples)
    lambdafft = np.fft.fft(lambdas * v, n=N)
    WPfft = np.fft.fft(WP * 1e12, n=N)
    ffft = np.fft.fftfreq(len(WPfft), ds / v)
    mask = np.logical_and(ffft >= 0, ffft < fmax)
    WPf = WPfft[mask] * ds
    lambdaf = lambdafft[mask] * ds
    return ffft[mask], -WPf / lambdaf, lambdaf


def trans_Z(s, WPx, WPy, lambdas, v, fmax, samples=1001):
    """``calc_trans_Z`` (wakeSolver.py:721-781)."""
    ds, N = dft_setup(s, v, fmax, samples)
    lambdafft = np.fft.fft(lambdas * v, n=N)
    ffft = np.fft.fftfreq(len(lambdafft), ds / v)
    mask = np.logical_and(ffft >= 0, ffft < fmax)
    lambdaf = lambdafft[mask] * ds
    WPxf = np.fft.fft(WPx * 1e12, n=N)[mask] * ds
    WPyf = np.fft.fft(WPy * 1e12, n=N)[mask] * ds
    return ffft[mask], 1j * WPxf / lambdaf, 1j * WPyf / lambdaf


def wake_solve(cap, q, sigmaz, beta, wakelength, skip_cells=0, fmax=None, ti=None,
               compute_plane="both", counter_moving=False):
    """``WakeSolver.solve`` (wakeSolver.py:210-254) on an in-memory capture."""
    v = beta * c_light
    ti = wake_ti(sigmaz, beta, ti)
    fmax = v / (3.0 * sigmaz) if fmax is None else fmax
    out = {}
    if compute_plane.lower() in ("both", "transverse"):
        s, WP, WP_3d = long_WP_3d(cap, q, v, ti, wakelength, skip_cells,
                                  counter_moving=counter_moving)
        WPx, WPy = trans_WP(s, WP_3d, cap["xf"], cap["yf"])
        lam = lambdas_analytic(s, sigmaz)
        f, Z, lambdaf = long_Z(s, WP, lam, v, fmax)
        _, Zx, Zy = trans_Z(s, WPx, WPy, lam, v, fmax)
        out.update(s=s, WP=WP, WP_3d=WP_3d, WPx=WPx, WPy=WPy, lambdas=lam, f=f, Z=Z,
                   Zx=Zx, Zy=Zy, lambdaf=lambdaf)
    else:
        s, WP = long_WP_1d(cap, q, v, ti, wakelength, skip_cells)
        lam = lambdas_analytic(s, sigmaz)
        f, Z, lambdaf = long_Z(s, WP, lam, v, fmax)
        out.update(s=s, WP=WP, lambdas=lam, f=f, Z=Z, lambdaf=lambdaf)
    out.update(ti=ti, v=v, fmax=fmax)
    return out
