"""Canonical oracle set-ups (test infrastructure only).

* ``test001``: the reference's tests/test_001_pec_cubic_cavity.py:71-146 case,
  with the STL solid replaced by its exact box decomposition (union of the
  cavity box and the square pipe; bounds as float32 values of the STL file).
* ``cavity_pipe``: PEC cubic cavity 50 mm with a round pipe r=5 mm along z in a
  60x60x160 mm domain (the 48x48x128 and 256x256x512 throughput workloads),
  PEC on x/y, ABC on z.
* ``pillbox``: cylindrical cavity with a beam pipe and a conductive shell wall in a
  PEC background (the lossy pillbox + PML workload).
"""

import json
from pathlib import Path

import numpy as np

from . import fit_oracle as fo

GOLDEN = Path(__file__).resolve().parents[1] / "tests" / "golden" / "test001_golden.json"


def test001_geometry():
    g = json.loads(GOLDEN.read_text())
    b = g["grid"]["bounds_float32"]
    lv = g["solid_levels_float32"]
    x, y, z = fo.uniform_axes(b[0], b[1], b[2], b[3], b[4], b[5],
                              g["grid"]["Nx"], g["grid"]["Ny"], g["grid"]["Nz"])
    cav_xy, pipe_xy = lv[0][3], lv[0][2]
    cav_z, pipe_z = lv[2][2], lv[2][3]

    def inside(X, Y, Z):
        cav = (np.abs(X) < cav_xy) & (np.abs(Y) < cav_xy) & (np.abs(Z) < cav_z)
        pipe = (np.abs(X) < pipe_xy) & (np.abs(Y) < pipe_xy) & (np.abs(Z) < pipe_z)
        return cav | pipe

    return x, y, z, fo.corner_any_inside(x, y, z, inside), g


def cavity_pipe_geometry(Nx, Ny, Nz, half_cav=25e-3, r_pipe=5e-3,
                         half_x=30e-3, half_z=80e-3):
    x, y, z = fo.uniform_axes(-half_x, half_x, -half_x, half_x, -half_z, half_z, Nx, Ny, Nz)

    def inside(X, Y, Z):
        cav = (np.abs(X) < half_cav) & (np.abs(Y) < half_cav) & (np.abs(Z) < half_cav)
        pipe = X ** 2 + Y ** 2 < r_pipe ** 2
        return cav | pipe

    return x, y, z, fo.corner_any_inside(x, y, z, inside)


def test001_solver(dtype=np.float32):
    x, y, z, mask, g = test001_geometry()
    s = fo.OracleSolver(x, y, z, g["bc"]["low"], g["bc"]["high"], bg="pec",
                        solids=[(mask, "vacuum")], dtype=dtype)
    return s, g


def cavity_pipe_solver(Nx, Ny, Nz, dtype=np.float64):
    x, y, z, mask = cavity_pipe_geometry(Nx, Ny, Nz)
    return fo.OracleSolver(x, y, z, ["pec", "pec", "abc"], ["pec", "pec", "abc"],
                           bg="pec", solids=[(mask, "vacuum")], dtype=dtype)


def pillbox_geometry(Nx, Ny, Nz, r_cav=40e-3, half_len=30e-3, r_pipe=10e-3, wall_cells=3,
                     half_x=45e-3, half_z=90e-3):
    """Lossy pillbox: vacuum cylinder (radius r_cav, |z| < half_len) plus a beam
    pipe (radius r_pipe, all z); a shell of ``wall_cells`` cells around the vacuum
    is the conductive wall, the rest is PEC background. Returns (x, y, z,
    vacuum_mask, wall_mask)."""
    x, y, z = fo.uniform_axes(-half_x, half_x, -half_x, half_x, -half_z, half_z, Nx, Ny, Nz)

    def inside(X, Y, Z):
        R2 = X ** 2 + Y ** 2
        return ((R2 < r_cav ** 2) & (np.abs(Z) < half_len)) | (R2 < r_pipe ** 2)

    vac = fo.corner_any_inside(x, y, z, inside)
    grown = vac.copy()
    for _ in range(wall_cells):  # 6-neighbour dilation, one cell per pass
        g = grown.copy()
        g[1:] |= grown[:-1]
        g[:-1] |= grown[1:]
        g[:, 1:] |= grown[:, :-1]
        g[:, :-1] |= grown[:, 1:]
        g[:, :, 1:] |= grown[:, :, :-1]
        g[:, :, :-1] |= grown[:, :, 1:]
        grown = g
    return x, y, z, vac, grown & ~vac
