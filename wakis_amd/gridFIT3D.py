"""Cartesian FIT grid with analytic staircase solids.

Mirrors ``wakis.gridFIT3D.GridFIT3D`` (gridFIT3D.py:32-300): node axes from
domain extents (``np.linspace``) or explicit arrays, cell sizes by
``np.diff``, and the primal/dual metrics of ``_compute_grid``. STL import
(pyvista), mesh refinement, MPI slabs and HDF5 save/load are not part of this
engine; solids are given as ``wakis_amd.solids`` objects and marked with the
reference's "any corner inside" rule.
"""

import time

import numpy as np

from .logger import Logger
from .materials import material_lib
from .solids import Solid


class GridFIT3D:
    def __init__(self, xmin=None, xmax=None, ymin=None, ymax=None, zmin=None, zmax=None,
                 Nx=None, Ny=None, Nz=None, x=None, y=None, z=None, use_mpi=False,
                 stl_solids=None, stl_materials=None, stl_colors=None, stl_tol=1e-3,
                 verbose=1, **kwargs):
        t0 = time.time()
        self.logger = Logger()
        if use_mpi:
            raise NotImplementedError("wakis_amd runs the full domain on one GPU (use_mpi=False)")
        unsupported = {k for k, v in kwargs.items()
                       if v not in (None, False, 1.0, [0.0, 0.0, 0.0], "insert", 1e-8, 1e-5)}
        if unsupported:
            raise NotImplementedError(f"GridFIT3D options not supported: {sorted(unsupported)}")
        self.verbose = verbose
        self.use_mpi = False
        if x is not None and y is not None and z is not None:
            self.x, self.y, self.z = np.asarray(x), np.asarray(y), np.asarray(z)
            self.Nx, self.Ny, self.Nz = len(self.x) - 1, len(self.y) - 1, len(self.z) - 1
            self.xmin, self.xmax = self.x[0], self.x[-1]
            self.ymin, self.ymax = self.y[0], self.y[-1]
            self.zmin, self.zmax = self.z[0], self.z[-1]
        elif all(v is not None for v in (xmin, xmax, ymin, ymax, zmin, zmax)):
            self.xmin, self.xmax, self.ymin, self.ymax = xmin, xmax, ymin, ymax
            self.zmin, self.zmax = zmin, zmax
            self.Nx, self.Ny, self.Nz = Nx, Ny, Nz
            self.x = np.linspace(self.xmin, self.xmax, self.Nx + 1)
            self.y = np.linspace(self.ymin, self.ymax, self.Ny + 1)
            self.z = np.linspace(self.zmin, self.zmax, self.Nz + 1)
        else:
            raise ValueError("[!] Error initializing GridFIT3D: provide x, y, z arrays or "
                             "xmin, xmax, ymin, ymax, zmin, zmax, Nx, Ny, Nz")
        self.N = self.Nx * self.Ny * self.Nz
        self.dx, self.dy, self.dz = np.diff(self.x), np.diff(self.y), np.diff(self.z)
        # dual axis: midpoints with the last one repeated (gridFIT3D.py:278-284)
        self.tx = np.append((self.x[1:] + self.x[:-1]) / 2, ((self.x[1:] + self.x[:-1]) / 2)[-1])
        self.ty = np.append((self.y[1:] + self.y[:-1]) / 2, ((self.y[1:] + self.y[:-1]) / 2)[-1])
        self.tz = np.append((self.z[1:] + self.z[:-1]) / 2, ((self.z[1:] + self.z[:-1]) / 2)[-1])
        self.tdx, self.tdy, self.tdz = np.diff(self.tx), np.diff(self.ty), np.diff(self.tz)

        self.stl_solids = {} if stl_solids is None else stl_solids
        self.stl_materials = {} if stl_materials is None else dict(stl_materials)
        self.stl_colors = stl_colors
        self.stl_tol = stl_tol
        if isinstance(self.stl_solids, (str, Solid)):
            self.stl_solids = {"Solid 1": self.stl_solids}
        if isinstance(self.stl_materials, (str, list, tuple)):
            self.stl_materials = {"Solid 1": self.stl_materials}
        self.grid = {}
        for key, solid in self.stl_solids.items():
            if isinstance(solid, str):
                raise NotImplementedError(
                    f"STL file import needs pyvista; give solid {key!r} as a wakis_amd.solids "
                    "object (Box, Cylinder, Sphere, or unions of them)")
            if isinstance(solid, np.ndarray):  # precomputed (Nx,Ny,Nz) cell mask
                if solid.shape != (self.Nx, self.Ny, self.Nz):
                    raise ValueError(f"solid {key!r}: mask shape {solid.shape} != grid cells")
                self.grid[key] = solid.astype(bool)
            else:
                if not isinstance(solid, Solid):
                    solid = Solid(solid)
                self.grid[key] = solid.cell_mask(self.x, self.y, self.z)
            mat = self.stl_materials.get(key)
            if isinstance(mat, str):
                lib = material_lib[mat.lower()]
                self.stl_materials[key] = list(lib)
        # the reference's grid log keys (gridFIT3D.py:181-208, 255-256); solids
        # are named by their description or mask shape instead of an STL path
        self.use_mesh_refinement = False
        self.update_logger(["Nx", "Ny", "Nz", "dx", "dy", "dz"])
        self.update_logger(["xmin", "xmax", "ymin", "ymax", "zmin", "zmax"])
        self.logger.grid["stl_solids"] = {
            k: (f"cell mask {tuple(v.shape)}, {int(np.count_nonzero(v))} cells"
                if isinstance(v, np.ndarray) else str(v)) for k, v in self.stl_solids.items()}
        self.update_logger(["stl_materials", "use_mesh_refinement"])
        self.gridInitializationTime = time.time() - t0
        self.update_logger(["gridInitializationTime"])

    def update_logger(self, attrs):
        """Copy grid attributes into ``logger.grid`` (gridFIT3D.py:1001-1011)."""
        for atr in attrs:
            self.logger.grid[atr] = getattr(self, atr)

    def _mark_cells_in_surface(self, key):
        """Keep only the surface cells of solid ``key`` (gridFIT3D.py:518-533), for the
        SIBC substitution of a highly conductive solid: a cell stays marked where the
        finite-difference gradient of the solid's 0/1 cell field is nonzero, i.e. where
        the field differs between its two neighbours along some axis (central
        difference), or from its one neighbour on the first and last cell of an axis.
        The reference takes the gradient with pyvista's ``compute_derivative``."""
        m = np.asarray(self.grid[key]).astype(np.int8)
        surf = np.zeros(m.shape, dtype=bool)
        for ax in range(3):
            if m.shape[ax] < 2:
                continue
            a = np.moveaxis(m, ax, 0)
            s = np.moveaxis(surf, ax, 0)  # view: writes land in surf
            s[1:-1] |= a[2:] != a[:-2]
            s[0] |= a[1] != a[0]
            s[-1] |= a[-1] != a[-2]
        self.grid[key] = surf
        return surf

    # Full 3N metric Fields (gridFIT3D.py:258-300), built on demand on the host.
    def _metric(self, which):
        Nx, Ny, Nz = self.Nx, self.Ny, self.Nz
        ax = {"L": (self.dx, self.dy, self.dz), "tL": (self.tdx, self.tdy, self.tdz)}
        if which in ax:
            vx, vy, vz = ax[which]
            return np.concatenate([
                np.broadcast_to(vx[:, None, None], (Nx, Ny, Nz)).reshape(-1, order="F"),
                np.broadcast_to(vy[None, :, None], (Nx, Ny, Nz)).reshape(-1, order="F"),
                np.broadcast_to(vz[None, None, :], (Nx, Ny, Nz)).reshape(-1, order="F")])
        L = self._metric("L" if which == "iA" else "tL")
        N = self.N
        Lx, Ly, Lz = L[:N], L[N:2 * N], L[2 * N:]
        out = np.zeros(3 * N)
        for c, aux in enumerate((Ly * Lz, Lx * Lz, Lx * Ly)):
            if which == "iA":
                out[c * N:(c + 1) * N] = np.divide(1.0, aux)
            else:
                out[c * N:(c + 1) * N] = np.divide(1.0, aux, out=np.zeros_like(aux),
                                                   where=aux != 0)
        return out

    @property
    def L(self):
        return self._metric("L")

    @property
    def tL(self):
        return self._metric("tL")

    @property
    def iA(self):
        return self._metric("iA")

    @property
    def itA(self):
        return self._metric("itA")
