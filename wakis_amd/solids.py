"""Analytic solids for staircase geometry fills.

The reference imports STL surfaces with pyvista and marks a cell when any of
its 8 corner nodes lies inside the surface (gridFIT3D.py:470-517,
``point_data_to_cell_data()['SelectedPoints'] > tol``). This module gives the
same "any corner inside" rule for solids described analytically (boxes,
cylinders, spheres and their unions/differences); a point is inside when it
lies strictly within the solid.
"""

import numpy as np


class Solid:
    """A point-membership test ``inside(X, Y, Z) -> bool array``."""

    def __init__(self, inside):
        self._inside = inside

    def inside(self, X, Y, Z):
        return self._inside(X, Y, Z)

    def __or__(self, other):
        return Solid(lambda X, Y, Z: self.inside(X, Y, Z) | other.inside(X, Y, Z))

    def __and__(self, other):
        return Solid(lambda X, Y, Z: self.inside(X, Y, Z) & other.inside(X, Y, Z))

    def __sub__(self, other):
        return Solid(lambda X, Y, Z: self.inside(X, Y, Z) & ~other.inside(X, Y, Z))

    def cell_mask(self, x, y, z):
        """(Nx,Ny,Nz) bool: cells with at least one corner node inside."""
        X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
        P = self.inside(X, Y, Z)
        return (P[:-1, :-1, :-1] | P[1:, :-1, :-1] | P[:-1, 1:, :-1] | P[1:, 1:, :-1]
                | P[:-1, :-1, 1:] | P[1:, :-1, 1:] | P[:-1, 1:, 1:] | P[1:, 1:, 1:])


class Box(Solid):
    def __init__(self, xmin, xmax, ymin, ymax, zmin, zmax):
        b = (xmin, xmax, ymin, ymax, zmin, zmax)
        self.bounds = b
        super().__init__(lambda X, Y, Z: (X > b[0]) & (X < b[1]) & (Y > b[2]) & (Y < b[3])
                         & (Z > b[4]) & (Z < b[5]))


class Cylinder(Solid):
    """Circular cylinder along ``axis`` with the given centre in the other two axes."""

    def __init__(self, radius, lo=-np.inf, hi=np.inf, axis="z", center=(0.0, 0.0)):
        r2 = radius * radius
        a = "xyz".index(axis)

        def inside(X, Y, Z):
            P = (X, Y, Z)
            u, w = [P[i] for i in range(3) if i != a]
            return ((u - center[0]) ** 2 + (w - center[1]) ** 2 < r2) & (P[a] > lo) & (P[a] < hi)

        super().__init__(inside)


class Sphere(Solid):
    def __init__(self, radius, center=(0.0, 0.0, 0.0)):
        r2 = radius * radius
        super().__init__(lambda X, Y, Z: (X - center[0]) ** 2 + (Y - center[1]) ** 2
                         + (Z - center[2]) ** 2 < r2)
