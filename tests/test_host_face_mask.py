"""CPU: the face masks of the fold and the fused-path checks.

``face_mask_kji`` writes the faces as planes in x, y, z order, so that the
last face wins on shared edges; checked against the rule written per cell
(each cell takes the last face in that order whose plane holds it), for every
component, every set of masked faces and grids with one-cell axes.
"""

import itertools

import numpy as np
import pytest
import torch

from wakis_amd.solverFIT3D import face_mask_kji


def _per_cell(comp, zbits, Nx, Ny, Nz):
    order = {0: [("j", 0, 1), ("j", Ny - 1, 4), ("k", 0, 2), ("k", Nz - 1, 5)],
             1: [("i", 0, 0), ("i", Nx - 1, 3), ("k", 0, 2), ("k", Nz - 1, 5)],
             2: [("i", 0, 0), ("i", Nx - 1, 3), ("j", 0, 1), ("j", Ny - 1, 4)]}[comp]
    keep = np.ones((Nz, Ny, Nx), dtype=bool)
    for k, j, i in itertools.product(range(Nz), range(Ny), range(Nx)):
        pos = {"i": i, "j": j, "k": k}
        face = -1
        for ax, val, f in order:
            if pos[ax] == val:
                face = f
        if face >= 0 and (zbits >> face) & 1:
            keep[k, j, i] = False
    return keep


@pytest.mark.parametrize("dims", [(4, 3, 5), (1, 3, 2), (3, 1, 1)])
def test_face_masks_match_the_per_cell_rule(dims):
    Nx, Ny, Nz = dims
    for comp in range(3):
        for zbits in range(64):
            got = face_mask_kji(comp, zbits, Nx, Ny, Nz, "cpu").numpy()
            assert np.array_equal(got, _per_cell(comp, zbits, Nx, Ny, Nz)), (comp, zbits)
