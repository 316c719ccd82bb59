"""Run log with the reference's keys and file layout (wakis/logger.py:40-97,
gridFIT3D.py:181-256, wakeSolver.py:1709-1723), on the host."""

import numpy as np

import wakis_amd as wa
from wakis_amd.logger import Logger

GRID_KEYS = ["Nx", "Ny", "Nz", "dx", "dy", "dz", "xmin", "xmax", "ymin", "ymax", "zmin", "zmax",
             "stl_solids", "stl_materials", "use_mesh_refinement", "gridInitializationTime"]
WAKE_KEYS = ["ti", "q", "sigmaz", "beta", "xsource", "ysource", "xtest", "ytest", "chargedist",
             "skip_cells", "results_folder"]


def test_grid_and_wake_logs_have_the_reference_keys():
    box = wa.solids.Box(-0.5, 0.5, -0.5, 0.5, -0.5, 0.5)
    g = wa.GridFIT3D(-1, 1, -1, 1, -2, 2, 6, 7, 8, stl_solids={"b": box},
                     stl_materials={"b": "vacuum"}, verbose=0)
    assert list(g.logger.grid) == GRID_KEYS
    assert g.logger.grid["Nz"] == 8 and g.logger.grid["zmin"] == -2
    assert g.logger.grid["stl_materials"] == {"b": [1.0, 1.0]}
    assert np.array_equal(g.logger.grid["dz"], np.diff(np.linspace(-2, 2, 9)))
    w = wa.WakeSolver(q=2e-9, sigmaz=5e-3, beta=0.9, xtest=1e-3, skip_cells=3, save=False,
                      verbose=0, results_folder="out")
    assert list(w.logger.wakeSolver) == WAKE_KEYS
    assert w.logger.wakeSolver["results_folder"] == "out/"
    assert w.logger.wakeSolver["ti"] == w.ti and w.logger.wakeSolver["xtest"] == 1e-3


def test_save_logs_layout_and_round_trip(tmp_path):
    lg = Logger()
    lg.grid = {"Nx": 4, "dx": np.array([0.5, 0.25]), "stl_materials": {"b": [1.0, 1.0]}}
    lg.solver = {"dt": np.float32(1.5e-12), "bc_low": ["pec", "pec", "abc"]}
    lg.wakeSolver = {"q": 1e-9, "results_folder": str(tmp_path / "res") + "/"}
    lg.save_logs()
    path = tmp_path / "res" / "wakis.log"
    text = path.read_text(encoding="utf-8")
    assert text.startswith("Simulation Parameters\n=====================\n\n")
    i_w, i_s, i_g = (text.index(f"## {t} Logs ##") for t in ("WakeSolver", "Solver", "Grid"))
    assert i_w < i_s < i_g
    back = Logger.read_logs(str(path))
    assert back["grid"] == {"Nx": 4, "dx": [0.5, 0.25], "stl_materials": {"b": [1.0, 1.0]}}
    # a float32 is not JSON: written as its str, as the reference does
    assert back["solver"] == {"dt": str(np.float32(1.5e-12)), "bc_low": ["pec", "pec", "abc"]}
    assert back["wakeSolver"]["q"] == 1e-9
    empty = Logger()
    empty.wakeSolver["results_folder"] = str(tmp_path / "e")
    empty.save_logs()
    assert Logger.read_logs(str(tmp_path / "e" / "wakis.log"))["grid"] is None


def test_wake_solver_copy_is_shallow():
    w = wa.WakeSolver(q=3e-9, sigmaz=4e-3, save=False, verbose=0)
    c = w.copy()
    assert c is not w and type(c) is wa.WakeSolver
    assert c.__dict__ == w.__dict__ and c.logger is w.logger
