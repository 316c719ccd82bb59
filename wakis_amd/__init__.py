"""wakis_amd: MI355X-native FIT time-stepping engine, drop-in for Wakis' hot path.

Exports the reference's public names for the FIT leapfrog and wakefield path
(wakis/__init__.py:6-24): Field, GridFIT3D, SolverFIT3D, WakeSolver, Logger, plus the
Beam, PlaneWave, WavePacket, Dipole and Pulse sources and analytic solids. The compute path is the HIP
library libwakis_fit.so (include/wakis_fit.h); there is no CPU fallback.
"""

from . import field_monitors, solids, sources
from ._version import __version__
from .field import Field
from .field_monitors import FieldMonitor
from .gridFIT3D import GridFIT3D
from .logger import Logger
from .materials import material_lib
from .solverFIT3D import SolverFIT3D
from .sources import Beam, Dipole, PlaneWave, Pulse, WavePacket
from .wakeSolver import WakeSolver

__all__ = ["Field", "FieldMonitor", "GridFIT3D", "Logger", "SolverFIT3D", "WakeSolver", "Beam",
           "PlaneWave", "WavePacket", "Dipole", "Pulse", "material_lib", "field_monitors", "solids", "sources", "__version__"]
