"""The DFSPH liquid plug-in (capability "sph", route "liquid_sph"): alloy/solvers/sph/__init__.py:4-19."""

from .config import DFSPHConfig, particle_mass, snap_down, stable_dt
from .model import GeoType, Model, ModelBuilder, ParticleFlags, State
from .types import SPHReport

__all__ = ["DFSPHConfig", "GeoType", "Model", "ModelBuilder", "ParticleFlags", "SPHReport", "SolverDFSPH", "State",
           "particle_mass", "snap_down", "stable_dt"]


def __getattr__(name):
    if name == "SolverDFSPH":
        from .solver import SolverDFSPH

        return SolverDFSPH
    raise AttributeError(name)
