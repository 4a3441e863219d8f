"""Heatflow-MPM air/heat substep on AMD MI355X (gfx950): HIP kernels behind a ctypes C ABI.

Public names follow the reference package (heatflow_mpm/__init__.py:2-7).  Newton's ThermalMPM is
not part of this path: particles come from a body object (StaticBody by default, KinematicBody
for prescribed motion); `mesh_particles` seeds material in a mesh (winding-number inside test on
the GPU) and `scenes.ellipsoid_particles` seeds synthetic material.
"""
from .solver import HeatflowMPM, KAPPA_AIR, K_PARTICLE, scene_inputs  # noqa: F401
from .grid import GridConfig, AirGrid  # noqa: F401
from .heat import HeatField  # noqa: F401
from .heater import Heater  # noqa: F401
from .mpm import KinematicBody, Material, StaticBody  # noqa: F401
from .scenes import ellipsoid_particles, plume_scene, coupled_scene, batched_scenes  # noqa: F401
from .batch import HeatflowBatch  # noqa: F401
from .serve import Request, Server, poisson_requests  # noqa: F401
from .frames import FrameWriter  # noqa: F401
from .shapes import load_mesh, sample_interior, mesh_particles, BUNNY  # noqa: F401

__version__ = "0.2.0"
