"""Material parameters and the particle body that feeds the substep.

The reference drives its particles with NVIDIA Newton's implicit MPM (mpm.py:25-80).  The
constitutive solve is outside this library: the substep takes particle positions, velocities
and temperatures and returns temperatures, melt fractions, yield stresses and viscosities.
`StaticBody` is the body used when nothing moves the particles (the benchmark scenes hold them
static); any object with `pos`/`vel` float32 (P, 3) arrays and a `step(dt)` method can take its
place.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Material:
    """Same fields and defaults as the reference dataclass (mpm.py:9-22)."""
    rho: float = 1300.0              # kg/m^3
    t_sol: float = 30.0              # fully solid below [C]
    t_liq: float = 36.0              # fully liquid above [C]
    yield_solid: float = 2.0e4       # Pa at f = 0
    viscosity_solid: float = 5.0     # Pa s at f = 0
    viscosity_liquid: float = 30.0   # Pa s at f = 1
    kappa: float = 2.0e-5            # thermal diffusivity in the material [m^2/s]
    ground_friction: float = 0.3
    strain_basis: str = "Q1"
    voxel_mult: float = 2.0
    max_iterations: int = 250
    tolerance: float = 1.0e-5


class StaticBody:
    """Particles that do not move: positions fixed, velocities zero (or as given).

    Stands in for ThermalMPM (mpm.py:25-80) on the parts the substep reads: pos, vel,
    positions(), velocities(), step(dt).
    """

    static = True   # the substep uploads particle positions only for bodies without this flag

    def __init__(self, particles: np.ndarray, velocities: np.ndarray | None = None):
        self.pos = np.ascontiguousarray(particles, dtype=np.float32).reshape(-1, 3)
        self.vel = (np.zeros_like(self.pos) if velocities is None
                    else np.ascontiguousarray(velocities, dtype=np.float32).reshape(-1, 3))
        if self.vel.shape != self.pos.shape:
            raise ValueError(f"velocities {self.vel.shape} do not match positions {self.pos.shape}")
        self.n = len(self.pos)
        self.moved = False

    def step(self, dt: float) -> None:
        """A static body does not move."""

    def positions(self) -> np.ndarray:
        return self.pos.copy()

    def velocities(self) -> np.ndarray:
        return self.vel.copy()


class KinematicBody:
    """Particles moved by a prescribed velocity field: each step(dt) advances positions by
    dt * velocity(pos, t) and stores that velocity.  A stand-in for ThermalMPM (mpm.py:25-80)
    where the material moves but its motion is given rather than solved, so the substep's
    moving-material paths (particle flux q and relaxation target m from nonzero velocities,
    re-binning after every move) run on known motion.  `velocity(pos, t)` returns a (P, 3)
    float32 array; the initial velocities are velocity(pos, 0)."""

    static = False

    def __init__(self, particles: np.ndarray, velocity):
        self.pos = np.ascontiguousarray(particles, dtype=np.float32).reshape(-1, 3)
        self.vel = np.ascontiguousarray(velocity(self.pos, 0.0), dtype=np.float32).reshape(-1, 3)
        self.n = len(self.pos)
        self.velocity, self.t = velocity, 0.0

    def step(self, dt: float) -> None:
        self.pos = np.ascontiguousarray(self.pos + np.float32(dt) * self.vel, dtype=np.float32)
        self.t += dt
        self.vel = np.ascontiguousarray(self.velocity(self.pos, self.t), dtype=np.float32).reshape(-1, 3)

    def positions(self) -> np.ndarray:
        return self.pos.copy()

    def velocities(self) -> np.ndarray:
        return self.vel.copy()
