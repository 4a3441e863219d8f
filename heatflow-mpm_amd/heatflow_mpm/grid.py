"""Air on a closed MAC grid: configuration and the host view of the device state.

GridConfig keeps the reference's fields and defaults (grid.py:13-22) and adds the heat-solve
settings that the reference hard-codes in HeatField.diffuse (heat.py:198).  AirGrid is the
host-side handle on the air part of a scene held by libhfm.so: the same attributes and
diagnostics the reference's AirGrid offers (grid.py:177-263), computed on the GPU.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _abi


@dataclass
class GridConfig:
    L: float = 0.48            # box edge [m]
    h: float = 0.005           # cell size [m]
    rho: float = 1.2           # air density [kg/m^3]
    t_amb: float = 25.0        # room temperature [C]; side and top walls are held here, the floor is adiabatic
    buoyancy: bool = True
    alpha_f_min: float = 0.05  # floor of the air fraction inside dense solids
    cg_tol: float = 1e-4       # pressure PCG relative tolerance
    cg_max_iter: int = 400
    heat_tol: float = 1e-6     # heat PCG relative tolerance (fixed at 1e-6 in the reference)
    heat_max_iter: int = 200

    @property
    def n(self) -> int:
        return int(round(self.L / self.h))


def lattice_shape(field: str, n: int) -> tuple:
    """Shape of a field on its lattice (cell, x-, y- or z-face)."""
    if field in ("U", "AX", "QX", "KX", "MX"):
        return (n + 1, n, n)
    if field in ("V", "AY", "QY", "KY", "MY"):
        return (n, n + 1, n)
    if field in ("W", "AZ", "QZ", "KZ", "MZ"):
        return (n, n, n + 1)
    return (n, n, n)


class DeviceField:
    """A field held on the GPU; `.numpy()` copies it to the host like a Warp array."""

    def __init__(self, scene, name: str):
        self._scene, self.name = scene, name
        self.shape = lattice_shape(name, scene.n)

    def numpy(self) -> np.ndarray:
        return self._scene.get_field(self.name)

    def assign(self, values: np.ndarray) -> None:
        self._scene.set_field(self.name, values)


class AirGrid:
    """Face velocities u, v, w, cell pressure p and the obstacle fields of one scene."""

    def __init__(self, cfg: GridConfig, scene):
        self.cfg, self._scene = cfg, scene
        self.n = cfg.n
        self.h, self.L = cfg.h, cfg.L
        for name, attr in (("U", "u"), ("V", "v"), ("W", "w"), ("P", "p"), ("RHS", "rhs"), ("ALPHA_S", "alpha_s"),
                           ("AX", "ax"), ("AY", "ay"), ("AZ", "az"), ("QX", "qx"), ("QY", "qy"), ("QZ", "qz"),
                           ("KX", "kx"), ("KY", "ky"), ("KZ", "kz"), ("MX", "mx"), ("MY", "my"), ("MZ", "mz"),
                           ("PDIAG", "diag")):
            setattr(self, attr, DeviceField(scene, name))

    @property
    def cg_iters(self) -> int:
        """Pressure PCG iterations as warp.optim.linear.cg reports them (grid.py:249)."""
        return int(self._scene.stats()["p_iters"])

    def set_static(self, alpha: np.ndarray, k_faces: list, m_faces: list) -> None:
        """Static obstacle fields added every substep (grid.py:198-200)."""
        self._scene.set_static(alpha, k_faces, m_faces)

    def cell_velocity(self) -> np.ndarray:
        """Face-averaged cell velocity, (n^3, 3) float32 (grid.py:253-255)."""
        return self._scene.cell_velocity()

    def divergence_max(self) -> float:
        """max |div(a_f u + q)| * h (grid.py:257-260)."""
        return self._scene.diagnostics()[0]

    def speed_max(self) -> float:
        """max over u, v, w of |face velocity| (grid.py:262-263)."""
        return self._scene.diagnostics()[1]
