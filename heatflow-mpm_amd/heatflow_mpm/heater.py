"""Electric heater as grid masks (host side, evaluated once per scene).

Restates heatflow_mpm/heater.py:1-65 of the reference: the trough heater is a reflector shell
and end caps (solid), two rods (solid, fixed temperature), an optional fan slab in front of the
opening and the trough air (both power sources).  Geometry is evaluated in float64 on the cell
centres with the same comparisons as the reference, so the masks are the same cell for cell.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

# Face relaxation weights [kg/s] and air heat capacity [J/(kg K)] (heater.py:7).
K_SOLID, K_FAN, CP_AIR = 1.0, 0.05, 1005.0

# Cell type codes (heater.py:34).
AIR, SHELL, ROD, FAN, TROUGH = 0, 1, 2, 3, 4


@dataclass
class Heater:
    """Same fields and defaults as the reference dataclass (heater.py:10-22)."""
    dist: float = 0.21                    # distance from the box centre [m]
    direction: tuple = (-0.423, 0.906)    # horizontal unit direction from the centre to the heater
    z: float = 0.12                       # height of the trough axis [m]
    tilt: float = 15.0                    # face tilted down [deg]
    radius: float = 0.055                 # reflector radius [m]
    length: float = 0.22                  # trough length [m]
    t_rod: float = 300.0                  # rod temperature [C], Dirichlet
    power: float = 400.0                  # heat injected into the trough air [W]
    t_cap: float = 120.0                  # the heated air never exceeds this [C]
    fan: float = 0.0                      # outlet air speed [m/s], 0 = no fan
    on: bool = True

    def frame(self, centre_xy):
        """Head centre and the unit vectors (facing, axis, up) of the trough (heater.py:24-31).

        facing points from the heater toward the box centre, tilted down by `tilt`; axis is the
        horizontal trough axis; up completes the right-handed frame as axis x facing.
        """
        dirn = np.asarray(self.direction, float)
        dirn = dirn / np.linalg.norm(dirn)
        centre = np.array([centre_xy[0] + self.dist * dirn[0], centre_xy[1] + self.dist * dirn[1], self.z])
        tilt = np.deg2rad(self.tilt)
        facing = np.array([-dirn[0] * np.cos(tilt), -dirn[1] * np.cos(tilt), -np.sin(tilt)])
        axis = np.array([dirn[1], -dirn[0], 0.0])
        return centre, facing, axis, np.cross(axis, facing)

    def masks(self, n: int, h: float, centre_xy, rho_air: float):
        """Cell types, solid/source masks, source rate [K/s], static face weights K and momenta m,
        and the frame (heater.py:33-65).

        Returns (cell int32 (n,n,n), solid bool, source bool, rate float, k_faces [3 float32],
        m_faces [3 float32], geom dict).
        """
        centre, facing, axis, up = self.frame(centre_xy)
        # local coordinates of every cell centre: s along the trough, q along the facing
        # direction, r along up; rho is the distance from the trough axis
        coords = (np.arange(n) + 0.5) * h
        rel = np.stack(np.meshgrid(coords, coords, coords, indexing="ij"), -1) - centre
        s_ax, q_fc, r_up = rel @ axis, rel @ facing, rel @ up
        rho = np.hypot(q_fc, r_up)
        cell = np.zeros((n, n, n), np.int32)
        if self.on:
            R, half = self.radius, self.length / 2
            behind = q_fc <= 0
            shell = (np.abs(s_ax) <= half) & behind & (rho >= R - h / 2) & (rho <= R + 0.006 + h / 2)
            caps = (np.abs(np.abs(s_ax) - half) <= h / 2) & behind & (rho <= R + 0.006)
            cell[shell] = SHELL
            cell[caps] = SHELL
            for rz in (-0.03, 0.03):
                rod = (np.abs(s_ax) <= half - 0.015) & (np.hypot(q_fc + 0.03, r_up - rz) <= 0.007 + h / 2)
                cell[rod] = ROD
            if self.fan > 0:
                slab = (np.abs(s_ax) <= half - 0.02) & (np.abs(r_up) <= R - 0.01) & (q_fc > 0) & (q_fc <= 0.015)
                cell[(cell == AIR) & slab] = FAN
            trough = (np.abs(s_ax) <= half - 0.015) & (rho <= R - 0.004) & behind
            cell[(cell == AIR) & trough] = TROUGH
        solid = (cell == SHELL) | (cell == ROD)
        source = (cell == FAN) | (cell == TROUGH)
        rate = self.power / (rho_air * CP_AIR * max(int(source.sum()), 1) * h ** 3) if self.on else 0.0
        k_faces, m_faces = [], []
        for a in range(3):
            k_a, m_a = face_weights(solid, cell == FAN, a, self.fan, facing[a])
            k_faces.append(k_a)
            m_faces.append(m_a)
        geom = dict(centre=centre.tolist(), facing=facing.tolist(), axis=axis.tolist(), up=up.tolist())
        return cell, solid, source, rate, k_faces, m_faces, geom


def face_weights(solid: np.ndarray, fan: np.ndarray, axis: int, fan_speed: float, facing_a: float):
    """Static relaxation weight K and momentum m on the faces normal to `axis` (heater.py:53-63).

    A face touching a solid cell on either side gets K_SOLID and no momentum; otherwise a face
    touching a fan cell gets K_FAN and momentum K_FAN * fan_speed * facing_a, facing_a being the
    facing component along `axis`.
    """
    pad = [(0, 0)] * 3
    pad[axis] = (1, 1)
    sol, fn = np.pad(solid, pad), np.pad(fan, pad)
    n_ax = solid.shape[axis]
    minus = [slice(None)] * 3
    plus = [slice(None)] * 3
    minus[axis], plus[axis] = slice(0, n_ax + 1), slice(1, n_ax + 2)
    on_solid = sol[tuple(minus)] | sol[tuple(plus)]
    on_fan = (fn[tuple(minus)] | fn[tuple(plus)]) & ~on_solid
    k = np.where(on_solid, K_SOLID, np.where(on_fan, K_FAN, 0.0)).astype(np.float32)
    m = np.where(on_fan, K_FAN * fan_speed * facing_a, 0.0).astype(np.float32)
    return k, m
