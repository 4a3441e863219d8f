"""Temperature on the grid: host view of the heat part of a scene (heat.py:177-210)."""
from __future__ import annotations

import numpy as np

from .grid import DeviceField


class HeatField:
    """Cell temperature T with its fixed cells and capped source, held by libhfm.so."""

    def __init__(self, n: int, h: float, scene):
        self.n, self.h, self._scene = n, h, scene
        self.rate, self.t_cap = 0.0, 0.0
        for name, attr in (("T", "T"), ("T2", "T2"), ("TW", "Tw"), ("KAPPA", "kap"), ("ALPHA_P", "alpha"),
                           ("TSUM", "tsum"), ("HRHS", "rhs"), ("HDIAG", "diag")):
            setattr(self, attr, DeviceField(scene, name))

    @property
    def cg_iters(self) -> int:
        """Heat PCG iterations as warp.optim.linear.cg reports them (heat.py:210)."""
        return int(self._scene.stats()["h_iters"])

    def set_fixed(self, fixed: np.ndarray, values: np.ndarray) -> None:
        """Dirichlet cells; they are applied to T at once (heat.py:189-192)."""
        self._scene.set_fixed(fixed, values)

    def set_source(self, mask: np.ndarray, rate: float, t_cap: float) -> None:
        """Capped volumetric source on the mask cells (heat.py:194-196)."""
        self.rate, self.t_cap = float(rate), float(t_cap)
        self._scene.set_source(mask, rate, t_cap)
