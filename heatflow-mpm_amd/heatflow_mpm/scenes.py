"""Synthetic scenes: seeded particle blobs and the workloads the path is measured on.

The reference seeds particles from a mesh (shapes.py:34-51: jittered lattice, inside test,
then placed on the floor at the box centre).  The workloads here use an ellipsoid instead of a
mesh, seeded the same way: lattice over the bounding box grown by one spacing, uniform jitter,
inside test, then shifted to stand on the floor at the box centre.
"""
from __future__ import annotations

import numpy as np

from .grid import GridConfig
from .heater import Heater
from .mpm import Material


def ellipsoid_particles(semi_axes=(0.05, 0.04, 0.05), spacing: float = 2.5e-3, jitter: float = 0.3e-3,
                        seed: int = 0, L: float = 0.48) -> np.ndarray:
    """Jittered-lattice particles filling an ellipsoid standing on the floor at the box centre.

    `jitter` is the absolute half-width [m] of the uniform jitter per axis.  The lattice covers
    [-a - spacing, a + spacing] per axis at spacing/2 + k*spacing like shapes.py:35-37; the
    jitter is drawn for every lattice point before the inside test (shapes.py:38).
    """
    a = np.asarray(semi_axes, float)
    lo, hi = -a - spacing, a + spacing
    axes = [np.arange(lo[d] + spacing / 2, hi[d], spacing) for d in range(3)]
    P = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3)
    P = P + np.random.default_rng(seed).uniform(-jitter, jitter, P.shape)
    inside = ((P / a) ** 2).sum(1) < 1.0
    P = P[inside]
    # lowest point of the ellipsoid at z = 0, centred in xy, then lifted by half a spacing (shapes.py:51)
    return (P + np.array([L / 2, L / 2, a[2] + spacing / 2])).astype(np.float32)


def plume_scene():
    """Heater plume with no material: n=48 (h = 10 mm), 400 W capped at 120 C, fan off, pressure
    and heat PCG to 1e-6.  Runs 200 substeps of 2 ms."""
    return dict(particles=np.zeros((0, 3), np.float32), mat=Material(),
                heater=Heater(power=400.0, t_cap=120.0, t_rod=300.0, fan=0.0),
                grid=GridConfig(L=0.48, h=0.01, cg_tol=1e-6), spacing=2.5e-3, t_init=25.0), 2e-3, 200


def coupled_scene(n: int = 96, particles_target: int | None = None, seed: int = 0):
    """Default coupled scene: box 0.48 m at n cells, a static ellipsoid (5 x 4 x 5 cm semi-axes,
    2.5 mm lattice, +-0.3 mm jitter) at 20 C in front of a 400 W heater (cap 120 C, fan 2 m/s,
    0.21 m away), melt window 28-34 C, pressure and heat PCG to 1e-6.  Runs 500 substeps of 2 ms.

    With `particles_target`, the lattice spacing is scaled so the ellipsoid holds about that
    many particles (spacing = 2.5 mm * (count at 2.5 mm / target)^(1/3)).
    """
    spacing = 2.5e-3
    if particles_target:
        base = len(ellipsoid_particles(spacing=2.5e-3, seed=seed))
        spacing = 2.5e-3 * (base / particles_target) ** (1.0 / 3.0)
    P = ellipsoid_particles(spacing=spacing, jitter=0.3e-3, seed=seed)
    return dict(particles=P, mat=Material(t_sol=28.0, t_liq=34.0),
                heater=Heater(power=400.0, t_cap=120.0, fan=2.0, dist=0.21),
                grid=GridConfig(L=0.48, h=0.48 / n, cg_tol=1e-6), spacing=spacing, t_init=20.0), 2e-3, 500


def random_scene(rng, n: int, name: str = "scene", cg_max_iter: int = 400):
    """One scene of the batched and served workloads at n cells per side (0.48 m box): heater
    power U(100, 800) W, fan U(0, 3) m/s, an ellipsoid of U{2000..50000} particles with semi-axes
    U(3, 7) cm each (lattice spacing (4/3 pi a b c / P)^(1/3), jitter 12 % of it) at 20 C; melt
    window 28-34 C, PCG tolerance 1e-6 with at most cg_max_iter pressure iterations (400, the
    reference's default).  Draw order: power, fan, P, semi-axes, particle seed."""
    power = float(rng.uniform(100.0, 800.0))
    fan = float(rng.uniform(0.0, 3.0))
    target = int(rng.integers(2000, 50001))
    axes = tuple(float(x) for x in rng.uniform(0.03, 0.07, 3))
    pseed = int(rng.integers(0, 2 ** 31))
    spacing = (4.0 / 3.0 * np.pi * axes[0] * axes[1] * axes[2] / target) ** (1.0 / 3.0)
    P = ellipsoid_particles(semi_axes=axes, spacing=spacing, jitter=0.12 * spacing, seed=pseed)
    kw = dict(particles=P, mat=Material(t_sol=28.0, t_liq=34.0),
              heater=Heater(power=power, t_cap=120.0, fan=fan, dist=0.21),
              grid=GridConfig(L=0.48, h=0.48 / n, cg_tol=1e-6, cg_max_iter=cg_max_iter), spacing=spacing,
              t_init=20.0)
    return f"{name}_n{n}_P{len(P)}", kw


def batched_scenes(count: int = 64, seed: int = 0):
    """Independent scenes of uneven size for batched stepping: n drawn from {32, 48, 64, 96, 128}
    (0.48 m box), heater power U(100, 800) W, fan U(0, 3) m/s, an ellipsoid of U{2000..50000}
    particles with semi-axes U(3, 7) cm each (lattice spacing derived from the count:
    (4/3 pi a b c / P)^(1/3), jitter 12 % of it) at 20 C; melt window 28-34 C, PCG tolerance 1e-6.
    Runs 200 substeps of 2 ms.  Draw order per scene: n, power, fan, P, semi-axes, then the
    particle seed.  Returns a list of (name, kwargs)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.choice([32, 48, 64, 96, 128]))
        out.append(random_scene(rng, n, name=f"s{i:02d}"))
    return out, 2e-3, 200
