"""Heat a mesh of material in front of the heater and write frames (scripts/run_scene.py of the reference).

    python -m heatflow_mpm.run_scene --out runs/bunny_fan --mesh path/to/bunny.obj --fan 2.0 --power 400
    python -m heatflow_mpm.run_scene --out runs/ellipsoid_fan --ellipsoid --fan 2.0

Same arguments and output as the reference script.  The mesh is seeded by mesh_particles (jittered
lattice, winding-number inside test on the GPU, shapes.py:47-51); no mesh ships with this library,
so --mesh names a file (default: the HFM_BUNNY environment variable, or the reference's
assets/bunny.obj copied next to the package).  --ellipsoid seeds an ellipsoid of --semi-axes on the
same lattice instead.  The particles are held static (Newton's MPM step is outside this library).
"""
from __future__ import annotations

import argparse
import sys

from . import BUNNY, GridConfig, Heater, HeatflowMPM, Material, ellipsoid_particles, mesh_particles
from .frames import run_scene


def parser() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", required=True)
    ap.add_argument("--seconds", type=float, default=20.0)
    ap.add_argument("--frame-dt", type=float, default=0.02)
    ap.add_argument("--dt", type=float, default=2e-3)
    ap.add_argument("--mesh", default=BUNNY)
    ap.add_argument("--height", type=float, default=0.10)
    ap.add_argument("--yaw", type=float, default=0.0)
    ap.add_argument("--ellipsoid", action="store_true", help="seed an ellipsoid of --semi-axes instead of the mesh")
    ap.add_argument("--semi-axes", type=float, nargs=3, default=(0.05, 0.04, 0.05))
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--spacing", type=float, default=2.5e-3)
    ap.add_argument("--L", type=float, default=0.48)
    ap.add_argument("--h", type=float, default=0.005)
    ap.add_argument("--t-amb", type=float, default=25.0)
    ap.add_argument("--t-init", type=float, default=20.0)
    ap.add_argument("--power", type=float, default=400.0)
    ap.add_argument("--t-cap", type=float, default=120.0)
    ap.add_argument("--t-rod", type=float, default=300.0)
    ap.add_argument("--fan", type=float, default=0.0)
    ap.add_argument("--heater-dist", type=float, default=0.21)
    ap.add_argument("--heater-z", type=float, default=0.12)
    ap.add_argument("--no-heater", action="store_true")
    ap.add_argument("--t-sol", type=float, default=30.0)
    ap.add_argument("--t-liq", type=float, default=36.0)
    ap.add_argument("--yield-solid", type=float, default=2e4)
    ap.add_argument("--viscosity-solid", type=float, default=5.0)
    ap.add_argument("--viscosity-liquid", type=float, default=30.0)
    ap.add_argument("--kappa", type=float, default=2e-5)
    ap.add_argument("--cg-tol", type=float, default=1e-4)
    ap.add_argument("--no-air-frames", action="store_true")
    ap.add_argument("--device", default="cuda:0")
    return ap


def build(a) -> HeatflowMPM:
    mat = Material(t_sol=a.t_sol, t_liq=a.t_liq, yield_solid=a.yield_solid, viscosity_solid=a.viscosity_solid,
                   viscosity_liquid=a.viscosity_liquid, kappa=a.kappa)
    heater = Heater(dist=a.heater_dist, z=a.heater_z, t_rod=a.t_rod, power=a.power, t_cap=a.t_cap, fan=a.fan,
                    on=not a.no_heater)
    if a.ellipsoid:
        P = ellipsoid_particles(semi_axes=tuple(a.semi_axes), spacing=a.spacing, jitter=0.2 * a.spacing, seed=a.seed,
                                L=a.L)
    else:
        P = mesh_particles(a.mesh, a.height, a.spacing, a.L, a.seed, a.yaw, a.device)
    grid = GridConfig(L=a.L, h=a.h, t_amb=a.t_amb, cg_tol=a.cg_tol)
    return HeatflowMPM(P, mat, heater, grid, a.spacing, a.t_init, device=a.device)


def main(argv=None) -> int:
    a = parser().parse_args(argv)
    sim = build(a)
    try:
        run_scene(sim, a.out, a.seconds, a.frame_dt, a.dt, not a.no_air_frames, meta=vars(a),
                  log=lambda s: print(s, flush=True))
    finally:
        sim.close()
    print("done", a.out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
