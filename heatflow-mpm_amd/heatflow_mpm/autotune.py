"""Autotuner of the pressure PCG layout: LDS tile depth of pass A and the CG form.

    python -m heatflow_mpm.autotune [--sizes 32 48 64 96 128 192] [--out path]

For every grid size and for two contexts -- a scene stepped alone, and eight scenes of that size
stepped as one batch -- each candidate layout runs the coupled scene (static ellipsoid, 400 W
heater with fan, PCG tolerance 1e-6) for a few substeps after a warmup, and is scored by wall
time per pressure-PCG iteration (launch gaps included, so the one-launch fused form is credited
for its fewer launches).  Candidates: pass A tiles of 16, 8 and 4 planes (those whose tile count
fits the pass-A workgroup cap), each with the two-pass PCG and with the one-launch iteration, the
persistent form of the one-launch iteration where a scene's tiles fit one per CU (a batch packs such
scenes into cooperative launches), and the untiled two-pass form.  The winners go to a JSON table (shipped as
tuning_mi355x.json); HeatflowMPM(layout="auto" / "auto-batch") reads it, using the entry of the
nearest measured size.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import time

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_mi355x.json")
_cache: dict = {}


def load_table(path: str | None = None) -> dict:
    p = path or TABLE
    if p not in _cache:
        if not os.path.exists(p):
            _cache[p] = {}
        else:
            with open(p) as f:
                _cache[p] = json.load(f)
    return _cache[p]


def lookup(n: int, context: str = "solo", table: dict | None = None) -> dict | None:
    """Layout for an n^3 scene: the entry of the measured size nearest to n (in log n), or None."""
    t = (table if table is not None else load_table()).get(context, {})
    if not t:
        return None
    best = min(t, key=lambda k: abs(math.log(int(k)) - math.log(n)))
    return dict(t[best]["choice"])


def candidates(n: int, context: str = "solo", cus: int = 256):
    tj = -(-n // 16)
    out = []
    for d in (16, 8, 4):
        tiles = tj * tj * -(-n // d)
        if tiles <= 4096:   # the pass-A workgroup cap of the library
            out.append(dict(dir_planes=d, cg="warp"))
            out.append(dict(dir_planes=d, cg="fused"))
            if tiles <= cus:   # one resident workgroup per tile (a batch packs scenes into such launches)
                out.append(dict(dir_planes=d, cg="persistent"))
    out.append(dict(dir_planes=-1, cg="warp"))
    return out


def probe(n: int, choice: dict, scenes: int = 1, warmup: int = 3, substeps: int = 4) -> float:
    """Wall milliseconds per pressure-PCG iteration of `scenes` coupled n^3 scenes."""
    from . import HeatflowBatch, HeatflowMPM, coupled_scene
    kw, dt, _ = coupled_scene(n)
    sims = [HeatflowMPM(**kw, layout=choice) for _ in range(scenes)]
    batch = HeatflowBatch(sims) if scenes > 1 else None
    step = (lambda: batch.substep(dt)) if batch else (lambda: sims[0].substep(dt))
    sync = (lambda: batch.synchronize()) if batch else (lambda: sims[0].synchronize())
    try:
        for _ in range(warmup):
            step()
        sync()
        it0 = sum(s.stats()["p_iters_total"] for s in sims)
        t0 = time.perf_counter()
        for _ in range(substeps):
            step()
        sync()
        wall = time.perf_counter() - t0
        its = sum(s.stats()["p_iters_total"] for s in sims) - it0
    finally:
        if batch:
            batch.close()
        for s in sims:
            s.close()
    return wall * 1e3 / max(its / scenes, 1)


def tune(sizes=(32, 48, 64, 96, 128, 192), batch_scenes: int = 8, log=print) -> dict:
    table = {"device": None, "solo": {}, "batch": {}}
    try:
        import torch
        table["device"] = torch.cuda.get_device_name(0)
    except Exception:
        pass
    for ctx, scenes in (("solo", 1), ("batch", batch_scenes)):
        for n in sizes:
            scores = []
            for c in candidates(n, ctx):
                ms = probe(n, c, scenes=scenes)
                scores.append((ms, c))
                if log:
                    log(f"{ctx:5s} n={n:3d} {c} {ms * 1e3:8.2f} us/iteration")
            ms, best = min(scores, key=lambda x: x[0])
            table[ctx][str(n)] = dict(choice=best, us_per_iteration=round(ms * 1e3, 2),
                                      candidates=[dict(c, us_per_iteration=round(m * 1e3, 2)) for m, c in scores])
    return table


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="autotune the pressure PCG layout")
    ap.add_argument("--sizes", type=int, nargs="+", default=[32, 48, 64, 96, 128, 192])
    ap.add_argument("--batch-scenes", type=int, default=8)
    ap.add_argument("--out", default=TABLE)
    a = ap.parse_args(argv)
    table = tune(tuple(a.sizes), a.batch_scenes, log=lambda s: print(s, flush=True))
    with open(a.out, "w") as f:
        json.dump(table, f, indent=1)
    print("wrote", a.out)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
