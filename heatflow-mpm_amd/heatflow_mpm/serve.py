"""Serving scene rollouts on one GPU with continuous batching.

No counterpart in the reference (which runs one HeatflowMPM to completion, run_scene.py:30-45).
Requests (a scene and a substep count) arrive over time.  The server keeps every admitted
scene's fields and solver state resident in device memory, admits waiting requests in arrival
order while their device footprint fits the budget (and a slot is free), and advances all
admitted scenes together, one batched substep at a time (HeatflowBatch): a request that arrives
while others run joins the batch at the next substep, and one that has run its substeps leaves
it with its result.  The host-side preparation of a scene (heater masks, walls, face weights)
runs on worker threads while the GPU steps the batch.

Each scene's arithmetic and reductions depend only on that scene, so a result is bit-identical
whether the request ran alone or in the mix.
"""
from __future__ import annotations

import collections
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Request:
    rid: int
    kwargs: dict              # HeatflowMPM constructor arguments
    substeps: int
    arrival: float            # seconds after the start of the run
    dt: float = 2e-3


@dataclass
class Result:
    rid: int
    n: int
    substeps: int
    arrival: float
    admitted: float
    finished: float
    device_bytes: int
    fields: dict = field(default_factory=dict)   # final T, U, V, W and particle temperatures if kept
    stats: dict = field(default_factory=dict)

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def footprint_estimate(n: int, particles: int, precision: str = "fp32", cg: str = "warp") -> int:
    """Device bytes of a scene before it is created: 21 face-sized and 29 cell-sized fp32
    arrays (fields, static parts, solver vectors, bins, contribution-list starts), 1 byte of
    flags per cell, an fp64 partial sum per 256 cells, 12 fp32 words and 4 ints per particle and
    its 8 contribution-list terms (an int and a float each), and fixed tables (partial sums with
    the persistent form's per-XCD copies, programs, about 1 MiB); the mixed form adds 5 FP16 cell
    arrays (inner solve), the multigrid
    preconditioner 6 fp32 arrays over its coarse levels (fewer than N / 7 cells)."""
    N = n ** 3
    F = (n + 1) * n * n
    extra = 2 * 5 * N if precision == "mixed" else 0
    if cg == "mg":
        extra += 24 * (N // 7 + 1)
    return 4 * (21 * F + 29 * N) + N + 8 * (N // 256 + 1) + 128 * particles + 1024 * 1024 + extra


class GpuBackend:
    """HeatflowMPM scenes stepped by HeatflowBatch on one device."""

    def __init__(self, cg: str = "warp", device: str = "cuda:0", precision: str = "fp32", layout=None):
        self.cg, self.device, self.precision, self.layout = cg, device, precision, layout

    def prepare(self, req: Request):
        from .solver import scene_inputs, KAPPA_AIR
        kw = req.kwargs
        from .mpm import Material
        from .heater import Heater
        from .grid import GridConfig
        P = np.asarray(kw["particles"], np.float32).reshape(-1, 3)
        return scene_inputs(len(P), kw.get("mat", Material()), kw.get("heater", Heater()),
                            kw.get("grid", GridConfig()), kw.get("spacing", 2.5e-3), kw.get("t_init", 20.0),
                            kw.get("kappa_air", KAPPA_AIR))

    def create(self, req: Request, prepared):
        from .solver import HeatflowMPM
        return HeatflowMPM(**req.kwargs, cg=self.cg, device=self.device, inputs=prepared, precision=self.precision,
                           layout=self.layout)

    def n_of(self, sim) -> int:
        return sim.grid.n

    def bytes(self, sim) -> int:
        return sim._scene.device_bytes()

    def make_batch(self, sims):
        from .batch import HeatflowBatch
        return HeatflowBatch(sims)

    def set_batch(self, batch, sims) -> None:
        batch.set_sims(sims)

    def substep(self, batch, dt: float) -> None:
        batch.substep(dt)

    def collect(self, sim, keep: bool):
        out = {}
        st = sim.stats()
        if keep:
            out["T"] = sim.grid_T()
            for f in ("U", "V", "W"):
                out[f] = sim._scene.get_field(f)
            out["temperatures"] = sim.temperatures().copy()
            st["divergence_max"] = sim.grid.divergence_max()
        return out, st

    def close_batch(self, batch) -> None:
        batch.close()

    def close_sim(self, sim) -> None:
        sim.close()


@dataclass
class _Slot:
    req: Request
    sim: object
    remaining: int
    admitted: float
    nbytes: int


class Server:
    """Continuous-batching request loop (see the module docstring).

    hbm_budget: device bytes the admitted scenes may hold together; max_active: most scenes in
    one batch (1 = the plain form: one request at a time, in arrival order).

    lanes: optional size classes served side by side, as [(largest n, most scenes in its batch), ...].
    A request goes to the first lane whose largest n it does not exceed; each lane runs its own
    continuous batch on its own stream from its own host thread, so a lane of small scenes keeps
    stepping while a large scene's solve runs on another (the GPU runs the lanes' kernels
    concurrently), and small requests do not wait on the slowest scene of one shared batch.  All
    lanes admit against the one device budget.  Results are the same bits as without lanes."""

    def __init__(self, backend=None, hbm_budget: int = 64 << 30, max_active: int = 64, keep_results: bool = False,
                 prepare_workers: int = 2, clock=time.perf_counter, sleep=time.sleep, lanes=None):
        self.backend = backend if backend is not None else GpuBackend()
        self.budget, self.max_active, self.keep = int(hbm_budget), int(max_active), keep_results
        self.workers, self.clock, self.sleep = prepare_workers, clock, sleep
        self.lanes = sorted((int(m), int(k)) for m, k in lanes) if lanes else None
        self.log = []   # (time, event, rid, active scenes of the lane, bytes in use)
        self._lock = threading.Lock()
        self._used = 0

    def run(self, requests) -> list:
        reqs = sorted(requests, key=lambda r: (r.arrival, r.rid))
        prec = getattr(self.backend, "precision", "fp32")
        cg = getattr(self.backend, "cg", "warp")
        dts = {r.dt for r in reqs}
        if len(dts) > 1:
            raise ValueError("all requests of one run share dt (they advance in the same batched substeps)")
        for r in reqs:
            if footprint_estimate(r.kwargs["grid"].n, len(r.kwargs["particles"]), prec, cg) > self.budget:
                raise ValueError(f"request {r.rid} does not fit the device budget on its own")
        dt = dts.pop() if dts else 2e-3
        self._used = 0
        pool = ThreadPoolExecutor(max_workers=max(1, self.workers))
        t0 = self.clock()
        try:
            if not self.lanes:
                results = self._serve(reqs, self.max_active, dt, t0, pool, prec, cg)
            else:
                parts = [[] for _ in self.lanes]
                for r in reqs:
                    n = r.kwargs["grid"].n
                    lane = next((i for i, (m, _) in enumerate(self.lanes) if n <= m), None)
                    if lane is None:
                        raise ValueError(f"request {r.rid} (n = {n}) is larger than every lane")
                    parts[lane].append(r)
                with ThreadPoolExecutor(max_workers=len(self.lanes)) as lanes:
                    futs = [lanes.submit(self._serve, part, k, dt, t0, pool, prec, cg)
                            for part, (_, k) in zip(parts, self.lanes) if part]
                    results = [x for f in futs for x in f.result()]
        finally:
            pool.shutdown(wait=True)
        return sorted(results, key=lambda r: r.rid)

    def _serve(self, reqs, max_active: int, dt: float, t0: float, pool, prec: str, cg: str) -> list:
        """One continuous batch over `reqs` (in arrival order), admitting against the shared budget."""
        pending = collections.deque(reqs)
        waiting = collections.deque()
        prepared = {}
        active, results = [], []
        batch, dirty = None, False
        try:
            while pending or waiting or active:
                now = self.clock() - t0
                while pending and pending[0].arrival <= now:
                    r = pending.popleft()
                    waiting.append(r)
                    prepared[r.rid] = pool.submit(self.backend.prepare, r)
                # admission: arrival order, while the head of the queue fits
                while waiting and len(active) < max_active:
                    r = waiting[0]
                    est = footprint_estimate(r.kwargs["grid"].n, len(r.kwargs["particles"]), prec, cg)
                    with self._lock:
                        if self._used + est > self.budget:
                            break
                        self._used += est
                    waiting.popleft()
                    sim = self.backend.create(r, prepared.pop(r.rid).result())
                    nbytes = self.backend.bytes(sim)
                    with self._lock:
                        self._used += nbytes - est
                        used = self._used
                    active.append(_Slot(r, sim, r.substeps, self.clock() - t0, nbytes))
                    dirty = True
                    self.log.append((self.clock() - t0, "admit", r.rid, len(active), used))
                if not active:
                    wait = (pending[0].arrival if pending else now + 0.01) - (self.clock() - t0)
                    if wait > 0:
                        self.sleep(min(wait, 0.01))
                    continue
                if dirty:
                    if batch is None:
                        batch = self.backend.make_batch([s.sim for s in active])
                    else:
                        self.backend.set_batch(batch, [s.sim for s in active])
                    dirty = False
                self.backend.substep(batch, dt)
                for s in active:
                    s.remaining -= 1
                done = [s for s in active if s.remaining <= 0]
                if done:
                    self.backend.set_batch(batch, [s.sim for s in active if s.remaining > 0])
                    dirty = True
                    fin = self.clock() - t0
                    for s in done:
                        fields, stats = self.backend.collect(s.sim, self.keep)
                        results.append(Result(s.req.rid, self.backend.n_of(s.sim), s.req.substeps, s.req.arrival,
                                              s.admitted, fin, s.nbytes, fields, stats))
                        self.backend.close_sim(s.sim)
                        with self._lock:
                            self._used -= s.nbytes
                            used = self._used
                        self.log.append((fin, "finish", s.req.rid, len(active) - len(done), used))
                    active = [s for s in active if s.remaining > 0]
        finally:
            if batch is not None:
                self.backend.close_batch(batch)
            for s in active:
                self.backend.close_sim(s.sim)
        return results


def poisson_requests(rate: float, count: int, seed: int = 0, sizes=(48, 64, 96, 128, 192), k_range=(50, 1000),
                     cg_max_iter: int = 400):
    """Seeded Poisson arrivals at `rate` per second; each request draws n from `sizes`, K
    uniformly from k_range, and heater / fan / particle-blob parameters as the batched scenes
    (heatflow_mpm.scenes.batched_scenes), with at most cg_max_iter pressure iterations.  Draw
    order: inter-arrival gap, n, K, then the scene."""
    from .scenes import random_scene
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for i in range(count):
        t += float(rng.exponential(1.0 / rate))
        n = int(rng.choice(list(sizes)))
        K = int(rng.integers(k_range[0], k_range[1] + 1))
        _, kw = random_scene(rng, n, name=f"r{i:03d}", cg_max_iter=cg_max_iter)
        out.append(Request(i, kw, K, t))
    return out


def summarize(results, makespan: float | None = None) -> dict:
    lat = np.array([r.latency for r in results])
    work = sum(r.n ** 3 * r.substeps for r in results)
    span = makespan if makespan is not None else max(r.finished for r in results) - min(r.arrival for r in results)
    return dict(requests=len(results), p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                mean_s=float(lat.mean()), cell_substeps=int(work), makespan_s=float(span),
                cell_substeps_per_s=float(work / span) if span > 0 else None)
