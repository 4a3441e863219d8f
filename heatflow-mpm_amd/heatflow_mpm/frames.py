"""Run output in the reference's on-disk format, written while the GPU keeps stepping.

A run directory holds (scripts/run_scene.py:3-4, 33-45):
  static.npz              HeatflowMPM.static_dict(): L, h, n, n_particles, spacing, t_amb, cell_type (u8),
                          heater_on/centre/facing/tilt/radius/length and mat_* (solver.py:98-103)
  meta.json               the run's arguments plus n_particles and the heater frame
  frames/frame_NNNN.npz   HeatflowMPM.frame_dict(t): t, pos, vel (f32), temp, melt, grid_T (n^3),
                          air_u (n^3 x 3) as f16 (solver.py:105-110); frame 0 is the initial state

The frame fields are read from the device synchronously (that is the only moment the GPU state
is consistent), but converting and writing them runs on a worker thread, so the next frame's
substeps are already on the GPU while the previous frame goes to disk.
"""
from __future__ import annotations

import json
import os
import queue
import threading
import time

import numpy as np


def _jsonable(x):
    if isinstance(x, dict):
        return {k: _jsonable(v) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        return [_jsonable(v) for v in x]
    if isinstance(x, np.ndarray):
        return x.tolist()
    if isinstance(x, np.generic):
        return x.item()
    return x


class FrameWriter:
    """Writes static.npz and meta.json at creation, then one npz per `write(k, t)`.

    `max_pending` frames may wait for the disk; a further `write` blocks until one is done, so a
    slow disk cannot pile up host memory.  Errors of the worker are raised by the next `write` or
    by `close`.
    """

    def __init__(self, out: str, sim, meta: dict | None = None, with_air: bool = True, max_pending: int = 4):
        self.out, self.sim, self.with_air = out, sim, with_air
        os.makedirs(os.path.join(out, "frames"), exist_ok=True)
        np.savez(os.path.join(out, "static.npz"), **sim.static_dict())
        info = dict(meta or {}, n_particles=int(sim.n), heater=sim.heater_geom)
        with open(os.path.join(out, "meta.json"), "w") as f:
            json.dump(_jsonable(info), f, indent=1)
        self._q: queue.Queue = queue.Queue(maxsize=max(1, int(max_pending)))
        self._error: BaseException | None = None
        self._thread = threading.Thread(target=self._work, name="frame-writer", daemon=True)
        self._thread.start()
        self.frames = 0

    def path(self, k: int) -> str:
        return os.path.join(self.out, "frames", f"frame_{k:04d}.npz")

    def _work(self):
        while True:
            item = self._q.get()
            if item is None:
                return
            k, d = item
            try:
                if self._error is None:
                    np.savez(self.path(k), **d)
            except BaseException as e:          # reported to the caller by write()/close()
                self._error = e

    def _raise(self):
        if self._error is not None:
            e, self._error = self._error, None
            raise RuntimeError(f"writing a frame under {self.out} failed: {e}") from e

    def write(self, k: int, t: float) -> None:
        """Read frame k (time t) from the device now and queue it for the disk."""
        self._raise()
        self._q.put((int(k), self.sim.frame_dict(t, self.with_air)))
        self.frames += 1

    def close(self) -> None:
        if self._thread.is_alive():
            self._q.put(None)
            self._thread.join()
        self._raise()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def run_scene(sim, out: str, seconds: float = 20.0, frame_dt: float = 0.02, dt: float = 2e-3, with_air: bool = True,
              meta: dict | None = None, log=print, log_every: int = 50) -> int:
    """The reference's run loop (scripts/run_scene.py:38-49): frame 0, then round(frame_dt/dt)
    substeps per frame for round(seconds/frame_dt) frames, each frame written as it completes.
    Every `log_every` frames (and at the end) prints the material height, the melted share, the
    air temperature above the material and the wall time per frame.  Returns the frame count."""
    sub, nf = int(round(frame_dt / dt)), int(round(seconds / frame_dt))
    if sub < 1 or nf < 0:
        raise ValueError(f"frame_dt {frame_dt} and seconds {seconds} must give at least one substep per frame")
    t0 = time.time()
    with FrameWriter(out, sim, meta, with_air) as fw:
        fw.write(0, 0.0)
        if log:
            log(f"{sim.n} particles, grid {sim.grid.n}^3")
        for k in range(1, nf + 1):
            for _ in range(sub):
                sim.substep(dt)
            fw.write(k, k * frame_dt)
            if log and (k % log_every == 0 or k == nf):
                n, h = sim.grid.n, sim.grid.h
                x, f, T = sim.positions(), sim.melt_fraction(), sim.grid_T()
                c = n // 2
                top = float(x[:, 2].max()) if len(x) else 0.0
                melted = 100.0 * float((f >= 1).mean()) if len(f) else 0.0
                above = T[c, c, min(int(top / h) + 1, n - 1)]
                log(f"t {k * frame_dt:5.1f} s | height {top * 100:4.1f} cm | melted {melted:5.1f} % | "
                    f"air above the material {above:5.1f} C | {(time.time() - t0) / k:.2f} s/frame")
    return nf + 1
