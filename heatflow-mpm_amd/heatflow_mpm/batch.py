"""HeatflowBatch: many HeatflowMPM simulations stepped together on one GPU.

No counterpart in the reference, which steps one HeatflowMPM at a time (solver.py:63-86).
A batch runs every stage of the substep as one kernel launch over all of its scenes
(hfm_batch_* in include/hfm.h), so grids too small to fill the GPU alone fill it together.
Scenes may differ in size, particles, heater and solver settings; each keeps its own PCG solves
and its results equal those of stepping it alone.
"""
from __future__ import annotations

import ctypes as C

from . import _abi


class HeatflowBatch:
    def __init__(self, sims):
        self.sims = list(sims)
        if not self.sims:
            raise ValueError("a batch needs at least one simulation")
        self.lib = _abi.load()
        arr = (C.c_void_p * len(self.sims))(*[s._scene.handle.value for s in self.sims])
        handle = C.c_void_p()
        _abi.check(self.lib.hfm_batch_create(arr, len(self.sims), C.byref(handle)), "hfm_batch_create")
        self.handle = handle

    def __len__(self) -> int:
        return len(self.sims)

    def set_sims(self, sims) -> None:
        """Replace the simulations the batch steps (possibly none), keeping its stream and buffers."""
        sims = list(sims)
        arr = (C.c_void_p * max(len(sims), 1))(*[s._scene.handle.value for s in sims])
        _abi.check(self.lib.hfm_batch_set_scenes(self.handle, arr, len(sims)), "hfm_batch_set_scenes")
        self.sims = sims

    def substep(self, dt: float) -> None:
        """One substep of every simulation (solver.py:63-88 for each), moving bodies included."""
        for s in self.sims:
            b = s.body
            if b is not None and not getattr(b, "static", False):
                s._scene.update_particles(b.pos, b.vel)
        _abi.check(self.lib.hfm_batch_substep(self.handle, float(dt)), "hfm_batch_substep")
        for s in self.sims:
            if s.body is not None:
                s.body.step(dt)

    def run(self, dt: float, substeps: int) -> float:
        """`substeps` substeps of static scenes back to back; returns the GPU time in ms."""
        if any(s.body is not None and not getattr(s.body, "static", False) for s in self.sims):
            raise ValueError("run() needs static bodies; call substep() for moving ones")
        ms = C.c_double(0.0)
        _abi.check(self.lib.hfm_batch_run(self.handle, float(dt), int(substeps), C.byref(ms)), "hfm_batch_run")
        return ms.value

    def synchronize(self) -> None:
        _abi.check(self.lib.hfm_batch_synchronize(self.handle), "hfm_batch_synchronize")

    def set_kernel_timing(self, on: bool) -> None:
        _abi.check(self.lib.hfm_batch_set_kernel_timing(self.handle, int(bool(on))), "hfm_batch_set_kernel_timing")

    def kernel_time(self, cls: int):
        ms, n = C.c_double(0.0), C.c_int64(0)
        _abi.check(self.lib.hfm_batch_kernel_time(self.handle, int(cls), C.byref(ms), C.byref(n)),
                   "hfm_batch_kernel_time")
        return ms.value, n.value

    def kernel_time_reset(self) -> None:
        _abi.check(self.lib.hfm_batch_kernel_time_reset(self.handle), "hfm_batch_kernel_time_reset")

    def close(self) -> None:
        if getattr(self, "handle", None):
            _abi.check(self.lib.hfm_batch_destroy(self.handle), "hfm_batch_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
