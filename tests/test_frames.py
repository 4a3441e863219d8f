"""Run output (heatflow_mpm.frames): the reference's static.npz / meta.json / frames layout
(scripts/run_scene.py:33-45, solver.py:98-110) and the run loop's frame and substep counts.

The CPU tests drive HeatflowMPM's own static_dict/frame_dict through a stand-in object that
answers the accessors from host arrays; the GPU test runs a real scene."""
import json
import os
from types import SimpleNamespace

import numpy as np
import pytest

from heatflow_mpm import GridConfig, Heater, HeatflowMPM, Material, ellipsoid_particles
from heatflow_mpm.frames import FrameWriter, run_scene
from heatflow_mpm.heater import Heater as _H

STATIC_KEYS = {"L", "h", "n", "n_particles", "spacing", "t_amb", "cell_type", "heater_on", "heater_centre",
               "heater_facing", "heater_tilt", "heater_radius", "heater_length", "mat_rho", "mat_t_sol", "mat_t_liq",
               "mat_yield_solid", "mat_viscosity_solid", "mat_viscosity_liquid", "mat_kappa", "mat_ground_friction",
               "mat_voxel_mult", "mat_max_iterations", "mat_tolerance"}


class _HostScene:
    """Answers what HeatflowMPM's record methods read, from host arrays."""

    def __init__(self, n, P):
        self.T = np.linspace(20, 30, n ** 3, dtype=np.float32).reshape(n, n, n)
        self.parts = [np.linspace(20, 40, P, dtype=np.float32), np.linspace(0, 1, P, dtype=np.float32),
                      np.zeros(P, np.float32), np.zeros(P, np.float32)]

    def particles(self):
        return [a.copy() for a in self.parts]

    def get_field(self, name):
        assert name == "T"
        return self.T.copy()


class _FakeSim:
    """HeatflowMPM's record methods bound to host data; substep() advances T and counts calls."""
    static_dict = HeatflowMPM.static_dict
    frame_dict = HeatflowMPM.frame_dict
    positions = HeatflowMPM.positions
    velocities = HeatflowMPM.velocities
    temperatures = HeatflowMPM.temperatures
    melt_fraction = HeatflowMPM.melt_fraction
    grid_T = HeatflowMPM.grid_T

    def __init__(self, n=8, P=50):
        from heatflow_mpm.mpm import StaticBody
        L = 0.16
        self.grid = SimpleNamespace(n=n, h=L / n, L=L, cfg=SimpleNamespace(t_amb=25.0),
                                    cell_velocity=lambda: np.full((n ** 3, 3), 0.25, np.float32))
        self.heater, self.mat, self.spacing, self.n = _H(), Material(), 2.5e-3, P
        cell, *_, geom = self.heater.masks(n, L / n, (L / 2, L / 2), 1.2)
        self.cell_type, self.heater_geom = cell, geom
        rng = np.random.default_rng(0)
        self.body = StaticBody(rng.uniform(0.05, 0.1, (P, 3)).astype(np.float32)) if P else None
        self._scene = _HostScene(n, P)
        self.substeps = 0

    def air_velocity(self):
        return self.grid.cell_velocity()

    def substep(self, dt):
        self.substeps += 1
        self._scene.T += 1.0


def test_static_meta_and_frame_layout(tmp_path):
    sim = _FakeSim()
    out = str(tmp_path / "run")
    with FrameWriter(out, sim, meta=dict(seconds=1.0, power=400.0)) as fw:
        fw.write(0, 0.0)
    st = np.load(os.path.join(out, "static.npz"))
    assert set(st.files) == STATIC_KEYS
    assert st["cell_type"].dtype == np.uint8 and st["cell_type"].shape == (8, 8, 8)
    assert int(st["n"]) == 8 and int(st["n_particles"]) == 50
    assert st["heater_centre"].dtype == np.float32 and st["heater_centre"].shape == (3,)
    meta = json.load(open(os.path.join(out, "meta.json")))
    assert meta["power"] == 400.0 and meta["n_particles"] == 50 and len(meta["heater"]["centre"]) == 3
    d = np.load(os.path.join(out, "frames", "frame_0000.npz"))
    assert set(d.files) == {"t", "pos", "vel", "temp", "melt", "grid_T", "air_u"}
    assert d["t"].dtype == np.float32 and d["pos"].dtype == np.float32 and d["vel"].dtype == np.float32
    for k in ("temp", "melt", "grid_T", "air_u"):
        assert d[k].dtype == np.float16, k
    assert d["grid_T"].shape == (512,) and d["air_u"].shape == (512, 3)
    assert d["pos"].shape == (50, 3) and d["temp"].shape == (50,)
    np.testing.assert_array_equal(d["temp"], sim._scene.parts[0].astype(np.float16))
    np.testing.assert_array_equal(d["grid_T"], sim._scene.T.reshape(-1).astype(np.float16))


def test_no_particles_and_no_air(tmp_path):
    sim = _FakeSim(P=0)
    out = str(tmp_path / "run")
    with FrameWriter(out, sim, with_air=False) as fw:
        fw.write(0, 0.0)
    d = np.load(os.path.join(out, "frames", "frame_0000.npz"))
    assert "air_u" not in d.files
    assert d["pos"].shape == (0, 3) and d["temp"].shape == (0,)


def test_run_loop_counts_frames_and_substeps(tmp_path):
    sim = _FakeSim()
    out = str(tmp_path / "run")
    lines = []
    frames = run_scene(sim, out, seconds=0.1, frame_dt=0.02, dt=2e-3, log=lines.append, log_every=2)
    assert frames == 6 and sim.substeps == 50                     # 5 frames of 10 substeps, plus frame 0
    names = sorted(os.listdir(os.path.join(out, "frames")))
    assert names == [f"frame_{k:04d}.npz" for k in range(6)]
    for k in range(6):
        d = np.load(os.path.join(out, "frames", names[k]))
        assert float(d["t"]) == pytest.approx(0.02 * k)
        # T advanced by one per substep: frame k was read after 10k substeps, before later ones
        assert float(d["grid_T"].astype(np.float32)[0]) == pytest.approx(20.0 + 10 * k, abs=0.05)
    assert lines[0].startswith("50 particles, grid 8^3") and len(lines) == 1 + 3   # frames 2, 4 and the last


def test_writer_reports_disk_errors(tmp_path):
    sim = _FakeSim()
    out = str(tmp_path / "run")
    fw = FrameWriter(out, sim)
    os.rename(os.path.join(out, "frames"), os.path.join(out, "moved"))   # the next save has nowhere to go
    fw.write(0, 0.0)
    with pytest.raises(RuntimeError, match="writing a frame"):
        fw.close()


@pytest.mark.gpu
def test_frames_of_a_gpu_run_match_the_scene(gpu, tmp_path):
    P = ellipsoid_particles(semi_axes=(0.03, 0.03, 0.03), spacing=4e-3, L=0.32)
    sim = HeatflowMPM(P, Material(t_sol=22.0, t_liq=28.0), Heater(power=400.0, t_cap=120.0, fan=1.0),
                      GridConfig(L=0.32, h=0.01, cg_tol=1e-5), spacing=4e-3, t_init=20.0)
    out = str(tmp_path / "run")
    frames = run_scene(sim, out, seconds=0.06, frame_dt=0.02, dt=2e-3, log=None)
    assert frames == 4 and sim.stats()["substeps"] == 30
    d = np.load(os.path.join(out, "frames", "frame_0003.npz"))
    np.testing.assert_array_equal(d["grid_T"], sim.grid_T().reshape(-1).astype(np.float16))
    np.testing.assert_array_equal(d["temp"], sim.temperatures().astype(np.float16))
    np.testing.assert_array_equal(d["melt"], sim.melt_fraction().astype(np.float16))
    np.testing.assert_array_equal(d["air_u"], sim.air_velocity().astype(np.float16))
    np.testing.assert_array_equal(d["pos"], P)
    d0 = np.load(os.path.join(out, "frames", "frame_0000.npz"))
    assert np.all(d0["temp"] == np.float16(20.0)) and float(np.abs(d0["air_u"]).max()) == 0.0
    st = np.load(os.path.join(out, "static.npz"))
    assert np.array_equal(st["cell_type"], sim.cell_type.astype(np.uint8))
    sim.close()
