"""HeatflowMPM: the air/heat substep of Heatflow-MPM on an MI355X.

Drop-in for heatflow_mpm.solver.HeatflowMPM of the reference (solver.py:41-110): same
constructor, substep(dt) and accessors.  The substep itself runs in libhfm.so (hand-written
HIP kernels for gfx950); this module only prepares the scene on the host (heater masks, walls,
fixed cells, parameters) and moves data across the C ABI.  The MPM constitutive solve stays
outside: the particles come from a body object (StaticBody by default) and the substep hands
back temperatures, melt fractions, yield stresses and viscosities.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import asdict

import numpy as np

from . import _abi
from .grid import AirGrid, GridConfig, lattice_shape
from .heat import HeatField
from .heater import ROD, Heater
from .mpm import Material, StaticBody

KAPPA_AIR = 2.2e-5     # thermal diffusivity of air [m^2/s] (solver.py:15)
K_PARTICLE = 1.0e-3    # face relaxation weight per particle [kg/s] (solver.py:16)
GRAVITY = 9.81         # grid.py:219
PRESSURE_CHECK_EVERY = 20   # grid.py:248
HEAT_CHECK_EVERY = 10       # heat.py:209


def scene_inputs(n_particles: int, mat: Material, heater: Heater, grid: GridConfig, spacing: float, t_init: float,
                 kappa_air: float) -> dict:
    """Everything HeatflowMPM.__init__ derives on the host (solver.py:42-61, heater.py:33-65).

    Returns the scene parameters (fields of hfm_params) and the static arrays: fixed cells with
    their values (rods at t_rod over room walls at t_amb on x = 0, n-1, y = 0, n-1 and the
    ceiling; the floor stays adiabatic), the source mask and rate, the static solid fraction and
    face weights, the cell types and the heater frame.
    """
    n, h = grid.n, grid.h
    cell, solid, source, rate, k_faces, m_faces, geom = heater.masks(n, h, (grid.L / 2, grid.L / 2), grid.rho)
    rods = cell == ROD
    wall = np.zeros((n, n, n), bool)
    wall[[0, -1], :, :] = True
    wall[:, [0, -1], :] = True
    wall[:, :, -1] = True
    tfix = np.where(rods, heater.t_rod, np.where(wall, grid.t_amb, 0.0)).astype(np.float32)
    fixed = (rods | wall).astype(np.int32)
    t_max = max(heater.t_rod, heater.t_cap, grid.t_amb, t_init)
    params = dict(n=n, buoyancy=int(bool(grid.buoyancy)), cg_max_iter=int(grid.cg_max_iter),
                  cg_check_every=PRESSURE_CHECK_EVERY, heat_max_iter=int(grid.heat_max_iter),
                  heat_check_every=HEAT_CHECK_EVERY, precision=_abi.PREC_FP32, flags=0,
                  h=float(h), rho=float(grid.rho), t_amb=float(grid.t_amb), gravity=GRAVITY,
                  alpha_f_min=float(grid.alpha_f_min), cg_tol=float(grid.cg_tol), heat_tol=float(grid.heat_tol),
                  kappa_air=float(kappa_air), kappa_mat=float(mat.kappa), vol=float(spacing) ** 3,
                  k_particle=K_PARTICLE, t_sol=float(mat.t_sol), t_liq=float(mat.t_liq),
                  yield_solid=float(mat.yield_solid), mu_solid=float(mat.viscosity_solid),
                  mu_liquid=float(mat.viscosity_liquid), t_max=float(t_max), dir_planes=0)
    return dict(params=params, cell_type=cell, fixed=fixed, tfix=tfix, source=source.astype(np.int32),
                rate=float(rate), t_cap=float(heater.t_cap), solid=solid.astype(np.float32),
                k_faces=[np.ascontiguousarray(k, np.float32) for k in k_faces],
                m_faces=[np.ascontiguousarray(m, np.float32) for m in m_faces], geom=geom, t_max=t_max)


def parse_device(device) -> int:
    """'cuda:0' / 'hip:1' / 'gpu' / 0 -> HIP device ordinal.  'cpu' is rejected: no CPU path."""
    if isinstance(device, int):
        return device
    d = str(device).strip().lower()
    if d.startswith("cpu"):
        raise ValueError("libhfm.so runs on the GPU only; device='cpu' has no implementation here")
    if ":" in d:
        return int(d.split(":", 1)[1])
    return 0


class _Scene:
    """Owner of one hfm_scene handle."""

    def __init__(self, params: dict, device: int):
        self.lib = _abi.load()
        self.n = int(params["n"])
        self.params = _abi.Params(**params)
        handle = C.c_void_p()
        _abi.check(self.lib.hfm_scene_create(C.byref(self.params), int(device), C.byref(handle)), "hfm_scene_create")
        self.handle = handle
        self.n_particles = 0

    def close(self) -> None:
        if getattr(self, "handle", None):
            _abi.check(self.lib.hfm_scene_destroy(self.handle), "hfm_scene_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _f32(self, a, shape) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.float32)
        if a.shape != tuple(shape):
            a = a.reshape(shape)
        return a

    def set_static(self, alpha, k_faces, m_faces) -> None:
        n = self.n
        sh = [lattice_shape(x, n) for x in ("U", "V", "W")]
        args = [self._f32(alpha, (n, n, n))] + [self._f32(k, s) for k, s in zip(k_faces, sh)] + \
               [self._f32(m, s) for m, s in zip(m_faces, sh)]
        _abi.check(self.lib.hfm_set_static(self.handle, *args), "hfm_set_static")

    def set_fixed(self, fixed, values) -> None:
        n = self.n
        f = np.ascontiguousarray(fixed, dtype=np.int32).reshape(n, n, n)
        v = self._f32(values, (n, n, n))
        _abi.check(self.lib.hfm_set_fixed(self.handle, f, v), "hfm_set_fixed")

    def set_source(self, mask, rate, t_cap) -> None:
        m = np.ascontiguousarray(mask, dtype=np.int32).reshape(self.n, self.n, self.n)
        _abi.check(self.lib.hfm_set_source(self.handle, m, float(rate), float(t_cap)), "hfm_set_source")

    def set_particles(self, pos, vel, temp) -> None:
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, np.float32).reshape(-1, 3)
        temp = np.ascontiguousarray(temp, np.float32).reshape(-1)
        P = len(pos)
        _abi.check(self.lib.hfm_set_particles(self.handle, P, _abi.ptr(pos), _abi.ptr(vel), _abi.ptr(temp)),
                   "hfm_set_particles")
        self.n_particles = P

    def update_particles(self, pos, vel) -> None:
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, np.float32).reshape(-1, 3)
        if len(pos) != self.n_particles or len(vel) != self.n_particles:
            raise ValueError("particle count changed; use set_particles")
        _abi.check(self.lib.hfm_update_particles(self.handle, _abi.ptr(pos), _abi.ptr(vel)), "hfm_update_particles")

    def substep(self, dt: float) -> None:
        _abi.check(self.lib.hfm_substep(self.handle, float(dt)), "hfm_substep")

    def run(self, dt: float, K: int) -> float:
        ms = C.c_double(0.0)
        _abi.check(self.lib.hfm_run(self.handle, float(dt), int(K), C.byref(ms)), "hfm_run")
        return ms.value

    def get_field(self, name: str) -> np.ndarray:
        out = np.empty(lattice_shape(name, self.n), np.float32)
        _abi.check(self.lib.hfm_get_field(self.handle, _abi.FIELD[name], out), f"hfm_get_field({name})")
        return out

    def set_field(self, name: str, values) -> None:
        a = self._f32(values, lattice_shape(name, self.n))
        _abi.check(self.lib.hfm_set_field(self.handle, _abi.FIELD[name], a), f"hfm_set_field({name})")

    def cell_velocity(self) -> np.ndarray:
        out = np.empty((self.n ** 3, 3), np.float32)
        _abi.check(self.lib.hfm_get_cell_velocity(self.handle, out), "hfm_get_cell_velocity")
        return out

    def particles(self):
        P = max(self.n_particles, 0)
        outs = [np.empty(P, np.float32) for _ in range(4)]
        if P:
            _abi.check(self.lib.hfm_get_particles(self.handle, *[_abi.ptr(o) for o in outs]), "hfm_get_particles")
        return outs

    def particle_bins(self):
        """(bin (P,), base (P, 4, 3) int32, frac (P, 4, 3) float32) as the device forms them (hfm.h:
        hfm_get_particle_bins); lattices in the order cells, x, y, z faces."""
        P = max(self.n_particles, 0)
        b = np.zeros(P, np.int32)
        base = np.zeros((P, 4, 3), np.int32)
        frac = np.zeros((P, 4, 3), np.float32)
        if P:
            _abi.check(self.lib.hfm_get_particle_bins(self.handle, b, base, frac), "hfm_get_particle_bins")
        return b, base, frac

    def diagnostics(self):
        dm, sm = C.c_double(0.0), C.c_double(0.0)
        _abi.check(self.lib.hfm_diagnostics(self.handle, C.byref(dm), C.byref(sm)), "hfm_diagnostics")
        return dm.value, sm.value

    def stats(self) -> dict:
        st = _abi.Stats()
        _abi.check(self.lib.hfm_get_stats(self.handle, C.byref(st)), "hfm_get_stats")
        return st.as_dict()

    def synchronize(self) -> None:
        _abi.check(self.lib.hfm_synchronize(self.handle), "hfm_synchronize")

    def device_bytes(self) -> int:
        return int(self.lib.hfm_scene_bytes(self.handle))

    def set_kernel_timing(self, on: bool) -> None:
        _abi.check(self.lib.hfm_set_kernel_timing(self.handle, int(bool(on))), "hfm_set_kernel_timing")

    def kernel_time(self, cls: int):
        ms, n = C.c_double(0.0), C.c_int64(0)
        _abi.check(self.lib.hfm_kernel_time(self.handle, int(cls), C.byref(ms), C.byref(n)), "hfm_kernel_time")
        return ms.value, n.value

    def kernel_time_reset(self) -> None:
        _abi.check(self.lib.hfm_kernel_time_reset(self.handle), "hfm_kernel_time_reset")


class HeatflowMPM:
    """Eulerian air and temperature coupled to a particle material; substep on the GPU.

    Same constructor as the reference (solver.py:42-43).  Extras, not in the reference: `body`
    replaces the default StaticBody built from `particles`; `cg="fused"` runs the pressure PCG in
    one launch per iteration (warp's Jacobi PCG in the single-reduction Chronopoulos-Gear form,
    HFM_FLAG_FUSED_CG in include/hfm.h) instead of warp's two-pass form; `cg="persistent"` runs
    that same iteration, with the same bits, as one cooperative launch per solve that keeps the
    vectors on chip (HFM_FLAG_PERSISTENT_CG: scenes stepped on their own whose 16^3 tiles the GPU
    holds resident at once, n <= 96 on MI355X; otherwise it is cg="fused");
    `precision="mixed"` solves both systems by iterative refinement with FP16 inner vectors
    (HFM_PREC_MIXED); the state and the advection stay fp32;
    `cg="mg"` preconditions the pressure PCG with a geometric multigrid V-cycle instead of
    Jacobi (HFM_FLAG_MG_PCG: n = m 2^k with m <= 8 and n >= 16; with precision="mixed" it
    preconditions the FP16 inner solves): the same stopping rule in far fewer iterations, fields
    equal to the Jacobi solve's to the solve tolerance, not in bits;
    `layout` picks the pressure PCG layout: None (chosen by grid size), "auto" / "auto-batch"
    (the autotuned table of heatflow_mpm.autotune for a scene stepped alone / in a batch), or
    dict(dir_planes=0|-1|4|8|16, cg="warp"|"fused"|"persistent"); with cg="mg" only its pass-A
    tiling applies;
    `advection="maccormack"` advects T and u, v, w by the limited MacCormack method (the
    reference's prepared but unused limiter, fields.py:51-65; HFM_FLAG_MACCORMACK, fp32 only)
    instead of the reference's semi-Lagrangian step.
    """

    CG_FORMS = ("warp", "fused", "persistent", "mg")
    ADVECTION = ("semi-lagrangian", "maccormack")
    PRECISIONS = {"fp32": _abi.PREC_FP32, "mixed": _abi.PREC_MIXED}

    def __init__(self, particles: np.ndarray, mat: Material = Material(), heater: Heater = Heater(),
                 grid: GridConfig = GridConfig(), spacing: float = 2.5e-3, t_init: float = 20.0,
                 kappa_air: float = KAPPA_AIR, device: str = "cuda:0", body=None, cg: str = "warp",
                 inputs: dict | None = None, precision: str = "fp32", layout=None,
                 advection: str = "semi-lagrangian"):
        if cg not in self.CG_FORMS:
            raise ValueError(f"cg must be one of {self.CG_FORMS}")
        dir_planes = 0
        if layout is not None:
            if isinstance(layout, str):
                from .autotune import lookup
                if layout not in ("auto", "auto-batch"):
                    raise ValueError("layout must be None, 'auto', 'auto-batch' or a dict")
                choice = lookup(grid.n, "batch" if layout == "auto-batch" else "solo")
            else:
                choice = dict(layout)
            if choice:
                dir_planes = int(choice.get("dir_planes", 0))
                if choice.get("cg") in ("warp", "fused", "persistent") and precision == "fp32" and cg != "mg":
                    cg = choice["cg"]
        self.layout = dict(dir_planes=dir_planes, cg=cg)
        if precision not in self.PRECISIONS:
            raise ValueError(f"precision must be one of {tuple(self.PRECISIONS)}")
        if advection not in self.ADVECTION:
            raise ValueError(f"advection must be one of {self.ADVECTION}")
        if precision == "mixed" and advection != "semi-lagrangian":
            raise ValueError("advection='maccormack' is available with precision='fp32' only")
        self.advection = advection
        if precision == "mixed" and cg not in ("warp", "mg"):
            raise ValueError(f"the {cg} PCG form is not available with precision='mixed'")
        self.precision = precision
        self.mat, self.heater, self.dev, self.spacing, self.kappa_air = mat, heater, device, spacing, kappa_air
        particles = np.asarray(particles, np.float32).reshape(-1, 3)
        self.n = len(particles)
        # `inputs`: scene_inputs(...) of these same arguments, prepared ahead (e.g. by a server)
        inp = inputs if inputs is not None else scene_inputs(self.n, mat, heater, grid, spacing, t_init, kappa_air)
        inp = dict(inp, params=dict(inp["params"]))
        if cg == "fused":
            inp["params"]["flags"] |= _abi.FLAG_FUSED_CG
        elif cg == "persistent":
            inp["params"]["flags"] |= _abi.FLAG_FUSED_CG | _abi.FLAG_PERSISTENT_CG
        elif cg == "mg":
            inp["params"]["flags"] |= _abi.FLAG_MG_PCG
        if advection == "maccormack":
            inp["params"]["flags"] |= _abi.FLAG_MACCORMACK
        inp["params"]["precision"] = self.PRECISIONS[precision]
        inp["params"]["dir_planes"] = dir_planes
        self._scene = _Scene(inp["params"], parse_device(device))
        self.grid = AirGrid(grid, self._scene)
        self.heat = HeatField(self.grid.n, self.grid.h, self._scene)
        self.cell_type, self.heater_geom, self.t_max = inp["cell_type"], inp["geom"], inp["t_max"]
        self.heat.set_fixed(inp["fixed"], inp["tfix"])
        self.heat.set_source(inp["source"], inp["rate"], inp["t_cap"])
        self.grid.set_static(inp["solid"], inp["k_faces"], inp["m_faces"])
        self.body = body if body is not None else (StaticBody(particles) if self.n else None)
        if self.body is not None:
            self._scene.set_particles(self.body.pos, self.body.vel, np.full(self.n, t_init, np.float32))

    def substep(self, dt: float) -> None:
        """One air/heat substep (solver.py:63-86), then the body's own step (solver.py:87-88)."""
        b = self.body
        if b is not None and not getattr(b, "static", False):
            self._scene.update_particles(b.pos, b.vel)
        self._scene.substep(dt)
        if b is not None:
            b.step(dt)

    def run(self, dt: float, substeps: int) -> float:
        """`substeps` substeps of a static scene back to back; returns the GPU time in ms."""
        if self.body is not None and not getattr(self.body, "static", False):
            raise ValueError("run() needs a static body; call substep() for a moving one")
        return self._scene.run(dt, substeps)

    # ---- state (solver.py:90-96)
    def positions(self):
        return self.body.positions() if self.body is not None else np.zeros((0, 3), np.float32)

    def velocities(self):
        return self.body.velocities() if self.body is not None else np.zeros((0, 3), np.float32)

    def temperatures(self):
        return self._scene.particles()[0][: self.n]

    def melt_fraction(self):
        return self._scene.particles()[1][: self.n]

    def yield_stress(self):
        return self._scene.particles()[2][: self.n]

    def viscosity(self):
        return self._scene.particles()[3][: self.n]

    def grid_T(self):
        return self._scene.get_field("T")

    def air_velocity(self):
        return self.grid.cell_velocity()

    def stats(self) -> dict:
        """Iteration counts (warp-quantised and exact), residuals and tolerances of the last substep."""
        return self._scene.stats()

    def synchronize(self) -> None:
        self._scene.synchronize()

    def close(self) -> None:
        self._scene.close()

    # ---- output records (solver.py:98-110)
    def static_dict(self) -> dict:
        hg = self.heater_geom
        return dict(L=self.grid.L, h=self.grid.h, n=self.grid.n, n_particles=self.n, spacing=self.spacing,
                    t_amb=self.grid.cfg.t_amb, cell_type=self.cell_type.astype(np.uint8),
                    heater_on=int(self.heater.on), heater_centre=np.array(hg["centre"], np.float32),
                    heater_facing=np.array(hg["facing"], np.float32), heater_tilt=self.heater.tilt,
                    heater_radius=self.heater.radius, heater_length=self.heater.length,
                    **{f"mat_{k}": v for k, v in asdict(self.mat).items() if not isinstance(v, str)})

    def frame_dict(self, t: float, with_air: bool = True) -> dict:
        if self.n:
            temp, melt = (a[: self.n] for a in self._scene.particles()[:2])    # one read for both
        else:
            temp = melt = np.zeros(0, np.float32)
        d = dict(t=np.float32(t), pos=self.positions().astype(np.float32), vel=self.velocities().astype(np.float32),
                 temp=temp.astype(np.float16), melt=melt.astype(np.float16),
                 grid_T=self.grid_T().reshape(-1).astype(np.float16))
        if with_air:
            d["air_u"] = self.air_velocity().astype(np.float16)
        return d
