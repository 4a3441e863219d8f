"""Host-side logic and the C ABI boundary (CPU only: no compute calls on a device)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from heatflow_mpm import GridConfig, Heater, HeatflowMPM, Material, coupled_scene, ellipsoid_particles, plume_scene
from heatflow_mpm import _abi
from heatflow_mpm.heater import FAN, ROD, SHELL, TROUGH
from heatflow_mpm.solver import parse_device, scene_inputs

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "hfm.h")


# ---- heater masks (heater.py:24-65) ----------------------------------------------------------

def test_heater_cell_counts_n48_no_fan():
    """Counts quoted by the survey of the reference for n=48 (h=10 mm), fan off."""
    cell, solid, source, rate, kf, mf, geom = Heater().masks(48, 0.01, (0.24, 0.24), 1.2)
    counts = np.bincount(cell.ravel(), minlength=5)
    assert counts[SHELL] == 462 and counts[ROD] == 136 and counts[FAN] == 0 and counts[TROUGH] == 541
    assert rate == pytest.approx(613.0, abs=0.5)
    assert rate == pytest.approx(400.0 / (1.2 * 1005.0 * source.sum() * 0.01 ** 3))


def test_heater_cell_counts_n96_with_fan():
    cell, solid, source, rate, kf, mf, geom = Heater(fan=2.0).masks(96, 0.005, (0.24, 0.24), 1.2)
    counts = np.bincount(cell.ravel(), minlength=5)
    assert (counts[SHELL], counts[ROD], counts[FAN], counts[TROUGH]) == (2742, 679, 1940, 4662)
    assert rate == pytest.approx(402.0, abs=0.5)


def test_heater_geometry_invariants():
    h = Heater(fan=2.0)
    c, f, a, up = h.frame((0.24, 0.24))
    for v in (f, a, up):
        assert np.linalg.norm(v) == pytest.approx(1.0)
    assert abs(f @ a) < 1e-12 and abs(f @ up) < 1e-12 and abs(a @ up) < 1e-12
    assert f[2] < 0                                     # tilted down
    n, hh = 96, 0.005
    cell, *_ = h.masks(n, hh, (0.24, 0.24), 1.2)
    g = (np.arange(n) + 0.5) * hh
    P = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1) - c
    q, r, s = P @ f, P @ up, P @ a
    rods = cell == ROD
    # every rod cell lies within 7 mm + h/2 of one of the two rod axes at (q, r) = (-0.03, +-0.03)
    d = np.minimum(np.hypot(q + 0.03, r - 0.03), np.hypot(q + 0.03, r + 0.03))
    assert d[rods].max() <= 0.007 + hh / 2 + 1e-12
    assert np.abs(s[rods]).max() <= h.length / 2 - 0.015 + 1e-12
    # the fan slab is in front of the opening, the trough air behind it
    assert (q[cell == FAN] > 0).all() and (q[cell == TROUGH] <= 0).all()


def test_face_weights_follow_the_masks():
    cell, solid, source, rate, kf, mf, geom = Heater(fan=2.0).masks(48, 0.01, (0.24, 0.24), 1.2)
    fan = cell == FAN
    for axis in range(3):
        assert kf[axis].shape[axis] == 49
        sl_lo = [slice(None)] * 3
        sl_hi = [slice(None)] * 3
        sl_lo[axis], sl_hi[axis] = slice(1, 48), slice(0, 47)
        inner = [slice(None)] * 3
        inner[axis] = slice(1, 48)
        touching_solid = solid[tuple(sl_lo)] | solid[tuple(sl_hi)]
        touching_fan = (fan[tuple(sl_lo)] | fan[tuple(sl_hi)]) & ~touching_solid
        k_in = kf[axis][tuple(inner)]
        assert np.all(k_in[touching_solid] == 1.0) and np.all(k_in[touching_fan] == np.float32(0.05))
        assert np.all(k_in[~touching_solid & ~touching_fan] == 0.0)
        m_in = mf[axis][tuple(inner)]
        assert np.allclose(m_in[touching_fan], 0.05 * 2.0 * geom["facing"][axis])
        assert np.all(m_in[~touching_fan] == 0.0)


def test_heater_off_has_no_cells_and_no_rate():
    cell, solid, source, rate, *_ = Heater(on=False).masks(24, 0.02, (0.24, 0.24), 1.2)
    assert cell.max() == 0 and rate == 0.0


# ---- scene inputs (solver.py:42-61) ---------------------------------------------------------

def test_walls_and_rods_fixed_floor_adiabatic():
    grid = GridConfig(L=0.48, h=0.01)
    inp = scene_inputs(0, Material(), Heater(t_rod=300.0), grid, 2.5e-3, 20.0, 2.2e-5)
    fixed, tfix, n = inp["fixed"], inp["tfix"], grid.n
    assert fixed[0].all() and fixed[-1].all() and fixed[:, 0].all() and fixed[:, -1].all()
    assert fixed[:, :, -1].all()
    inner_floor = fixed[1:-1, 1:-1, 0]
    rods_on_floor = inp["cell_type"][1:-1, 1:-1, 0] == ROD
    assert not inner_floor[~rods_on_floor].any()               # floor k = 0 not fixed
    rods = inp["cell_type"] == ROD
    assert np.all(tfix[rods] == 300.0)                          # rods override walls
    wall_only = fixed.astype(bool) & ~rods
    assert np.all(tfix[wall_only] == 25.0)
    assert inp["t_max"] == 300.0
    p = inp["params"]
    assert p["n"] == 48 and p["vol"] == pytest.approx(2.5e-3 ** 3) and p["cg_check_every"] == 20
    assert p["heat_check_every"] == 10 and p["heat_tol"] == 1e-6 and p["heat_max_iter"] == 200


def test_grid_config_defaults_match_reference():
    g = GridConfig()
    assert (g.L, g.h, g.rho, g.t_amb, g.buoyancy, g.alpha_f_min, g.cg_tol, g.cg_max_iter) == \
        (0.48, 0.005, 1.2, 25.0, True, 0.05, 1e-4, 400)
    assert g.n == 96 and GridConfig(L=0.30).n == 60


def test_ellipsoid_seeding():
    P = ellipsoid_particles()
    assert 26500 < len(P) < 27100                                # 4/3 pi abc / s^3 = 26808
    assert P.dtype == np.float32
    assert P[:, 2].min() > 0 and abs(P[:, 0].mean() - 0.24) < 1e-3 and abs(P[:, 1].mean() - 0.24) < 1e-3
    assert np.array_equal(P, ellipsoid_particles())             # seeded


def test_scene_factories():
    kw, dt, K = plume_scene()
    assert kw["grid"].n == 48 and len(kw["particles"]) == 0 and (dt, K) == (2e-3, 200)
    kw, dt, K = coupled_scene(96)
    assert kw["grid"].n == 96 and kw["heater"].fan == 2.0 and K == 500 and kw["mat"].t_sol == 28.0


def test_cpu_device_is_rejected():
    with pytest.raises(ValueError):
        parse_device("cpu")
    assert parse_device("cuda:0") == 0 and parse_device("hip:1") == 1


# ---- C ABI -------------------------------------------------------------------------------------

def _declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hfm_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _abi.load()
    declared = _declared_symbols()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(_abi.SIGNATURES), "ctypes signatures and header disagree"


def test_struct_layouts_match_header(tmp_path):
    """Every field offset and the size of hfm_params and hfm_stats as gcc lays them out from
    include/hfm.h equal the ctypes mirrors."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{REPO}/include/hfm.h"', "int main(void) {"]
    for cname, cls in (("hfm_params", _abi.Params), ("hfm_stats", _abi.Stats)):
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in cls._fields_:
            lines.append(f'printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n"):
        if line:
            sname, field, val = line.split()
            got[(sname, field)] = int(val)
    for cname, cls in (("hfm_params", _abi.Params), ("hfm_stats", _abi.Stats)):
        assert got[(cname, "size")] == C.sizeof(cls), cname
        for field, _ in cls._fields_:
            assert got[(cname, field)] == getattr(cls, field).offset, (cname, field)


def test_errors_surface_without_a_device():
    """Creating a scene with no device visible fails with HFM_ERR_ARG and a message."""
    lib = _abi.load()
    if lib.hfm_device_count() > 0:
        pytest.skip("a device is visible")
    p = _abi.Params(n=8, cg_max_iter=10, cg_check_every=2, heat_max_iter=10, heat_check_every=2, h=0.01)
    h = C.c_void_p()
    rc = lib.hfm_scene_create(C.byref(p), 0, C.byref(h))
    assert rc == -1 and b"device" in lib.hfm_last_error()
    with pytest.raises(RuntimeError):
        HeatflowMPM(np.zeros((0, 3), np.float32), grid=GridConfig(L=0.08, h=0.01))


def test_mixed_precision_argument_checks():
    """The mixed form excludes the fused PCG form (Python and C checks; no device needed)."""
    with pytest.raises(ValueError, match="fused"):
        HeatflowMPM(np.zeros((0, 3), np.float32), grid=GridConfig(L=0.08, h=0.01), cg="fused", precision="mixed")
    with pytest.raises(ValueError, match="precision"):
        HeatflowMPM(np.zeros((0, 3), np.float32), grid=GridConfig(L=0.08, h=0.01), precision="fp16")
    lib = _abi.load()
    p = _abi.Params(n=8, cg_max_iter=10, cg_check_every=2, heat_max_iter=10, heat_check_every=2, h=0.01,
                    precision=_abi.PREC_MIXED, flags=_abi.FLAG_FUSED_CG)
    h = C.c_void_p()
    assert lib.hfm_scene_create(C.byref(p), 0, C.byref(h)) == -1
    assert b"FUSED" in lib.hfm_last_error()
    p = _abi.Params(n=8, cg_max_iter=10, cg_check_every=2, heat_max_iter=10, heat_check_every=2, h=0.01,
                    precision=7)
    assert lib.hfm_scene_create(C.byref(p), 0, C.byref(h)) == -1
    assert b"precision" in lib.hfm_last_error()


def test_autotune_table_lookup_and_candidates():
    """The autotuner's table lookup takes the nearest measured size (in log n); candidates respect
    the pass-A workgroup cap (4096 workgroups) and offer both PCG forms on every tiling; the
    persistent form only for a scene alone whose tiles fit one per CU."""
    from heatflow_mpm.autotune import candidates, lookup
    t = {"solo": {"64": {"choice": {"dir_planes": 8, "cg": "fused"}},
                  "128": {"choice": {"dir_planes": 16, "cg": "warp"}}}}
    assert lookup(96, "solo", t) == {"dir_planes": 16, "cg": "warp"}
    assert lookup(70, "solo", t) == {"dir_planes": 8, "cg": "fused"}
    assert lookup(96, "batch", t) is None
    c = candidates(192)
    assert {"dir_planes": 16, "cg": "warp"} in c and {"dir_planes": -1, "cg": "warp"} in c
    assert all(x["dir_planes"] != 4 for x in c)          # 12 x 12 x 48 tiles exceed the cap
    assert {"dir_planes": 8, "cg": "fused"} in c         # 12 x 12 x 24 tiles fit it
    assert all(x["cg"] == "warp" or x["dir_planes"] in (4, 8, 16) for x in candidates(64))
    assert all(x["cg"] != "persistent" for x in c)        # 576 tiles of 16 planes > 256 CUs
    assert {"dir_planes": 16, "cg": "persistent"} in candidates(96)        # 6 x 6 x 6 tiles
    assert {"dir_planes": 8, "cg": "persistent"} not in candidates(96)     # 432 tiles
    assert {"dir_planes": 16, "cg": "persistent"} in candidates(96, "batch")


def test_layout_argument_checks():
    lib = _abi.load()
    p = _abi.Params(n=8, cg_max_iter=10, cg_check_every=2, heat_max_iter=10, heat_check_every=2, h=0.01,
                    dir_planes=5)
    h = C.c_void_p()
    assert lib.hfm_scene_create(C.byref(p), 0, C.byref(h)) == -1
    assert b"dir_planes" in lib.hfm_last_error()
    with pytest.raises(ValueError, match="layout"):
        HeatflowMPM(np.zeros((0, 3), np.float32), grid=GridConfig(L=0.08, h=0.01), layout="fastest")


def test_multigrid_argument_checks():
    lib = _abi.load()
    h = C.c_void_p()
    for n, flags, prec in ((36, _abi.FLAG_MG_PCG, 0), (8, _abi.FLAG_MG_PCG, 0),
                           (32, _abi.FLAG_MG_PCG | _abi.FLAG_FUSED_CG, 0),
                           (32, _abi.FLAG_MG_PCG | _abi.FLAG_FUSED_CG, _abi.PREC_MIXED)):
        p = _abi.Params(n=n, cg_max_iter=10, cg_check_every=2, heat_max_iter=10, heat_check_every=2, h=0.01,
                        flags=flags, precision=prec)
        assert lib.hfm_scene_create(C.byref(p), 0, C.byref(h)) == -1
        assert b"HFM_FLAG_MG_PCG" in lib.hfm_last_error() or b"HFM_FLAG_FUSED_CG" in lib.hfm_last_error(), n


def test_maccormack_argument_checks():
    lib = _abi.load()
    h = C.c_void_p()
    p = _abi.Params(n=16, cg_max_iter=10, cg_check_every=2, heat_max_iter=10, heat_check_every=2, h=0.01,
                    flags=_abi.FLAG_MACCORMACK, precision=_abi.PREC_MIXED)
    assert lib.hfm_scene_create(C.byref(p), 0, C.byref(h)) == -1
    assert b"HFM_FLAG_MACCORMACK" in lib.hfm_last_error()
    with pytest.raises(ValueError, match="maccormack"):
        HeatflowMPM(np.zeros((0, 3), np.float32), grid=GridConfig(L=0.16, h=0.01), advection="maccormack",
                    precision="mixed")
    with pytest.raises(ValueError, match="advection"):
        HeatflowMPM(np.zeros((0, 3), np.float32), grid=GridConfig(L=0.16, h=0.01), advection="upwind")
