"""CPU checks of the C ABI: the library loads, exports every declared symbol, the Uniforms
mirror matches the reference layout asserts, and the workload builder is deterministic."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import numericdream_amd as nd
from numericdream_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in ("nd_hip.h", "nd_config.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(nd_[a-z0-9_]+)\s*\(", txt):
            names.add(m.group(1))
    return names


def test_exports_every_declared_symbol():
    decl = declared_symbols()
    assert decl == set(abi.EXPORTS), decl ^ set(abi.EXPORTS)
    for name in decl:
        assert hasattr(nd.lib, name), name


def test_uniforms_layout_matches_reference_asserts():
    U = abi.Uniforms
    assert C.sizeof(U) == 12016
    assert C.sizeof(abi.ColorStop) == 80 and abi.ColorStop.roughness.offset == 44
    assert abi.ColorStop.film_thickness.offset == 48
    assert C.sizeof(abi.FractalInstance) == 1760
    for f in ("trap_center", "trap_box", "mixins", "frame"):
        assert getattr(abi.FractalInstance, f).offset % 16 == 0
    assert abi.Warp.frame.offset == 96 and C.sizeof(abi.Sky) == 96
    expect = {"instances": 144, "lights": 7184, "accel_params": 7936, "march_verify": 8076, "warps": 8080,
              "photon_enabled": 8656, "sky": 8704, "debug_parts": 8828, "carves": 8880,
              "particle_systems": 9136, "approx_flags": 11792, "adapt_enabled": 11824,
              "fog_vis_light_layer": 11952}
    for k, v in expect.items():
        assert getattr(U, k).offset == v, k
    for k in ("accel_params", "warps", "photon_enabled", "photon_centre", "sky", "photon_pass", "carves",
              "particle_systems", "approx_flags", "adapt_enabled", "flat_coloring", "fog_tracking",
              "fog_vis_origin", "fog_vis_dims", "fog_vis_light_layer"):
        assert getattr(U, k).offset % 16 == 0, k


def test_defaults_match_reference_struct_defaults():
    u = nd.default_uniforms()
    assert u.epsilon_coefficient == pytest.approx(0.0005) and u.epsilon_floor == pytest.approx(0.00005)
    assert u.refine_fast == 4 and u.refine_hq == 20 and u.hq_footprint_budget_px == 2
    assert u.sky.sun_cos == -2 and list(u.sky.sun_dir) == [0, 1, 0]
    assert all(u.instances[i].code == -1 for i in range(4))
    assert list(u.fog_vis_light_layer[0]) == [-1] * 4


def test_config_builder_deterministic_and_seeded():
    a = nd.uniforms_bytes(nd.config_uniforms(2, 5))
    b = nd.uniforms_bytes(nd.config_uniforms(2, 5))
    c = nd.uniforms_bytes(nd.config_uniforms(2, 6))
    d = nd.uniforms_bytes(nd.config_uniforms(2, 5, seed=7))
    assert a == b and a != c and a != d
    assert [nd.config_job_count(k) for k in range(7)] == [0, 1, 16, 1, 256, 8, 0]


def test_config_cameras():
    u = nd.config_uniforms(1, 0)
    assert list(u.camera_pos) == [0, 0, -2.5]
    up = np.array(u.camera_up)
    assert np.degrees(2 * np.arctan(np.linalg.norm(up))) == pytest.approx(60.0, abs=1e-4)
    # right = normalize(forward x up) as camera.zig:103-106
    f, r = np.array(u.camera_forward), np.array(u.camera_right)
    assert np.allclose(r / np.linalg.norm(r), np.cross(f, up / np.linalg.norm(up)), atol=1e-6)
    theta_px = 2 * np.tan(np.radians(30)) / 360
    assert u.epsilon_coefficient == pytest.approx(1e-3 * theta_px, rel=1e-6) and u.epsilon_floor == 0
    for i in range(16):
        v = nd.config_uniforms(2, i)
        assert np.linalg.norm(np.array(v.camera_pos)) == pytest.approx(4.0, rel=1e-6)
        assert np.dot(np.array(v.camera_forward), -np.array(v.camera_pos) / 4.0) == pytest.approx(1.0, abs=1e-6)


def test_config4_distribution():
    sides, kinds = [], []
    for i in range(256):
        u = nd.config_uniforms(4, i)
        sides.append(int(u.resolution[0]))
        kinds.append(int(u.instances[0].code))
        assert u.resolution[0] == u.resolution[1]
        it = u.instances[0].params[1] if kinds[-1] == 0 else u.instances[0].params[4]
        assert 8 <= it <= 20 and it == int(it)
        if kinds[-1] == 0:
            assert 2 <= u.instances[0].params[0] <= 12
        else:
            s = u.instances[0].params[0]
            assert (-2.5 <= s <= -1.2) or (1.5 <= s <= 3.0)
        assert 2.5 - 1e-5 <= np.linalg.norm(np.array(u.camera_pos)) <= 5.0 + 1e-5
    assert min(sides) >= 64 and max(sides) <= 1024
    assert 0.35 < np.mean(kinds) < 0.65


def test_context_without_gpu_reports_device_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    assert nd.lib.nd_ctx_create(0, C.byref(h)) == abi.ND_E_DEVICE


def test_formula_ids_match_the_header():
    """abi.FORMULA_FILES, the id order of nd_builtin_formula and the test table agree."""
    import re
    from tests import formula_table as ft
    hdr = open(os.path.join(ROOT, "include", "nd_hip.h")).read()
    ids = {int(m.group(2)): m.group(3) for m in re.finditer(r"ND_FORMULA_(\w+) = (\d+),?\s*/\* (\S+) \*/", hdr)}
    assert len(ids) == abi.FORMULA_COUNT == int(re.search(r"#define ND_FORMULA_COUNT (\d+)", hdr).group(1))
    for i, f in enumerate(abi.FORMULA_FILES):
        assert ids[i] == f, (i, ids[i], f)
    for name in ft.FILES:
        assert abi.FORMULA_FILES[ft.ID[name]] == "formulas/%s.wgsl" % name
