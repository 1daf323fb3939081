"""GPU parity of subsurface scattering in the path trace (subsurface_setup, diffuse_weight,
subsurface_light_path; template.wgsl:3088-3270) against the MC oracle, through the C ABI.

Scenes: the four-fractal scene with its Mandelbox (the instance that fills the view) made
subsurface (amount 0.6, radius 0.05, its strip colours kept); the same with two shadowed
lights and the Monte Carlo light pick (subsurface hits evaluate every light, others pick
one), with the empty-space accelerator, and with clear glass on the same Mandelbox's other
stops; a lit unit sphere of radius-0.3 subsurface. Full transport is compared like the
opaque path trace: MSE against the FP64 oracle at equal spp within twice the FP32 oracle
restatement's own MSE, and the mean radiance per channel within 3%. The part mask and an
amount at or below 0.001 give the render without subsurface exactly.
"""
import numpy as np
import pytest

import numericdream_amd as nd
from tests import formula_table as ft
from tests.test_gpu_transmission import compare_full, two_lights

pytestmark = pytest.mark.gpu


def sss_c3(w=64, h=36, amount=0.6, radius=0.05, glass=False):
    u = nd.resize(nd.config_uniforms(3, 0), w, h)
    inst = u.instances[1]
    for s in range(int(inst.color_count)):
        c = inst.colors[s]
        if glass and s % 2 == 1:
            c.transparency, c.ior, c.inner_max_steps, c.reflectiveness = 0.9, 1.5, 64.0, 0.08
        else:
            c.subsurface, c.subsurface_radius = amount, radius
    if glass:
        u.has_transparency = 1.0
    return u


def test_subsurface_mandelbox(renderer):
    compare_full(renderer, sss_c3(), 8, "subsurface")


def test_subsurface_with_light_pick(renderer):
    compare_full(renderer, two_lights(sss_c3()), 8, "subsurface + light pick")


def test_subsurface_with_accelerator(renderer):
    u = sss_c3()
    nd.enable_accel(u, res=32, levels=5, safety=0.92)
    compare_full(renderer, u, 8, "subsurface + accelerator")


def test_subsurface_with_glass(renderer):
    compare_full(renderer, sss_c3(glass=True), 8, "subsurface + glass")


@pytest.fixture
def sphere_bound(renderer):
    renderer.bind_formulas([ft.ID["sphere_primitive"]])
    yield renderer
    renderer.bind_formulas([0, 1, 2, 3])


def lit_sphere(subsurface, parts=0, size=48):
    u, cm = ft.scene("sphere_primitive", params=[1.0], cam=(0.0, 0.0, -4.0), size=(size, size), max_steps=256,
                     max_dist=20.0)
    inst = u.instances[0]
    inst.color_count = 1.0
    c = inst.colors[0]
    c.color[0], c.color[1], c.color[2] = 0.9, 0.5, 0.2
    c.subsurface, c.subsurface_radius = subsurface, 0.3
    L = u.lights[0]
    L.light_type = 1.0
    L.position_or_direction[0], L.position_or_direction[1], L.position_or_direction[2] = 0.3, 1.0, 0.4
    L.color[0] = L.color[1] = L.color[2] = 1.0
    L.brightness, L.cast_shadows, L.shadow_softness, L.hard_shadows = 1.0, 1.0, 16.0, 1.0
    u.light_count = 1.0
    u.debug_parts = float(int(u.debug_parts) | parts)
    u.mc_enabled = 1.0
    return u, cm


def test_subsurface_sphere(sphere_bound):
    u, cm = lit_sphere(0.8)
    compare_full(sphere_bound, u, 8, "subsurface sphere", code_map=cm)


def test_subsurface_gating_is_exact(sphere_bound):
    r = sphere_bound
    plain, _, _, _ = r.path_trace(lit_sphere(0.0)[0], 4)
    tiny, _, _, _ = r.path_trace(lit_sphere(0.001)[0], 4)
    masked, _, _, _ = r.path_trace(lit_sphere(0.8, parts=1 << 9)[0], 4)  # PART_SUBSURFACE off
    sss, _, _, _ = r.path_trace(lit_sphere(0.8)[0], 4)
    assert np.array_equal(plain, tiny)
    assert np.array_equal(plain, masked)
    assert sss.sum() > plain.sum()
