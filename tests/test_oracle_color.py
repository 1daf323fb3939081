"""Pins of the oracle's G-buffer colour attachments: fs_simple's simple_geometry and
render_lite/shade_lite (template.wgsl:4417-4538) and fs_aov's albedo (:4853-4896).

- undo_tonemap is the inverse of the tone curve c / (1 + c), up to its 1/255 clamp;
- simple_geometry on a unit sphere (exact SDF): the key-lit, depth-cued grey of the
  ray-sphere intersection's normal, and the background on a miss;
- shade_lite on the diffuse sphere under a point light: (1 - F) albedo (0.15 AO + L n.l), the
  same closed form as the path trace's direct light but with no shadow march;
- aov_albedo: the stop colour at reflectiveness 0, white at 1, the midpoint at 0.5; the sky on
  a miss.
"""
import math

import numpy as np
import pytest

import oracle
from tests import formula_table as ft
from tests.test_oracle_mc_pins import diffuse_sphere_scene, sphere_closed_form, unit


def test_undo_tonemap_inverts_the_tone_curve():
    for c in (0.0, 0.01, 0.3, 1.0, 7.5, 120.0, 254.0):
        x = c / (1 + c)
        got = oracle.helper(39, [x, x, x], 3)
        assert got == pytest.approx([c] * 3, rel=1e-9, abs=1e-12)
    # past 1 - 1/255 the divisor is clamped: white maps to 255, and inputs clamp to [0, 1]
    assert oracle.helper(39, [1.0, 2.0, -1.0], 3) == pytest.approx([255.0, 255.0, 0.0])


def geometry_view(u):
    u.debug_parts = float(1 << 16)  # PART_GEOMETRY_ONLY: fs_simple draws simple_geometry
    return u


def test_simple_geometry_on_a_sphere():
    u, cm = diffuse_sphere_scene()
    geometry_view(u)
    hit, depth, _, _, cosi = sphere_closed_form(u)
    col, _ = oracle.gbuffer_color(u, code_map=cm)
    o = np.array(u.camera_pos[:], np.float64)
    up, right = np.array(u.camera_up[:], np.float64), np.array(u.camera_right[:], np.float64)
    bg = np.array([0.10, 0.11, 0.13])
    tone_inv = lambda c: np.clip(c, 0, 1) / np.maximum(1 - np.clip(c, 0, 1), 1 / 255)
    H, W = hit.shape
    checked = 0
    for y in range(H):
        for x in range(W):
            if not hit[y, x]:
                assert np.allclose(col[y, x], tone_inv(bg), rtol=1e-6)
                continue
            if cosi[y, x] < 0.3:
                continue
            d = oracle.primary_dir(u, x, y)
            n = unit(o + d * depth[y, x])  # the sphere's normal at the intersection
            key = unit(-d + up * 0.45 - right * 0.35)
            shade = 0.18 + 0.82 * max(n @ key, 0.0)
            cue = min(max(depth[y, x] / float(u.max_dist), 0.0), 1.0) * 0.55
            a = np.array([0.72, 0.72, 0.73]) * shade
            assert np.allclose(col[y, x], tone_inv(a + (bg - a) * cue), rtol=2e-3), (x, y)
            checked += 1
    assert checked > 200


def test_shade_lite_on_the_diffuse_sphere():
    u, cm = diffuse_sphere_scene()
    u.debug_parts = 0.0  # render_lite; shade_lite has no reflection continuation to switch off
    hit, depth, ndl, rad, cosi = sphere_closed_form(u)
    col, alb = oracle.gbuffer_color(u, code_map=cm)
    # away from grazing views, where the hit (up to eps / |d.n| before the surface) moves
    # Schlick's steep grazing term: 3e-4 at |d.n| > 0.5, 1e-3 at 0.3
    sel = hit & (cosi > 0.5)
    assert sel.sum() > 150 and (sel & (ndl > 0.2)).sum() > 80 and (sel & (ndl == 0)).sum() > 10
    for k in range(3):
        assert np.allclose(col[..., k][sel], rad[sel], rtol=1e-3), k
        assert np.allclose(alb[..., k][hit], 0.6, atol=1e-6), k  # reflectiveness = transparency = 0
    # the sky on a miss, clamped for the albedo
    miss = ~hit
    sky = oracle.sky(u, np.array([oracle.primary_dir(u, x, y) for y, x in zip(*np.nonzero(miss))]))
    assert np.allclose(col[miss], sky, rtol=1e-6, atol=1e-9)
    assert np.allclose(alb[miss], np.clip(sky, 0, 1), rtol=1e-6, atol=1e-9)


@pytest.mark.parametrize("refl,transp,expect", [(0.0, 0.0, 0.6), (1.0, 0.0, 1.0), (0.5, 0.0, 0.8), (0.2, 0.5, 0.8)])
def test_aov_albedo(refl, transp, expect):
    u, cm = diffuse_sphere_scene(16, 12)
    u.debug_parts = 0.0  # the part mask would clear reflectiveness with PART_REFLECTIONS off
    c = u.instances[0].colors[0]
    c.reflectiveness, c.transparency = refl, transp
    hit, *_ = sphere_closed_form(u)
    _, alb = oracle.gbuffer_color(u, code_map=cm)
    assert np.allclose(alb[hit], expect, atol=1e-6)


def test_aov_albedo_follows_the_part_mask():
    """apply_part_mask comes before aov_albedo: with PART_REFLECTIONS off a mirror stop keeps
    its colour, with PART_COLOR_STRIPS off the stop colour becomes SIMPLE_ALBEDO."""
    u, cm = diffuse_sphere_scene(16, 12)
    u.instances[0].colors[0].reflectiveness = 1.0
    hit, *_ = sphere_closed_form(u)
    u.debug_parts = float(1 << 6)
    assert np.allclose(oracle.gbuffer_color(u, code_map=cm)[1][hit], 0.6, atol=1e-6)
    u.debug_parts = float(1 << 0)
    _, alb = oracle.gbuffer_color(u, code_map=cm)
    assert np.allclose(alb[hit], 1.0, atol=1e-6)  # a mirror: white whatever the colour
    u.instances[0].colors[0].reflectiveness = 0.0
    _, alb = oracle.gbuffer_color(u, code_map=cm)
    assert np.allclose(alb[hit], [0.72, 0.72, 0.73], atol=1e-6)
