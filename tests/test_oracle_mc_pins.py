"""Pins of the Monte Carlo oracle's direct shading, each chosen so that a plausible slip (a
swapped exponent range, a dropped factor, a wrong sign or vector, a biased survival rule)
fails it:

- light_terms (template.wgsl:3076-3087): Blinn-Phong at known half-angles;
- calc_ao (template.wgsl:1825-1841): exactly 1 on open and convex surfaces, and the closed
  form of the 5 taps in a right-angle corner of two exact boxes;
- sky_color (template.wgsl:2050-2112): zenith, horizon, ground, the falloff exponent, yaw and
  the sun disk;
- the Russian roulette (template.wgsl:4107-4115): unbiased in expectation;
- perturb_normal (template.wgsl:3396-3401): the alpha -> 0 and alpha = 1 limits;
- trace_path on a diffuse sphere under a point light with indirect light off: the closed form
  (1 - F) albedo (0.15 AO + L n.l shadow), with the hit point and normal from the ray-sphere
  intersection and Schlick's F at index 1.
"""
import math

import numpy as np
import pytest

import numericdream_amd as nd
import oracle
from tests import formula_table as ft

SPEC_MIN, SPEC_MAX = 4.0, 128.0  # SPEC_POWER_MIN / _MAX, template.wgsl:3071-3072


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def rot_about_z(v, deg):
    c, s = math.cos(math.radians(deg)), math.sin(math.radians(deg))
    return np.array([c * v[0] - s * v[1], s * v[0] + c * v[1], v[2]])


def light_terms(l, n, v, gloss):
    out = oracle.helper(36, list(unit(l)) + list(unit(n)) + list(unit(v)) + [gloss], 2)
    return out[0], out[1]


# ----------------------------------------------------------------------------- Blinn-Phong
def test_light_terms_geometry_and_exponent():
    n = np.array([0.0, 0.0, 1.0])
    for gloss in (0.0, 0.1, 0.5, 1.0):
        power = SPEC_MIN + (SPEC_MAX - SPEC_MIN) * gloss
        for theta in (10.0, 30.0, 60.0):
            # light and view mirrored about n: the half vector is n, the lobe's peak = glossiness
            l = rot_about_z(np.array([math.sin(math.radians(theta)), 0, math.cos(math.radians(theta))]), 20)
            v = np.array([-l[0], -l[1], l[2]])
            ndotl, spec = light_terms(l, n, v, gloss)
            assert ndotl == pytest.approx(math.cos(math.radians(theta)), abs=1e-12)
            assert spec == pytest.approx(gloss, abs=1e-12)
            # view along n, light at 2 theta: the half vector is theta off n
            l2 = np.array([math.sin(math.radians(2 * theta)), 0, math.cos(math.radians(2 * theta))])
            _, spec2 = light_terms(l2, n, n, gloss)
            assert spec2 == pytest.approx(gloss * math.cos(math.radians(theta)) ** power, rel=1e-9, abs=1e-300)
            # light and view along the same direction: h = l
            _, spec3 = light_terms(l, n, l, gloss)
            assert spec3 == pytest.approx(gloss * math.cos(math.radians(theta)) ** power, rel=1e-9, abs=1e-300)
        # the exponent read back from two half-angles is 4 at glossiness 0 and 128 at 1
        if gloss > 0:
            a = light_terms([math.sin(math.radians(40)), 0, math.cos(math.radians(40))], n, n, gloss)[1]
            b = light_terms([math.sin(math.radians(20)), 0, math.cos(math.radians(20))], n, n, gloss)[1]
            p = math.log(a / b) / math.log(math.cos(math.radians(20)) / math.cos(math.radians(10)))
            assert p == pytest.approx(power, rel=1e-9)
    # behind the surface: no diffuse term, and a half vector below the surface gives no specular
    ndotl, spec = light_terms([0.3, 0, -1], n, [0.2, 0, -1], 1.0)
    assert ndotl == 0.0 and spec == 0.0


# ----------------------------------------------------------------------------- ambient occlusion
def corner_scene():
    """Two exact boxes of half-extent 1 (cube_primitive, union): a floor [-1,1] x [-2,0] x
    [-1,1] and a wall [-2,0] x [0,2] x [-1,1], meeting in a right-angle corner along x = y = 0."""
    u, cm = ft.scene("cube_primitive", params=[1.0])
    u.instance_count = 2.0
    u.instances[1] = u.instances[0]
    for i, off in enumerate([(0.0, -1.0, 0.0), (-1.0, 1.0, 0.0)]):
        for k in range(3):
            u.instances[i].offset[k] = off[k]
    u.instances[1].combine_mode = 0.0
    return u, cm


def ao_closed(w, t):
    """Taps above the floor at x = w: the distance to the union is min(height, w)."""
    r = min(max(0.04 * t, 0.003), 0.25)
    occ = sum(max(s - min(s, w), 0.0) * 0.6 ** i for i, s in enumerate(r * (0.2 + 0.2 * np.arange(5))))
    return min(max(1.0 - 2.0 * occ / r, 0.0), 1.0)


def test_calc_ao_open_plane_convex_and_corner():
    u, cm = corner_scene()
    up = np.array([0.0, 1.0, 0.0])
    # open floor, far from the wall: no tap is occluded
    assert oracle.calc_ao(u, [[0.9, 0.0, 0.3]], [up], 6.25, code_map=cm)[0] == pytest.approx(1.0, abs=1e-12)
    # a convex edge (the floor's outer edge at x = 1): exactly 1 as well
    assert oracle.calc_ao(u, [[0.999, 0.0, 0.0]], [up], 6.25, code_map=cm)[0] == pytest.approx(1.0, abs=1e-12)
    # the inside corner: closed form of the 5 taps for several distances and radii (t sets the
    # radius, clamped to [0.003, 0.25])
    vals = []
    for t in (0.01, 1.0, 3.0, 6.25, 20.0):
        for w in (0.002, 0.01, 0.03, 0.05, 0.1, 0.2):
            got = oracle.calc_ao(u, [[w, 0.0, 0.0]], [up], t, code_map=cm)[0]
            assert got == pytest.approx(ao_closed(w, t), abs=1e-12), (w, t)
            vals.append(got)
    assert min(vals) == 0.0 and max(vals) > 1 - 1e-12 and any(0.1 < v < 0.9 for v in vals)
    # a sphere (exact SDF) is convex everywhere: 1 at every point of its surface
    us, cms = ft.scene("sphere_primitive", params=[1.0])
    rng = np.random.default_rng(3)
    n = rng.normal(size=(50, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    assert np.allclose(oracle.calc_ao(us, n * (1 + 1e-9), n, 4.0, code_map=cms), 1.0, atol=1e-12)


# ----------------------------------------------------------------------------- sky
def sky_uniforms():
    u = nd.default_uniforms()
    u.sky.mode = 0.0
    u.sky.intensity = 1.0
    for k, (h, z, g) in enumerate(zip((0.4, 0.5, 0.6), (0.1, 0.2, 0.9), (0.3, 0.2, 0.1))):
        u.sky.horizon[k], u.sky.zenith[k], u.sky.ground[k] = h, z, g
    u.sky.falloff = 1.0
    u.sky.sun_cos = -2.0
    return u


def test_sky_gradient_limits_falloff_and_yaw():
    u = sky_uniforms()
    hz, ze, gr = np.array(u.sky.horizon[:]), np.array(u.sky.zenith[:]), np.array(u.sky.ground[:])
    f32 = lambda a: np.asarray(a, np.float32).astype(np.float64)
    hz, ze, gr = f32(hz), f32(ze), f32(gr)
    for yaw in (0.0, 0.7, 2.5):
        u.sky.yaw = yaw
        assert np.allclose(oracle.sky(u, [[0, 1, 0]])[0], ze, atol=1e-12)
        assert np.allclose(oracle.sky(u, [[0, -1, 0]])[0], gr, atol=1e-12)
        az = np.linspace(0, 2 * math.pi, 7)
        horiz = np.stack([np.cos(az), np.zeros_like(az), np.sin(az)], 1)
        assert np.allclose(oracle.sky(u, horiz), hz, atol=1e-12)
        # below the horizon the blend toward the ground is linear in -y
        d = unit([0.6, -0.5, 0.2])
        assert np.allclose(oracle.sky(u, [d])[0], hz + (gr - hz) * (-d[1]), atol=1e-12)
    # falloff f: elevation y maps to y^(1/f); f = 2 at y = 1/4 is the midpoint
    u.sky.yaw = 0.0
    u.sky.falloff = 2.0
    y = 0.25
    d = [math.sqrt(1 - y * y), y, 0.0]
    assert np.allclose(oracle.sky(u, [d])[0], 0.5 * (hz + ze), atol=1e-12)
    # intensity scales everything, and negative colours clamp at 0
    u.sky.falloff = 1.0
    u.sky.intensity = 2.0
    assert np.allclose(oracle.sky(u, [[0, 1, 0]])[0], 2 * ze, atol=1e-12)
    u.sky.zenith[0] = -1.0
    assert oracle.sky(u, [[0, 1, 0]])[0][0] == 0.0


def test_sky_sun_disk():
    u = sky_uniforms()
    sun = unit([0.3, 0.8, -0.52])
    for k in range(3):
        u.sky.sun_dir[k] = sun[k] * 3.0  # any length: the reference normalises it
        u.sky.sun_color[k] = (1.0, 0.9, 0.7)[k]
    u.sky.sun_intensity = 5.0
    u.sky.sun_cos = math.cos(math.radians(4.0))
    sc = np.array(u.sky.sun_color[:], np.float64)
    si = 5.0
    for yaw in (0.0, 1.1):
        u.sky.yaw = yaw
        # the direction that sky_rotate maps onto the sun
        c, s = math.cos(yaw), math.sin(yaw)
        d = np.array([c * sun[0] + s * sun[2], sun[1], -s * sun[0] + c * sun[2]])
        u_nosun = sky_uniforms()
        u_nosun.sky.yaw = yaw
        base = oracle.sky(u_nosun, [d])[0]
        # the disk centre: the whole sun is added
        assert np.allclose(oracle.sky(u, [d])[0], base + sc * si, rtol=1e-6)
        # 5 degrees off the sun (outside the 4 degree disk): nothing is added
        off = unit(d + 0.0875 * unit(np.cross(d, [0, 0, 1])))
        assert np.allclose(oracle.sky(u, [off])[0], oracle.sky(u_nosun, [off])[0], atol=1e-12)


# ----------------------------------------------------------------------------- Russian roulette
@pytest.mark.parametrize("thr", [0.01, 0.3, 0.8, 1.0, 2.5])
def test_roulette_is_unbiased(thr):
    rng = np.random.default_rng(11)
    n = 20000
    seeds = rng.uniform(0, 1000, (n, 2)).astype(np.float32)
    surv = np.zeros(n)
    scale = np.zeros(n)
    for i, (a, b) in enumerate(seeds):
        surv[i], scale[i] = oracle.helper(38, [thr, float(a), float(b)], 2)
    p = min(max(thr, 0.05), 1.0)
    # survivors carry exactly 1 / p; the expected weight is 1 (unbiased)
    assert np.all(scale[surv > 0] == pytest.approx(1.0 / p, rel=1e-12))
    assert np.all(scale[surv == 0] == 0)
    sigma = math.sqrt(p * (1 - p) / n)
    assert abs(surv.mean() - p) <= 4 * sigma + 1e-12
    assert abs((surv * scale).mean() - 1.0) <= 4 * sigma / p + 1e-12
    if thr >= 1.0:
        assert surv.all()


# ----------------------------------------------------------------------------- roughness lobe
def test_perturb_normal_limits():
    rng = np.random.default_rng(5)
    n = unit([0.2, -0.4, 0.9])
    seeds = rng.uniform(0, 500, (4000, 2)).astype(np.float32)
    pert = lambda alpha, s: oracle.helper(37, list(n) + [alpha, float(s[0]), float(s[1])], 3)
    # at alpha <= MIN_ROUGHNESS (1e-4) the normal is returned unchanged
    assert np.array_equal(pert(1e-4, seeds[0]), n)
    means = []
    for alpha in (0.25, 0.5, 0.75, 1.0):
        d = np.array([pert(alpha, s) for s in seeds])
        assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-12)
        c = d @ n
        assert c.min() > 0
        # the tangential part averages out: the lobe is centred on n
        tang = d - np.outer(c, n)
        assert np.linalg.norm(tang.mean(0)) < 0.03
        means.append(c.mean())
    # alpha = 1 is the cosine-weighted hemisphere: E[cos] = 2/3, E[cos^2] = 1/2
    d1 = np.array([pert(1.0, s) for s in seeds])
    c1 = d1 @ n
    assert abs(c1.mean() - 2 / 3) < 0.015 and abs((c1 ** 2).mean() - 0.5) < 0.015
    assert all(a > b for a, b in zip(means, means[1:]))  # the lobe widens with alpha
    # In between: mix(n, h, alpha) lies on the arc from n toward h, at the angle
    # atan2(alpha sin(th), 1 - alpha + alpha cos(th)) from n when h is th from n. Its mean cosine
    # over the cosine-weighted hemisphere (density 2 sin cos), by quadrature:
    th = np.linspace(0.0, math.pi / 2, 20001)
    for alpha, got in zip((0.25, 0.5, 0.75), means[:3]):
        phi = np.arctan2(alpha * np.sin(th), 1 - alpha + alpha * np.cos(th))
        expect = np.trapezoid(np.cos(phi) * 2 * np.sin(th) * np.cos(th), th)
        assert abs(got - expect) < 0.005, (alpha, got, expect)


# ----------------------------------------------------------------------------- closed-form render
def diffuse_sphere_scene(w=40, h=30):
    """A unit sphere (exact SDF) of one grey, matte, index-1 colour stop, seen from
    (0, 0.35 * 3.5, -3.5) under one point light; Monte Carlo indirect light off, so the
    ambient term is 0.15 times the ambient occlusion."""
    u, cm = ft.scene("sphere_primitive", params=[1.0], size=(w, h), cam=(0.0, 0.35 * 3.5, -3.5), eps_px=1e-3)
    inst = u.instances[0]
    inst.color_count = 1.0
    c = inst.colors[0]
    for k in range(3):
        c.color[k] = 0.6
    c.position = 0.0
    c.glossiness = 0.0
    c.reflectiveness = 0.0
    c.transparency = 0.0
    c.ior = 1.0  # Schlick's r0 = 0: the reflected share is the grazing term (1 - |cos|)^5 alone
    c.roughness = 0.0
    u.debug_parts = float(1 << 6)  # PART_REFLECTIONS off: no mirror continuation
    u.mc_enabled = 0.0
    u.light_count = 1.0
    L = u.lights[0]
    for k, v in enumerate((2.0, 3.0, -2.0)):
        L.position_or_direction[k] = v
        L.color[k] = 1.0
    L.brightness = 10.0
    L.light_type = 0.0
    L.cast_shadows = 1.0
    L.shadow_softness = 16.0
    return u, cm


def sphere_closed_form(u):
    """Per pixel: hit flag, depth, n.l and the closed-form radiance of the diffuse sphere."""
    W, H = int(u.resolution[0]), int(u.resolution[1])
    o = np.array(u.camera_pos[:], np.float64)
    lp = np.array(u.lights[0].position_or_direction[:], np.float64)
    hit = np.zeros((H, W), bool)
    depth = np.zeros((H, W))
    ndl = np.zeros((H, W))
    rad = np.zeros((H, W))
    cosi = np.zeros((H, W))
    for y in range(H):
        for x in range(W):
            d = oracle.primary_dir(u, x, y)
            b = o @ d
            disc = b * b - (o @ o - 1.0)
            if disc <= 0:
                continue
            t = -b - math.sqrt(disc)
            p = o + d * t
            to = lp - p
            dist = np.linalg.norm(to)
            col = float(u.lights[0].brightness) / max(dist * dist, 0.01)  # point light falloff
            nl = max(p @ (to / dist), 0.0)
            hit[y, x], depth[y, x], ndl[y, x], cosi[y, x] = True, t, nl, abs(p @ d)
            # AO = 1 and shadow = 1 on a convex sphere; the local share is 1 - F with Schlick's
            # F = (1 - |d.n|)^5 at r0 = 0
            fres = (1.0 - abs(p @ d)) ** 5
            rad[y, x] = (1.0 - fres) * 0.6 * (0.15 * 1.0 + col * nl * 1.0)
    return hit, depth, ndl, rad, cosi


def test_trace_path_diffuse_sphere_closed_form():
    u, cm = diffuse_sphere_scene()
    hit, depth, ndl, rad, cosi = sphere_closed_form(u)
    out = oracle.render_mc(u, 4, code_map=cm)
    got = out["rgb"].astype(np.float64)
    # where the surface faces the light beyond the penumbra's reach (k n.l > 1), and on the
    # unlit side, where only the ambient term remains; away from grazing view angles, where the
    # hit (up to eps / |d.n| before the surface) would move Schlick's steep grazing term
    lit = hit & (ndl > 0.2) & (cosi > 0.3)
    dark = hit & (ndl == 0.0) & (cosi > 0.3)
    assert lit.sum() > 100 and dark.sum() > 20
    # the march stops once the DE falls below eps(t), up to eps / |d.n| before the surface
    eps = float(u.epsilon_coefficient) * depth[hit] + float(u.epsilon_floor)
    assert np.all(np.abs(out["depth"][hit] - depth[hit]) <= 2 * eps / cosi[hit] + 1e-6)
    for k in range(3):
        assert np.allclose(got[..., k][lit], rad[lit], rtol=1e-3), k
        assert np.allclose(got[..., k][dark], rad[dark], rtol=1e-3), k
