"""Pins of the oracle's subsurface scattering (oracle/nd_oracle_mc.hpp: subsurface_setup,
diffuse_weight, subsurface_transmit, subsurface_light_path and their use in shade_local;
template.wgsl:3088-3218, 3234-3270), by what the model fixes: Lambert at amount 0, full
light unchanged when the surface faces it, the back-lit weight equal to the amount, |n.l|
without reach, Beer-Lambert attenuation, chord lengths through exact spheres, the depth cap,
opaque occluders, and the render without subsurface."""
import numpy as np

import oracle
from tests import formula_table as ft


def H(op, args, n=8):
    return oracle.helper(op, args, n)


def dw(amount, ndl, ndl_blur, reach):
    return H(34, [amount, ndl, ndl_blur] + list(reach))


REACH = [0.05, 0.1, 0.2, 0.35, 0.5, 0.7, 0.9, 1.0]


# --------------------------------------------------------------------------- diffuse weight
def test_diffuse_weight_limits():
    rng = np.random.default_rng(11)
    for _ in range(200):
        ndl, b = rng.uniform(-1, 1, 2)
        a = rng.uniform(0, 1)
        # amount 0 is Lambert, whatever the reach and the blurred normal
        assert np.allclose(dw(0.0, ndl, b, REACH), max(ndl, 0.0), atol=0)
        # no reach: the wrap follows the fine normal and the profile is |n.l|
        assert np.allclose(dw(1.0, ndl, b, [1e-12] * 8), abs(ndl), atol=1e-9)
    for a in (0.1, 0.5, 1.0):
        # facing the light (both normals): the full light, unchanged by the amount
        assert np.allclose(dw(a, 1.0, 1.0, REACH), 1.0, atol=1e-12)
        # light straight from behind: the profile's back term alone, weighted by the amount
        assert np.allclose(dw(a, -1.0, -1.0, REACH), a, atol=1e-12)
        # at the terminator some light wraps past it, more for a longer reach
        w = dw(a, 0.0, 0.0, REACH)
        assert np.all(w > 0) and np.all(np.diff(w) > 0)


def test_diffuse_weight_reach_blends_toward_the_blurred_normal():
    """With full reach the wrapped part sees only the blurred normal: changing the fine
    n.l within the lit side moves the result by the Lambert and back terms only."""
    a = 1.0
    for b in (-0.5, 0.0, 0.3, 0.9):
        w1 = dw(a, 0.2, b, [1.0] * 8)
        w2 = dw(a, 0.7, b, [1.0] * 8)
        assert np.allclose(w1, w2, atol=1e-12)  # amount 1: no Lambert share; n.l > 0: no back term
    # monotone in the blurred n.l for a partial reach
    prev = None
    for b in np.linspace(-1, 1, 21):
        w = dw(0.6, 0.1, b, REACH)
        if prev is not None:
            assert np.all(w >= prev - 1e-15)
        prev = w


# --------------------------------------------------------------------------- transmission
def test_transmit_is_beer_lambert():
    r = 0.3
    t0 = H(35, [r, 0.0] + REACH)
    assert np.allclose(t0, 1.0, atol=0)
    for d1, d2 in ((0.1, 0.2), (0.05, 0.7)):
        assert np.allclose(H(35, [r, d1 + d2] + REACH), H(35, [r, d1] + REACH) * H(35, [r, d2] + REACH), rtol=1e-12)
    # one mean free path (radius * reach) leaves 1/e in every lane
    for k, re in enumerate(REACH):
        assert np.isclose(H(35, [r, r * re] + REACH)[k], np.exp(-1.0), rtol=1e-12)


# --------------------------------------------------------------------------- light path
def sphere_scene(subsurface=0.0, radius=0.5):
    """A unit sphere (exact SDF) at the origin with one colour stop and one global light
    straight up (+y), shadows hard."""
    u, cm = ft.scene("sphere_primitive", params=[1.0], cam=(0.0, 0.0, -4.0), size=(9, 9), max_steps=256,
                     max_dist=20.0)
    inst = u.instances[0]
    inst.color_count = 1.0
    c = inst.colors[0]
    c.color[0], c.color[1], c.color[2] = 0.9, 0.5, 0.2
    c.subsurface, c.subsurface_radius = subsurface, radius
    L = u.lights[0]
    L.light_type = 1.0
    L.position_or_direction[0], L.position_or_direction[1], L.position_or_direction[2] = 0.0, 1.0, 0.0
    L.color[0] = L.color[1] = L.color[2] = 1.0
    L.brightness, L.cast_shadows, L.shadow_softness, L.hard_shadows = 1.0, 1.0, 16.0, 1.0
    u.light_count = 1.0
    return u, cm


ALBEDO = (0.9, 0.5, 0.2)
HERO = 0.37


def tetra_sphere(q, e):
    """Normalised tetrahedral-stencil gradient of |x| - 1 at q with tap offset e."""
    K = np.array([[1, -1, -1], [-1, -1, 1], [-1, 1, -1], [1, 1, 1]], np.float64)
    g = sum(k * (np.linalg.norm(q + k * e) - 1.0) for k in K)
    return g / np.linalg.norm(g)


def depth_of(trans, radius, reach):
    return -np.log(trans) * radius * reach


def test_light_path_from_inside_crosses_the_chord():
    """From a point on the unlit side (n.l < 0) the light path runs through the sphere to
    its far side: the depth read back from each lane's exp(-depth / (radius reach)) is the
    chord -2 p.l, past it by at most the exit slack and one bisected floor step."""
    u, cm = sphere_scene()
    rng = np.random.default_rng(2)
    radius = 0.5
    for _ in range(25):
        p = rng.normal(size=3)
        p /= np.linalg.norm(p)
        if p[1] > -0.15:
            p[1] = -abs(p[1]) - 0.15
            p /= np.linalg.norm(p)
        chord = -2.0 * p[1]  # light direction (0, 1, 0)
        tr, ev, blur, reach = oracle.subsurface_path(u, HERO, p, p, True, 0, False, 0.8, radius, ALBEDO, code_map=cm)
        d = depth_of(tr, radius, reach)
        assert np.allclose(d, d[0], rtol=1e-9)  # one depth for every wavelength
        # the last inside step is a floor step (an exact SDF never oversteps the exit), of at
        # most 0.05 radius grown by 1 + 6/64 per sample; 7 bisections cut it by 128
        floor = 0.05 * radius * (1 + 6 / 64) ** ev
        assert chord - 1e-9 <= d[0] <= chord + 0.0005 * (chord + floor) + 0.00005 + floor / 128, (chord, d[0], ev)
        # the blurred normal: the 4-tap tetrahedral gradient of the sphere's exact distance,
        # taken radius/2 off the surface with taps radius/4 (tends to the radial direction)
        for r in (radius, 0.002):
            _, _, blur, _ = oracle.subsurface_path(u, HERO, p, p, True, 0, False, 0.8, r, ALBEDO, code_map=cm)
            assert np.allclose(blur, tetra_sphere(p + p * r / 2, r / 4), atol=1e-12)
            assert np.allclose(blur, p, atol=r)


def test_light_path_cap_and_outside_light():
    u, cm = sphere_scene()
    down = np.array([0.0, -1.0, 0.0])
    # depth beyond 6 radii is cut to black: the 2-long diameter against radius 0.2 (cap 1.2)
    tr, _, _, _ = oracle.subsurface_path(u, HERO, down, down, True, 0, False, 0.8, 0.2, ALBEDO, code_map=cm)
    assert np.all(tr == 0.0)
    tr, _, _, _ = oracle.subsurface_path(u, HERO, down, down, True, 0, False, 0.8, 0.4, ALBEDO, code_map=cm)
    assert np.all(tr > 0.0)
    # a lit point without occlusion: 1 exactly, no march (the 4 evaluations are the blur taps)
    up = np.array([0.0, 1.0, 0.0])
    tr, ev, _, _ = oracle.subsurface_path(u, HERO, up, up, False, 0, False, 0.8, 0.4, ALBEDO, code_map=cm)
    assert np.all(tr == 1.0) and ev == 4


def test_light_path_through_occluders():
    """Below the sphere looking up at the light: an opaque sphere (no subsurface) blocks it;
    a subsurface sphere passes exp(-diameter / (radius reach)) of it."""
    p = np.array([0.0, -3.0, 0.0])
    n = np.array([0.0, 1.0, 0.0])
    u, cm = sphere_scene(subsurface=0.0)
    tr, _, _, _ = oracle.subsurface_path(u, HERO, p, n, False, 0, True, 0.8, 0.5, ALBEDO, code_map=cm)
    assert np.all(tr == 0.0)
    u, cm = sphere_scene(subsurface=0.7, radius=0.1)  # the occluder's own radius does not enter
    tr, _, _, reach = oracle.subsurface_path(u, HERO, p, n, False, 0, True, 0.8, 0.5, ALBEDO, code_map=cm)
    d = depth_of(tr, 0.5, reach)
    assert np.allclose(d, d[0], rtol=1e-9)
    assert 2.0 - 0.0006 <= d[0] <= 2.0 + 0.0005 * 4 + 0.00005 + 0.05 * 0.5 * 2.0 / 128, d[0]
    # without occlusion the outside march never starts
    tr, _, _, _ = oracle.subsurface_path(u, HERO, p, n, False, 0, False, 0.8, 0.5, ALBEDO, code_map=cm)
    assert np.all(tr == 1.0)


# --------------------------------------------------------------------------- the path trace
def lit_sphere(subsurface, parts=0):
    u, cm = sphere_scene(subsurface=subsurface, radius=0.3)
    L = u.lights[0]
    L.position_or_direction[0], L.position_or_direction[1], L.position_or_direction[2] = 0.3, 1.0, 0.4
    u.debug_parts = float(int(u.debug_parts) | parts)
    u.mc_enabled = 1.0
    u.resolution[0], u.resolution[1] = 12.0, 12.0
    return u, cm


def test_subsurface_below_epsilon_or_masked_is_the_plain_render():
    ref = oracle.render_mc(*lit_sphere(0.0)[:1], 4, code_map=lit_sphere(0.0)[1])["rgb"]
    tiny = oracle.render_mc(*lit_sphere(0.0009)[:1], 4, code_map=lit_sphere(0.0009)[1])["rgb"]
    assert np.array_equal(ref, tiny)
    u, cm = lit_sphere(0.8, parts=1 << 9)  # PART_SUBSURFACE off
    assert np.array_equal(ref, oracle.render_mc(u, 4, code_map=cm)["rgb"])
    u, cm = lit_sphere(0.8)
    sss = oracle.render_mc(u, 4, code_map=cm)["rgb"]
    hit = oracle.render_mc(u, 1, code_map=cm)["depth"] > 0
    # the unlit side of the sphere receives light through the material
    assert not np.array_equal(ref, sss)
    assert sss[hit].sum() > ref[hit].sum()
