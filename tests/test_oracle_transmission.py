"""Pins of the oracle's transmissive path trace (oracle/nd_oracle_mc.hpp): refraction,
the inside march, Cauchy dispersion, shadows through transparent occluders and the light
pick, by what the optics and the estimators fix (Snell's law, the Abbe number, chord
lengths, invisibility of index-1 glass, Fresnel probabilities, unbiasedness)."""
import numpy as np
import pytest

import oracle
from tests import formula_table as ft


def H(op, args, n=8):
    return oracle.helper(op, args, n)


# --------------------------------------------------------------------------- dispersion
def test_cauchy_ior_meets_the_abbe_definition():
    """n(lambda_d) = n_d and n_F - n_C = (n_d - 1) / V (Fraunhofer d, F, C lines, um)."""
    for nd_ in (1.3, 1.5, 1.9):
        for v in (20.0, 40.0, 64.0):
            n = lambda lam: H(26, [nd_, v, lam])[0]
            assert abs(n(0.5876) - nd_) < 2e-5
            assert np.isclose(n(0.4861) - n(0.6563), (nd_ - 1.0) / v, rtol=2e-3)
            assert n(0.40) > n(0.55) > n(0.70)  # normal dispersion: blue bends most
    # Abbe numbers below 1 are clamped to 1
    assert np.isclose(H(26, [1.5, 0.2, 0.45])[0], H(26, [1.5, 1.0, 0.45])[0], rtol=1e-12)


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def test_refract_safe_follows_snell_and_total_internal_reflection():
    rng = np.random.default_rng(3)
    for _ in range(300):
        n = unit(rng.normal(size=3))
        d = unit(rng.normal(size=3))
        eta = float(rng.uniform(0.5, 1.6))
        out = H(27, list(d) + list(n) + [eta], 4)
        r, tir = out[:3], out[3] > 0.5
        nf = -n if d @ n > 0 else n  # the normal facing the ray
        cos_i = -(d @ nf)
        sin_i = np.sqrt(max(0.0, 1 - cos_i * cos_i))
        if eta * sin_i < 0.999:
            assert not tir
            assert np.isclose(np.linalg.norm(r), 1.0, atol=1e-9)
            cos_t = -(r @ nf)
            assert np.isclose(np.sqrt(max(0.0, 1 - cos_t * cos_t)), eta * sin_i, atol=1e-9)  # Snell
            assert cos_t > 0  # through the surface
            assert abs(np.linalg.det(np.array([d, nf, r]))) < 1e-9  # plane of incidence
        elif eta * sin_i > 1.001:
            assert tir
            assert np.allclose(r, d - 2 * (d @ nf) * nf, atol=1e-12)  # mirror direction
    # the normal's orientation does not matter
    d, n = unit([0.3, -1.0, 0.2]), unit([0.1, 1.0, -0.3])
    assert np.allclose(H(27, list(d) + list(n) + [0.7], 4), H(27, list(d) + list(-n) + [0.7], 4), atol=1e-14)


def test_dispersion_weights():
    """Gaussian weights of each wavelength's refracted direction against the hero's: they
    sum to 8; wide -> all 1; narrow -> all weight on the wavelength whose refraction is the
    hero direction; at normal incidence every wavelength goes straight -> all 1."""
    hero, nd_, v = 0.37, 1.5, 30.0
    lam = H(25, [hero])
    d, n = unit([0.5, -1.0, 0.1]), unit([0.0, 1.0, 0.0])
    for k in (0, 3, 7):
        ior_k = H(26, [nd_, v, lam[k] * 1e-3])[0]
        hero_dir = H(27, list(d) + list(n) + [1.0 / ior_k], 4)[:3]
        for sigma in (1e-4, 0.01, 0.1, 10.0):
            w = H(28, [hero] + list(d) + list(n) + [nd_, v, 0.0] + list(hero_dir) + [sigma])
            assert np.isclose(w.sum(), 8.0, rtol=1e-12)
            assert np.argmax(w) == k
        narrow = H(28, [hero] + list(d) + list(n) + [nd_, v, 0.0] + list(hero_dir) + [1e-5])
        assert np.isclose(narrow[k], 8.0, rtol=1e-6)
        wide = H(28, [hero] + list(d) + list(n) + [nd_, v, 0.0] + list(hero_dir) + [10.0])
        assert np.allclose(wide, 1.0, atol=1e-4)
    straight = H(28, [hero] + [0, -1, 0] + [0, 1, 0] + [nd_, v, 0.0] + [0, -1, 0] + [0.01])
    assert np.allclose(straight, 1.0, atol=1e-12)


# --------------------------------------------------------------------------- inside march
def sphere(radius=1.0, size=(9, 9), cam=(0.0, 0.0, -4.0)):
    return ft.scene("sphere_primitive", params=[radius], cam=cam, size=size, max_steps=256, max_dist=20.0)


def test_march_inside_exits_a_sphere_after_its_chord():
    """From a point inside a unit sphere the inside march exits at the far surface: t is past
    the chord length by less than one (grown) floor step, never before it."""
    u, cm = sphere()
    rng = np.random.default_rng(5)
    for _ in range(40):
        o = unit(rng.normal(size=3)) * rng.uniform(0.2, 0.95)
        d = unit(rng.normal(size=3))
        b = o @ d
        chord = -b + np.sqrt(b * b - (o @ o - 1.0))  # distance to the surface along d
        exited, t, ev = oracle.march_inside(u, o, d, 64, code_map=cm)
        assert exited
        if chord > 0.02:
            assert chord <= t <= chord + 0.004 * (1 + 6 / 64) ** ev + 1e-9
        assert ev >= 3  # the exit test starts at the third sample


def test_march_inside_warmup_and_budget():
    """Starting outside, the exit test still waits for the third sample: t = two steps of
    max(|DE|, floor), the floor growing by 1 + 6 / steps. A budget too short for the chord
    ends without an exit after exactly that many samples."""
    u, cm = sphere()
    o, d = np.array([0.0, 0.0, -2.0]), np.array([0.0, 0.0, -1.0])  # outside, moving away
    steps = 16
    g = 1 + 6 / steps
    exited, t, ev = oracle.march_inside(u, o, d, steps, code_map=cm)
    t1 = max(1.0, 0.004)           # |DE| at z = -2 is 1
    t2 = t1 + max(2.0, 0.004 * g)  # |DE| at z = -3 is 2
    assert exited and ev == 3 and np.isclose(t, t2, rtol=1e-12)
    exited, t, ev = oracle.march_inside(u, [0.0, 0.0, -0.999], [0.0, 0.0, 1.0], 4, code_map=cm)
    assert not exited and ev == 4 and t < 1.9


# --------------------------------------------------------------------------- shadows
def slab(transparency, color=(1.0, 1.0, 1.0), has_transparency=1.0):
    """A cube of half size 1 at the origin; shadow rays start below it and go up."""
    u, cm = ft.scene("cube_primitive", params=[1.0], cam=(0.0, 0.0, -4.0), size=(9, 9))
    inst = u.instances[0]
    inst.color_count = 1.0
    c = inst.colors[0]
    for k in range(3):
        c.color[k] = color[k]
    c.transparency = transparency
    u.has_transparency = has_transparency
    return u, cm


def test_transparent_shadow_tint():
    """Through a transparent occluder the visibility is the opaque-free march's factor times
    from_rgb(mix(1, c, a)) * a per crossing; the march itself does not depend on a or c."""
    o, d = [0.1, -3.0, 0.2], [0.0, 1.0, 0.0]
    hero = 0.41
    ref = oracle.visibility(*slab(1.0)[:1], hero, o, d, 10.0, 8.0, 40, code_map=slab(1.0)[1])
    assert ref.max() > 0.0
    for a, col in ((0.5, (0.2, 0.6, 1.0)), (0.9, (1.0, 0.3, 0.1)), (0.02, (0.5, 0.5, 0.5))):
        u, cm = slab(a, col)
        v = oracle.visibility(u, hero, o, d, 10.0, 8.0, 40, code_map=cm)
        # per wavelength of the hero's basis; the tolerance is the FP32 rounding of the
        # colour as the Uniforms hold it
        mixed = [1 + (col[k] - 1) * a for k in range(3)]
        lanes = oracle.spectral_basis(hero, mixed) * a
        assert np.allclose(v, ref * lanes, rtol=1e-6, atol=1e-12)
    # below SHADOW_OPAQUE_ALPHA, and in scenes without transparency, occluders are opaque
    assert np.all(oracle.visibility(*slab(0.005)[:1], hero, o, d, 10.0, 8.0, 40, code_map=slab(0.005)[1]) == 0)
    u, cm = slab(0.9, has_transparency=0.0)
    assert np.all(oracle.visibility(u, hero, o, d, 10.0, 8.0, 40, code_map=cm) == 0)


# --------------------------------------------------------------------------- trace_path
def glass(ior=1.0, transparency=1.0, abbe=0.0, soft=0.0, size=(9, 9), hidden=False):
    """A clear sphere seen from above at a slant: its centre ray runs through the sphere's
    centre (normal incidence), looks down at the ground colour of the sky and, reflected,
    up at the zenith, so forward and back scattered samples differ."""
    u, cm = sphere(size=size, cam=(0.0, 3.0, -2.5))
    inst = u.instances[0]
    inst.color_count = 1.0
    c = inst.colors[0]
    for k in range(3):
        c.color[k] = 1.0
    c.transparency, c.reflectiveness, c.ior, c.abbe = transparency, 0.0, ior, abbe
    c.roughness, c.glossiness, c.inner_max_steps = 0.0, 0.0, 128.0
    u.has_transparency = 1.0
    u.light_count = 0.0
    u.mc_enabled = 0.0
    u.photon_dispersion_soft = soft
    u.max_reflection_bounces = 3.0
    inst.hidden = 1.0 if hidden else 0.0
    return u, cm


def centre(u, cm, spp, **kw):
    x, y = int(u.resolution[0]) // 2, int(u.resolution[1]) // 2
    r = oracle.render_mc(u, spp, rect=(x, y, 1, 1), code_map=cm, per_sample=True, **kw)
    return r["samples"][:, 0, 0, :].astype(np.float64)


def test_index_one_clear_glass_is_invisible_on_axis():
    """A fully transparent sphere of index 1 on the optical axis: no Fresnel reflection at
    normal incidence, no bending, so every sample of the centre pixel is the sky behind it
    (the render with the sphere hidden, sample for sample)."""
    g = centre(*glass(ior=1.0), 64)
    h = centre(*glass(ior=1.0, hidden=True), 64)
    assert np.allclose(g, h, rtol=1e-6, atol=1e-7)


def test_fresnel_split_at_normal_incidence():
    """Index 1.5 at normal incidence reflects with probability r0 = 0.04 at each surface, and
    the transmitted path goes straight on: the share of centre samples that come out forward
    unchanged is 0.96^2 (+ 0.0015 for two internal reflections)."""
    spp = 4096
    g = centre(*glass(ior=1.5), spp)
    h = centre(*glass(ior=1.5, hidden=True), spp)
    # the tetrahedral normal of a 9-pixel image tilts by ~0.01 rad, so forward samples see
    # the sky a hair off axis (< 1e-4 relative); back-reflected ones see the zenith (~50% off)
    forward = np.all(np.isclose(g, h, rtol=1e-2, atol=0.0), axis=1).mean()
    want = 0.96 ** 2 + 0.96 * 0.04 * 0.04 * 0.96
    assert abs(forward - want) < 4 * np.sqrt(want * (1 - want) / spp), (forward, want)


def test_dispersion_collapse_is_unbiased_in_luminance():
    """Without soft weights a dispersive hit keeps one of the 8 wavelengths at weight 8. Each
    sample's colour is gamut-mapped (trace_path returns gamut_fit(rgb)), which keeps its
    luminance, and luminance is linear in the spectrum: on average over samples it equals the
    non-dispersive image's (index 1: every wavelength takes the same path)."""
    spp = 2048
    g = centre(*glass(ior=1.0, abbe=30.0, soft=0.0), spp)
    h = centre(*glass(ior=1.0, abbe=0.0), spp)
    assert not np.allclose(g, h)  # single-wavelength samples differ one by one
    Y = np.array([0.2126, 0.7152, 0.0722])
    yg, yh = g @ Y, h @ Y
    assert abs(yg.mean() - yh.mean()) < 4 * yg.std() / np.sqrt(spp)


def two_lights(pick):
    u, cm = ft.scene("sphere_primitive", params=[1.0], cam=(0.0, 0.0, -4.0), size=(16, 16))
    u.mc_enabled = 1.0
    u.mc_light_pick = 1.0 if pick else 0.0
    u.light_count = 2.0
    for i, (pos, col) in enumerate((((3.0, 3.0, -3.0), (1.0, 0.8, 0.6)), ((-3.0, 1.0, -2.0), (0.3, 0.5, 1.0)))):
        L = u.lights[i]
        L.light_type = 0.0
        for k in range(3):
            L.position_or_direction[k] = pos[k]
            L.color[k] = col[k]
        L.brightness = 10.0
        L.cast_shadows = 1.0
        L.shadow_softness = 8.0
        L.hard_shadows = 0.0
    return u, cm


def test_light_pick_is_unbiased():
    """With mc_light_pick one of two shadowed lights is chosen per shading point and scaled
    by total / weight: the mean image equals shading both lights."""
    spp = 256
    a = oracle.render_mc(*two_lights(True)[:1], spp, code_map=two_lights(True)[1], per_sample=True)
    b = oracle.render_mc(*two_lights(False)[:1], spp, code_map=two_lights(False)[1], per_sample=True)
    hit = b["depth"] > 0
    assert hit.mean() > 0.2
    sa, sb = a["samples"].astype(np.float64), b["samples"].astype(np.float64)
    assert not np.allclose(sa, sb)  # one light per sample differs from both
    da = sa.mean(0)[hit] - sb.mean(0)[hit]
    se = np.sqrt(sa.var(0)[hit] / spp + sb.var(0)[hit] / spp) + 1e-6
    assert np.abs(da / se).mean() < 1.2  # |z| averages 0.8 for unbiased estimates
    assert np.abs(sa.mean() - sb.mean()) < 0.01 * sb.mean()


# --------------------------------------------------------------------------- thin film
def film_reflectance_tm(n1, n2, n3, d_nm, lam_nm, cos1, pol):
    """Reflectance of one film (index n2, thickness d) between n1 and n3 by the
    characteristic (transfer) matrix method, for s or p polarisation."""
    s1 = np.sqrt(max(0.0, 1 - cos1 * cos1))
    c2 = np.sqrt(1 - (n1 * s1 / n2) ** 2 + 0j)
    c3 = np.sqrt(1 - (n1 * s1 / n3) ** 2 + 0j)
    # tilted admittances
    eta = (lambda n, c: n * c) if pol == "s" else (lambda n, c: n / c)
    e1, e2, e3 = eta(n1, cos1), eta(n2, c2), eta(n3, c3)
    delta = 2 * np.pi * n2 * d_nm * c2 / lam_nm
    M = np.array([[np.cos(delta), 1j * np.sin(delta) / e2], [1j * e2 * np.sin(delta), np.cos(delta)]])
    B, Cc = M @ np.array([1.0, e3])
    r = (e1 * B - Cc) / (e1 * B + Cc)
    return abs(r) ** 2


def test_thin_film_tint_is_the_airy_reflectance():
    """thin_film_tint = 1 + (R / R_peak - 1) strength, R the s/p-averaged reflectance of the
    film on its base (transfer-matrix method, independent of the Airy sum the oracle uses)
    and R_peak its maximum over the film phase; no interference (film index = base index or
    1 on 1) leaves 1."""
    hero = 0.23
    lam = H(25, [hero])
    for cosv in (1.0, 0.8, 0.45):
        for n2, n3, d in ((1.8, 1.5, 400.0), (1.33, 1.0, 250.0), (2.2, 1.6, 620.0)):
            tint = H(30, [hero, cosv, n3, n2, d, 1.0, 1.0])
            R = np.array([0.5 * (film_reflectance_tm(1.0, n2, n3, d, l, cosv, "s") +
                                 film_reflectance_tm(1.0, n2, n3, d, l, cosv, "p")) for l in lam])
            ratio = tint  # = R / peak at strength 1
            # R / ratio is the same peak for every wavelength
            peak = R / ratio
            assert np.allclose(peak, peak[0], rtol=1e-9)
            assert np.all(ratio <= 1 + 1e-12)
            half = H(30, [hero, cosv, n3, n2, d, 0.5, 1.0])
            assert np.allclose(half, 1 + (ratio - 1) * 0.5, rtol=1e-12)
    assert np.allclose(H(30, [hero, 0.7, 1.5, 1.5, 400.0, 1.0, 1.0]), 1.0, atol=1e-12)  # film = base
    assert np.allclose(H(30, [hero, 0.7, 1.0, 1.0, 400.0, 1.0, 1.0]), 1.0, atol=1e-12)  # no contrast
    assert np.allclose(H(30, [hero, 0.7, 1.5, 1.8, 400.0, 0.0, 1.0]), 1.0, atol=1e-12)  # strength 0


def test_film_noise_is_value_noise_on_the_hash33_lattice():
    """At lattice points the noise is 2 hash33(p) - 1; in between it is continuous, in
    [-1, 1], and changes per cell."""
    rng = np.random.default_rng(11)
    for _ in range(50):
        i = rng.integers(-20, 20, size=3).astype(np.float64)
        h = H(32, list(i), 3)
        assert np.all((h >= 0) & (h < 1))
        assert np.allclose(H(31, list(i), 3), 2 * h - 1, atol=1e-7)
        p = i + rng.random(3)
        e = 1e-7 * unit(rng.normal(size=3))
        assert np.allclose(H(31, list(p), 3), H(31, list(p + e), 3), atol=1e-5)
        assert np.all(np.abs(H(31, list(p), 3)) <= 1)
    # distinct neighbouring cells
    assert not np.allclose(H(32, [1.0, 2.0, 3.0], 3), H(32, [2.0, 2.0, 3.0], 3))


def test_film_of_the_base_index_is_invisible():
    """A film whose index equals its base's has no interference: the path trace of a mirror
    ball (reflective, film strength 1, film perturbation on) equals the one without film."""
    def mirror(film):
        u, cm = sphere(size=(12, 12), cam=(0.0, 1.5, -3.5))
        inst = u.instances[0]
        inst.color_count = 1.0
        c = inst.colors[0]
        for k in range(3):
            c.color[k] = 0.9
        c.reflectiveness, c.ior, c.glossiness = 1.0, 1.5, 0.5
        c.film_ior, c.film_thickness, c.film_angle_scale = 1.5, 300.0, 1.0
        c.film_strength = 1.0 if film else 0.0
        c.film_perturb, c.film_perturb_scale = 0.5, 0.2
        return u, cm
    a = oracle.render_mc(*mirror(True)[:1], 4, code_map=mirror(True)[1])["rgb"]
    b = oracle.render_mc(*mirror(False)[:1], 4, code_map=mirror(False)[1])["rgb"]
    assert np.allclose(a, b, rtol=1e-6, atol=1e-8)
    u, cm = mirror(True)
    u.instances[0].colors[0].film_ior = 2.2
    c = oracle.render_mc(u, 4, code_map=cm)["rgb"]
    assert not np.allclose(c, b, rtol=1e-3)  # a contrasting film shows


# --------------------------------------------------------------------------- hit cache
def test_octahedral_normal_packing():
    """oct_encode/oct_decode through two 16-bit snorm values: unit length, within the
    lattice's angular resolution (~3e-5 rad, bounded here by 1e-4) over both hemispheres,
    exact on the axes."""
    rng = np.random.default_rng(7)
    worst = 0.0
    for _ in range(2000):
        n = unit(rng.normal(size=3))
        q = H(33, list(n), 3)
        assert np.isclose(np.linalg.norm(q), 1.0, atol=1e-12)
        worst = max(worst, np.arccos(np.clip(q @ n, -1, 1)))
        assert np.sign(q[2]) == np.sign(n[2]) or abs(n[2]) < 1e-3
    assert worst < 1e-4, worst
    for a in np.eye(3).tolist() + (-np.eye(3)).tolist():
        assert np.allclose(H(33, a, 3), a, atol=1e-12)


def test_primary_cache_only_quantizes_normals():
    """With the primary hit cache on, the primary normal (and the normal after a mirror
    bounce at depth 0) is the packed one: the image moves by far less than its noise."""
    u, cm = ft.test_scene("mandelboxsort_box_fold", size=(24, 18))
    u.mc_enabled = 1.0
    a = oracle.render_mc(u, 4, code_map=cm)
    u.primary_cache_key = 1.0
    b = oracle.render_mc(u, 4, code_map=cm)
    assert np.array_equal(a["depth"], b["depth"])
    hit = a["depth"] > 0
    assert hit.mean() > 0.2
    na, nb = a["normal"][hit].astype(np.float64), b["normal"][hit].astype(np.float64)
    dn = np.degrees(np.arccos(np.clip((na * nb).sum(-1) / np.linalg.norm(na, axis=-1) / np.linalg.norm(nb, axis=-1),
                                      -1, 1)))
    assert 0 < dn.max() < 0.01
    assert np.abs(a["rgb"] - b["rgb"]).mean() < 0.02 * np.abs(a["rgb"]).mean()


# --------------------------------------------------------------------------- adaptive sampling
def adaptive(threshold, min_samples):
    u, cm = ft.test_scene("mandelboxsort_box_fold", size=(20, 14))
    u.mc_enabled = 1.0
    u.adapt_enabled = 1.0
    u.adapt_threshold = threshold
    u.adapt_min_samples = min_samples
    u.adapt_row_width = 1e6
    return u, cm


def test_adaptive_sampling_limits():
    """A pixel stops sampling once converged and keeps the mean of the samples it took. With
    an unbounded threshold every pixel converges at its min-samples count k: the image is the
    plain k-sample render. With a negative threshold no pixel converges (the standard error is
    never negative; at threshold 0 a pixel whose FP32 variance cancels to 0 still would):
    the image is the plain render of all samples. Pixels past adapt_row_width are never
    adaptive."""
    for k in (2, 5, 8):
        u, cm = adaptive(1e9, float(k))
        a = oracle.render_mc(u, 12, code_map=cm)["rgb"]
        u.adapt_enabled = 0.0
        b = oracle.render_mc(u, k, code_map=cm)["rgb"]
        assert np.allclose(a, b, rtol=1e-6, atol=1e-9), k
    u, cm = adaptive(-1.0, 2.0)
    a = oracle.render_mc(u, 6, code_map=cm)["rgb"]
    u.adapt_enabled = 0.0
    b = oracle.render_mc(u, 6, code_map=cm)["rgb"]
    assert np.allclose(a, b, rtol=1e-6, atol=1e-9)
    u, cm = adaptive(1e9, 2.0)
    u.adapt_row_width = 10.0  # columns 10.. are outside the adaptive rows
    a = oracle.render_mc(u, 6, code_map=cm)["rgb"]
    u.adapt_enabled = 0.0
    full, two = oracle.render_mc(u, 6, code_map=cm)["rgb"], oracle.render_mc(u, 2, code_map=cm)["rgb"]
    assert np.allclose(a[:, :10], two[:, :10], rtol=1e-6, atol=1e-9)
    assert np.allclose(a[:, 10:], full[:, 10:], rtol=1e-6, atol=1e-9)
