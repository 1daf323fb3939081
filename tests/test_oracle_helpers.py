"""Pins of the oracle's DE helper library and formulas (template.wgsl:947-1109, the four
formula files) by closed forms and invariants chosen so that a plausible slip (a wrong sign,
index or operand order, a dropped term) fails at least one of them. CPU only.

Which check pins which oracle function:
  box_fold               identity inside [-L, L]^3, mirror 2L - z outside
  sphere_fold            branch closed forms, continuity at r^2 = m^2 and r^2 = F^2
  sphere_invert_offset   |d'| |d| = R^2 inside, identity outside, continuity at |d| = R
  lattice_wrap           period = cell, range [-cell/2, cell/2], ties round half to even
  sort_desc_abs          sign-flip and permutation invariance, sorted output
  rotate_xyz_deg         norm preserved, right-handed quarter turns, identity at 0
  bulb_power_step        power-2 triplex closed form, |z^n| = r^n, freeze beyond r = 2
  Mandelbulb DE          (n-1)-fold rotational symmetry about z, mirror y -> -y
  Mandelbox DE           48-fold cube symmetry, single-iteration closed form
  Menger3 DE             0 iterations = unit box SDF, level-1 centre distance sqrt(2)/3
  Kleinian DE            lattice period with CScale = 0 inside the fold limit
  smooth combines        hard limits (k <= 1e-4), symmetric, exact outside the seam
  orbit traps            shape distances and the five accumulation modes
"""
import math

import numpy as np
import pytest

import numericdream_amd as nd
import oracle

rng = np.random.default_rng(1234)


def h(op, args, n=4):
    return oracle.helper(op, args, n)


# ---------------------------------------------------------------- folds
def test_box_fold():
    L = 0.8
    for _ in range(200):
        z = rng.uniform(-3, 3, 3)
        out = h(0, list(z) + [L])[:3]
        exp = np.where(np.abs(z) <= L, z, np.sign(z) * 2 * L - z)
        assert np.allclose(out, exp, atol=1e-12)


def test_sphere_fold_branches_and_continuity():
    m2, F2 = 0.25, 1.0
    for _ in range(200):
        z = rng.normal(size=3) * rng.uniform(0.01, 2.0)
        dr = rng.uniform(0.5, 3)
        out = h(1, list(z) + [dr, m2, F2])
        r2 = z @ z
        t = F2 / m2 if r2 < m2 else (F2 / r2 if r2 < F2 else 1.0)
        assert np.allclose(out[:3], z * t, rtol=1e-12) and out[3] == pytest.approx(dr * t, rel=1e-12)
    d = rng.normal(size=3); d /= np.linalg.norm(d)
    for r in (math.sqrt(m2), math.sqrt(F2)):
        a = h(1, list(d * r * (1 - 1e-9)) + [1.0, m2, F2])
        b = h(1, list(d * r * (1 + 1e-9)) + [1.0, m2, F2])
        assert np.allclose(a, b, atol=1e-7)


def test_sphere_invert_offset():
    c = np.array([0.3, -0.2, 0.5]); R2 = 0.64
    for _ in range(200):
        z = c + rng.normal(size=3) * rng.uniform(0.05, 1.5)
        out = h(2, list(z) + [2.0] + list(c) + [R2])
        d = z - c
        if d @ d < R2:
            dn = out[:3] - c
            assert np.linalg.norm(dn) * np.linalg.norm(d) == pytest.approx(R2, rel=1e-12)
            assert np.allclose(np.cross(dn, d), 0, atol=1e-9) and dn @ d > 0
            assert out[3] == pytest.approx(2.0 * R2 / (d @ d), rel=1e-12)
        else:
            assert np.allclose(out, list(z) + [2.0])
    u = rng.normal(size=3); u /= np.linalg.norm(u)
    a = h(2, list(c + u * 0.8 * (1 - 1e-10)) + [1.0] + list(c) + [R2])
    b = h(2, list(c + u * 0.8 * (1 + 1e-10)) + [1.0] + list(c) + [R2])
    assert np.allclose(a, b, atol=1e-8)


def test_lattice_wrap_period_range_and_ties():
    cell = 2.0
    for _ in range(100):
        p = rng.uniform(-10, 10, 3)
        k = rng.integers(-5, 6, 3)
        a = h(3, list(p) + [cell])[:3]
        b = h(3, list(p + k * cell) + [cell])[:3]
        assert np.allclose(a, b, atol=1e-9)
        assert np.all(np.abs(a) <= cell / 2 + 1e-12)
    # round half to even: 0.5 -> 0, 1.5 -> 2, 2.5 -> 2
    out = h(3, [1.0, 3.0, 5.0, cell])[:3]  # p / cell = 0.5, 1.5, 2.5
    assert np.allclose(out, [1.0, 3.0 - 4.0, 5.0 - 4.0])


def test_sort_desc_abs():
    for _ in range(100):
        z = rng.normal(size=3)
        out = h(4, list(z))[:3]
        assert np.allclose(out, sorted(np.abs(z), reverse=True))
        perm = rng.permutation(3); sg = rng.choice([-1, 1], 3)
        assert np.allclose(h(4, list(sg * z[perm]))[:3], out)


def test_rotate_xyz_deg():
    for _ in range(50):
        z = rng.normal(size=3); deg = rng.uniform(-180, 180, 3)
        out = h(5, list(z) + list(deg))[:3]
        assert np.linalg.norm(out) == pytest.approx(np.linalg.norm(z), rel=1e-12)
    assert np.allclose(h(5, [1, 0, 0, 0, 0, 90])[:3], [0, 1, 0], atol=1e-12)   # about z: x -> y
    assert np.allclose(h(5, [0, 1, 0, 90, 0, 0])[:3], [0, 0, 1], atol=1e-12)   # about x: y -> z
    assert np.allclose(h(5, [0, 0, 1, 0, 90, 0])[:3], [1, 0, 0], atol=1e-12)   # about y: z -> x
    # order: x first, then y, then z
    z = np.array([0.3, -0.7, 0.2])
    a = h(5, list(z) + [30, 0, 0])[:3]
    b = h(5, list(a) + [0, 40, 0])[:3]
    c = h(5, list(b) + [0, 0, 50])[:3]
    assert np.allclose(h(5, list(z) + [30, 40, 50])[:3], c, atol=1e-12)
    assert np.allclose(h(5, list(z) + [0, 0, 0])[:3], z)


# ---------------------------------------------------------------- Mandelbulb
def test_bulb_power2_triplex_closed_form():
    for _ in range(100):
        z = rng.uniform(-0.9, 0.9, 3); add = rng.uniform(-1, 1, 3); dr = rng.uniform(1, 3)
        out = h(6, list(z) + [dr] + list(add) + [2.0])
        x, y, w = z
        rho = math.hypot(x, y)
        exp = np.array([2 * w * (x * x - y * y) / rho, 4 * x * y * w / rho, w * w - rho * rho]) + add
        assert np.allclose(out[:3], exp, atol=1e-12)
        assert out[3] == pytest.approx(2 * np.linalg.norm(z) * dr + 1, rel=1e-12)


def test_bulb_power_norm_and_freeze():
    for n in (3.0, 8.0, 5.5):
        z = rng.uniform(-0.8, 0.8, 3)
        out = h(6, list(z) + [1.0, 0, 0, 0, n])
        assert np.linalg.norm(out[:3]) == pytest.approx(np.linalg.norm(z) ** n, rel=1e-12)
    z = [1.5, 1.5, 0.3]  # r > 2: carry returned unchanged
    assert np.allclose(h(6, z + [3.0, 1, 1, 1, 8.0]), z + [3.0])


def bulb_u(power=8.0, iters=8):
    u = nd.config_uniforms(1, 0)
    u.instances[0].params[0] = power
    u.instances[0].params[1] = iters
    return u


def test_mandelbulb_rotational_symmetry():
    u = bulb_u(8.0, 8)
    pts = rng.uniform(-1.2, 1.2, (200, 3))
    a = 2 * math.pi / 7  # (n - 1)-fold symmetry about z for z -> z^n + c
    R = np.array([[math.cos(a), -math.sin(a), 0], [math.sin(a), math.cos(a), 0], [0, 0, 1]])
    d0 = oracle.scene_de(u, pts)
    d1 = oracle.scene_de(u, pts @ R.T)
    d2 = oracle.scene_de(u, pts * np.array([1, -1, 1]))
    assert np.allclose(d0, d1, rtol=1e-6, atol=1e-9)
    assert np.allclose(d0, d2, rtol=1e-9, atol=1e-12)


# ---------------------------------------------------------------- Mandelbox
def box_u(scale=-1.5, iters=15):
    u = nd.config_uniforms(2, 0)
    u.instances[0].params[0] = scale
    u.instances[0].params[4] = iters
    return u


def test_mandelbox_cube_symmetry():
    u = box_u()
    pts = rng.uniform(-2.5, 2.5, (300, 3))
    d0 = oracle.scene_de(u, pts)
    for perm in ([1, 0, 2], [2, 1, 0], [0, 2, 1]):
        assert np.allclose(d0, oracle.scene_de(u, pts[:, perm]), rtol=1e-9, atol=1e-12)
    for sg in ([-1, 1, 1], [1, -1, 1], [1, 1, -1]):
        assert np.allclose(d0, oracle.scene_de(u, pts * np.array(sg)), rtol=1e-9, atol=1e-12)


def test_mandelbox_single_iteration_closed_form():
    S = -1.5
    u = box_u(S, 1)
    m2, F2 = 0.25, 1.0
    for _ in range(100):
        p = rng.normal(size=3)
        p *= rng.uniform(0.05, 0.45) / np.linalg.norm(p)  # inside the fold limit and min radius
        t = F2 / m2
        z = p * t * S + p
        dr = t * abs(S) + 1
        assert oracle.scene_de(u, p)[0] == pytest.approx(np.linalg.norm(z) / dr, rel=1e-12)


# ---------------------------------------------------------------- Menger3 / Kleinian
def menger_u(iters):
    u = nd.config_uniforms(2, 0)
    inst = u.instances[0]
    inst.code = 2
    for k in range(16):
        inst.params[k] = 0
    inst.params[0] = iters
    inst.params[1] = inst.params[2] = inst.params[3] = 1
    inst.params[4] = 3
    return u


def test_menger_zero_iterations_is_box_sdf():
    u = menger_u(0)
    pts = rng.uniform(-3, 3, (300, 3))
    exp = np.linalg.norm(np.maximum(np.abs(pts) - 1, 0), axis=1)
    assert np.allclose(oracle.scene_de(u, pts), exp, atol=1e-12)


def test_menger_level1_center_distance():
    # centre of the level-1 sponge: the nearest solid sub-cube is an edge cube centred at
    # (2/3, 2/3, 0) with half side 1/3, at distance sqrt(2) / 3
    u = menger_u(1)
    assert oracle.scene_de(u, [0, 0, 0])[0] == pytest.approx(math.sqrt(2) / 3, rel=1e-12)
    pts = rng.uniform(-1.5, 1.5, (200, 3))
    d0 = oracle.scene_de(u, pts)
    assert np.allclose(d0, oracle.scene_de(u, pts[:, [2, 0, 1]] * np.array([-1, 1, -1])), atol=1e-12)


def test_kleinian_lattice_period():
    u = nd.config_uniforms(3, 0)
    k = u.instances[3]
    u.instance_count = 1
    u.instances[0] = k
    inst = u.instances[0]
    inst.params[0] = 50.0   # fold limit beyond every test point: box_fold is the identity
    inst.params[7] = 0.0    # CScale = 0: pos is not added back
    pts = rng.uniform(-3, 3, (200, 3))
    d0 = oracle.scene_de(u, pts)
    d1 = oracle.scene_de(u, pts + np.array([2.0, 0, -4.0]))  # cell = 2
    assert np.allclose(d0, d1, rtol=1e-9, atol=1e-12)


# ---------------------------------------------------------------- combines
@pytest.mark.parametrize("op,hard", [(7, min), (8, max), (9, lambda c, b: max(-c, b)), (10, min), (11, min),
                                     (12, min)])
def test_combines_hard_limit(op, hard):
    for _ in range(50):
        a, b = rng.uniform(-1, 1, 2)
        assert h(op, [a, b, 1e-5], 1)[0] == pytest.approx(hard(a, b))


def test_smooth_min_properties():
    k = 0.3
    for _ in range(200):
        a, b = rng.uniform(-1, 1, 2)
        s = h(7, [a, b, k], 1)[0]
        assert s == pytest.approx(h(7, [b, a, k], 1)[0], abs=1e-12)
        assert s <= min(a, b) + 1e-12
        if abs(a - b) >= k:
            assert s == pytest.approx(min(a, b), abs=1e-12)
        assert min(a, b) - s <= k / 4 + 1e-12  # the polynomial seam dips at most k/4
        mx = h(8, [a, b, k], 1)[0]
        assert mx >= max(a, b) - 1e-12


def test_combine_de_weights():
    # union: h = 1 when the existing distance is closer, 0 when the new one is
    assert list(h(13, [0, 0.2, 0.5, 0.0], 2)) == [0.2, 1.0]
    assert list(h(13, [0, 0.5, 0.2, 0.0], 2)) == [0.2, 0.0]
    d, w = h(13, [1, 0.0, 0.05, 0.1], 2)
    assert w == pytest.approx(0.5 + 0.5 * 0.05 / 0.1)
    assert d == pytest.approx(h(7, [0.0, 0.05, 0.1], 1)[0])


# ---------------------------------------------------------------- traps
def test_trap_shapes():
    c = [0.1, -0.2, 0.3]
    q = np.array([0.5, 0.4, -0.6])
    d = q - c
    assert h(14, [1] + list(q) + c + [1, 1, 1, 0.7, 0.0], 1)[0] == pytest.approx(abs(np.linalg.norm(d) - 0.7))
    box = [0.2, 0.3, 0.1]
    e = np.abs(d) - box
    exp = abs(np.linalg.norm(np.maximum(e, 0)) + min(e.max(), 0))
    assert h(14, [2] + list(q) + c + box + [0.7, 0.0], 1)[0] == pytest.approx(exp)
    assert h(14, [3] + list(q) + c + [1, 1, 1, 0.7, 0.0], 1)[0] == pytest.approx(np.abs(d).min())
    tor = abs(math.hypot(math.hypot(d[0], d[2]) - 0.7, d[1]) - 0.1)
    assert h(14, [4] + list(q) + c + [1, 1, 1, 0.7, 0.1], 1)[0] == pytest.approx(tor)
    assert h(14, [0] + list(q) + c + [1, 1, 1, 0.7, 0.0], 1)[0] == pytest.approx(np.linalg.norm(d))


def test_trap_modes():
    zs = rng.normal(size=(6, 3))
    start = np.zeros(3)
    dist = np.abs(np.linalg.norm(zs, axis=1) - 1.0)  # sphere trap, radius 1, centre 0
    args = lambda mode: [mode, 1, 1.0, len(zs)] + list(start) + list(zs.ravel())
    assert h(18, args(0), 1)[0] == pytest.approx(dist.min())
    assert h(18, args(1), 1)[0] == pytest.approx(dist.max())
    assert h(18, args(2), 1)[0] == pytest.approx(dist.mean())
    assert h(18, args(3), 1)[0] == pytest.approx(dist[-1])
    assert h(18, args(4), 1)[0] == pytest.approx((np.argmin(dist) + 1) / len(zs))
