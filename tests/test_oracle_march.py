"""Pins of the oracle's camera, march, normals and soft shadow against closed forms on scenes
whose distance estimate is exact: Menger3 with 0 iterations is the unit-box SDF, so placed,
rotated and scaled instances of it give boxes whose ray intersections, normals and shadows
are known analytically. CPU only.

  primary_dir       centre pixel looks along forward; corner angle = the fov; row 0 looks
                    along -up (template.wgsl:5119-5127)
  march_to          hit depth of the slab intersection, within the on-hit bisection window
                    [t_true - eps/|d.n| - eps, t_true]; misses where the ray misses the box
  instance transform  offset / rotation frame / scale of the box
  normals           tetra, central and analytic normals = the face normal away from edges
  visibility_march  unoccluded = 1, fully blocked = 0, penumbra monotone across the edge
  FP32 restatement  agrees with FP64 on depth to 1e-5 relative
"""
import math

import numpy as np
import pytest

import numericdream_amd as nd
import oracle

rng = np.random.default_rng(99)


def box_scene(offset=(0, 0, 0), scale=(1, 1, 1), euler=(0, 0, 0), res=(64, 48), cam=(0.3, 0.4, -4.0)):
    u = nd.config_uniforms(1, 0)
    nd.resize(u, *res)
    inst = u.instances[0]
    inst.code = 2  # Menger3
    for k in range(16):
        inst.params[k] = 0
    inst.params[0] = 0  # 0 iterations: exact unit box
    inst.params[1] = inst.params[2] = inst.params[3] = 1
    inst.params[4] = 3
    for k in range(3):
        inst.offset[k] = offset[k]
        inst.scale[k] = scale[k]
    R = euler_matrix(*euler)
    for c in range(3):
        for r in range(3):
            inst.frame[c][r] = R[r, c]
    nd.camera_look_at(u, cam, (0, 0, 0), 60.0)
    nd.set_pixel_epsilon(u, 60.0)
    return u, R


def euler_matrix(rx, ry, rz):
    cx, sx, cy, sy, cz, sz = math.cos(rx), math.sin(rx), math.cos(ry), math.sin(ry), math.cos(rz), math.sin(rz)
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def slab(o, d, lo, hi):
    """Ray / axis-aligned box intersection: (t_enter, face normal) or (None, None)."""
    t0 = (lo - o) / d
    t1 = (hi - o) / d
    tmin = np.minimum(t0, t1)
    tmax = np.maximum(t0, t1)
    te, tx = tmin.max(), tmax.min()
    if te > tx or tx < 0:
        return None, None
    ax = int(np.argmax(tmin))
    n = np.zeros(3)
    n[ax] = -np.sign(d[ax])
    return te, n


def test_primary_dir_camera_model():
    u, _ = box_scene(res=(64, 48))
    f = np.array(u.camera_forward); up = np.array(u.camera_up); rt = np.array(u.camera_right)
    # the four pixels around the centre average to forward
    c = sum(oracle.primary_dir(u, x, y) for x in (31, 32) for y in (23, 24))
    assert np.allclose(c / np.linalg.norm(c), f, atol=1e-12)
    # row 0 looks along -up, column 0 along -right (template.wgsl:4523-4525)
    assert oracle.primary_dir(u, 32, 0) @ up < 0 and oracle.primary_dir(u, 0, 24) @ rt < 0
    # vertical fov: the pixel-centre rays of the top and bottom rows span 60 deg less one pixel
    a = oracle.primary_dir(u, 32, 0); b = oracle.primary_dir(u, 32, 47)
    ang = math.degrees(math.acos(np.clip(a @ b, -1, 1)))
    half = math.tan(math.radians(30))
    exp = 2 * math.degrees(math.atan(half * (1 - 1 / 48)))
    assert ang == pytest.approx(exp, abs=0.2)


@pytest.mark.parametrize("cfg", [dict(), dict(offset=(0.4, -0.3, 0.2)), dict(scale=(0.6, 1.3, 0.9)),
                                 dict(euler=(0.3, -0.5, 0.8), offset=(0.1, 0.2, -0.1)),
                                 dict(euler=(0.2, 0.7, 0.0), scale=(0.5, 0.8, 1.2))])
def test_march_depth_matches_slab_intersection(cfg):
    u, R = box_scene(**cfg)
    off = np.array(cfg.get("offset", (0, 0, 0)))
    sc = np.array(cfg.get("scale", (1, 1, 1)))
    o = np.array(u.camera_pos)
    W, H = int(u.resolution[0]), int(u.resolution[1])
    dirs = np.array([oracle.primary_dir(u, x, y) for y in range(H) for x in range(W)])
    hit, t, steps, ev = oracle.march(u, np.tile(o, (len(dirs), 1)), dirs)
    c = u.epsilon_coefficient
    checked = 0
    for i, d in enumerate(dirs):
        # world -> local: R^T (p - off) / s; the box is [-1, 1]^3 in local space
        ol = R.T @ (o - off) / sc
        dl = R.T @ d / sc
        te, nl = slab(ol, dl, -np.ones(3), np.ones(3))
        if te is None:
            assert not hit[i]
            continue
        nw = R @ (nl / sc); nw /= np.linalg.norm(nw)
        cosn = abs(d @ nw)
        if cosn < 0.05:
            continue  # grazing: outside the window the on-hit bisection guarantees
        # the DE is the box SDF scaled by min|s|: a lower bound, exact only for uniform scale
        de_scale = sc.min() / sc.max() if np.ptp(sc) > 0 else 1.0
        eps = c * te
        assert hit[i], (i, te)
        lo_bound = te - eps / (cosn * de_scale) - 2 * eps - 1e-12
        assert lo_bound <= t[i] <= te + 1e-9, (i, t[i], te)
        checked += 1
    assert checked > 100


def test_march_miss_far_plane_and_budget():
    u, _ = box_scene()
    o = np.array(u.camera_pos)
    d = -o / np.linalg.norm(o)
    away = -d
    hit, t, steps, ev = oracle.march(u, [o], [away])
    assert not hit[0] and t[0] > u.max_dist
    u.max_steps = 3
    hit, t, steps, ev = oracle.march(u, [o + 0.0], [np.array([1.0, 0, 0])])
    assert not hit[0] and steps[0] == 3 and ev[0] <= 3


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_normals_are_face_normals(mode):
    u, R = box_scene(euler=(0.4, 0.3, -0.2))
    for _ in range(60):
        ax = rng.integers(0, 3); sg = rng.choice([-1, 1])
        pl = rng.uniform(-0.8, 0.8, 3); pl[ax] = sg * 1.0
        nl = np.zeros(3); nl[ax] = sg
        p = R @ pl
        e = 1e-4
        # the unit-box DE is 0 inside, so every tap stays outside: 3 e off the face
        n = oracle.normal(u, p + (R @ nl) * 3e-4, mode, e)[0]
        assert np.allclose(n, R @ nl, atol=2e-3), (mode, n, R @ nl)


def test_shadow_closed_forms():
    # a ground slab (box scaled to half-thickness 0.2, wide) and a unit occluder above it
    u = nd.config_uniforms(2, 0)
    nd.resize(u, 32, 32)
    u.instance_count = 2
    for i in range(2):
        inst = u.instances[i]
        inst.code = 2
        for k in range(16):
            inst.params[k] = 0
        inst.params[1] = inst.params[2] = inst.params[3] = 1
        inst.params[4] = 3
        inst.step_safety = 1
        inst.hidden = 0
        for c in range(3):
            for r in range(4):
                inst.frame[c][r] = 1.0 if c == r else 0.0
    u.instances[0].scale[0] = 8; u.instances[0].scale[1] = 0.2; u.instances[0].scale[2] = 8
    u.instances[0].offset[1] = -1.2
    u.instances[1].offset[1] = 1.0
    u.instances[1].scale[0] = u.instances[1].scale[1] = u.instances[1].scale[2] = 0.5
    u.instances[1].combine_mode = 0
    up = np.array([0.0, 1.0, 0.0])
    # fully blocked: straight below the occluder
    vis, _ = oracle.shadow(u, [0, -0.996, 0], up, 20.0, 16.0, 128)
    assert vis == 0.0
    # unoccluded far away from the occluder: every k d / t >= 1
    vis, _ = oracle.shadow(u, [6, -0.996, 6], up, 20.0, 16.0, 128)
    assert vis == 1.0
    # monotone penumbra moving out from under the occluder edge (x = 0.5)
    xs = np.linspace(0.3, 1.5, 13)
    v = [oracle.shadow(u, [x, -0.996, 0], up, 20.0, 16.0, 128)[0] for x in xs]
    assert all(b >= a - 1e-12 for a, b in zip(v, v[1:])), v
    assert v[0] == 0.0 and v[-1] > 0.5


def test_fp32_restatement_agrees_with_fp64():
    u = nd.resize(nd.config_uniforms(2, 0), 48, 27)
    a = oracle.render_gbuffer(u, prec=0, normal_mode=0, shadows=True)
    b = oracle.render_gbuffer(u, prec=1, normal_mode=0, shadows=True)
    both = (a["hit"] > 0) & (b["hit"] > 0)
    assert (a["hit"] == b["hit"]).mean() > 0.98
    rel = np.abs(a["depth"][both] - b["depth"][both]) / a["depth"][both]
    assert np.median(rel) < 1e-5
