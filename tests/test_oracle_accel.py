"""Pins of the oracle's empty-space accelerator (placeCascades, cs_build_accel, accel_skip and
the skip steps of march_to / visibility_march) on the exact box SDF (Menger3 with 0
iterations), where every safe distance has a closed form. CPU only."""
import ctypes as C

import numpy as np
import pytest

import numericdream_amd as nd
import oracle
from tests.test_oracle_march import box_scene, slab

L = oracle.lib()
L.ndo_place_cascades.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_void_p]
L.ndo_accel.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]


def box_sdf(p, half=1.0):
    q = np.abs(p) - half
    return np.linalg.norm(np.maximum(q, 0.0), axis=-1) + np.minimum(q.max(axis=-1), 0.0)


def oracle_grid(u, pts=None, prec=0):
    res, levels = int(u.accel_res), int(u.accel_levels)
    g = np.zeros((levels, res, res, res))
    pts = np.zeros((0, 3)) if pts is None else np.ascontiguousarray(pts, np.float64)
    skip = np.zeros(len(pts))
    r = L.ndo_accel(C.addressof(u), None, prec, g.ctypes.data, pts.ctypes.data if len(pts) else None, len(pts),
                    skip.ctypes.data if len(pts) else None)
    assert r == 0
    return g, skip


def test_cascades_match_place_cascades():
    """nd_accel_setup stamps what placeCascades computes; corners sit on their cell lattice,
    cells halve from the coarsest (2 max_dist / res) to level 0, every level holds the camera."""
    u, _ = box_scene(cam=(0.37, -1.2, -3.3))
    nd.enable_accel(u, res=32, levels=5, safety=0.9)
    cam = np.array(list(u.camera_pos), np.float64)
    ref = np.zeros((8, 4))
    L.ndo_place_cascades(cam.ctypes.data, float(u.max_dist), 5, 32, ref.ctypes.data)
    prm = np.array([list(u.accel_params[l]) for l in range(8)])
    assert np.array_equal(prm.astype(np.float32), ref.astype(np.float32))
    cells = prm[:5, 3]
    assert cells[4] == pytest.approx(2 * u.max_dist / 32, rel=1e-6)
    assert np.allclose(cells[1:] / cells[:-1], 2.0)
    assert np.all(prm[5:] == 0)
    for l in range(5):
        corner, cell = prm[l, :3], prm[l, 3]
        assert np.allclose(np.round(corner / cell) * cell, corner, atol=1e-6 * cell)
        assert np.all(cam >= corner) and np.all(cam < corner + 32 * cell)
    assert int(u.accel_enabled) == 1 and int(u.accel_levels) == 5 and int(u.accel_res) == 32


def test_grid_is_the_safe_distance_closed_form():
    """Each voxel holds max(SDF(centre) - sqrt(3)/2 cell, 0) * safety, in [level][z][y][x] order
    (a transposed index or a wrong corner fails here)."""
    u, _ = box_scene(cam=(0.4, 0.9, -3.5))
    nd.enable_accel(u, res=16, levels=3, safety=0.92)
    g, _ = oracle_grid(u)
    for l in range(3):
        corner, cell = np.array(list(u.accel_params[l])[:3]), float(u.accel_params[l][3])
        idx = (np.arange(16) + 0.5) * cell
        zz, yy, xx = np.meshgrid(idx, idx, idx, indexing="ij")
        centres = np.stack([corner[0] + xx, corner[1] + yy, corner[2] + zz], -1)
        want = np.maximum(box_sdf(centres) - 0.8660254 * cell, 0.0) * float(np.float32(0.92))
        assert np.allclose(g[l], want, rtol=1e-9, atol=1e-12), l


def test_skip_never_passes_the_surface():
    """For an exact SDF the skip at p is at most the distance from p to the surface."""
    u, _ = box_scene(cam=(0.2, 0.3, -3.0))
    nd.enable_accel(u, res=16, levels=4, safety=1.0)
    rng = np.random.default_rng(5)
    pts = rng.uniform(-4, 4, (20000, 3))
    _, skip = oracle_grid(u, pts)
    assert np.all(skip <= np.maximum(box_sdf(pts), 0.0) + 1e-12)
    assert (skip > 0).mean() > 0.3  # the grid does skip


def test_march_with_accelerator_hits_the_same_faces():
    """Depths with the accelerator equal the slab intersections within the bisection window
    (as without it), with fewer DE evaluations."""
    u, _ = box_scene(cam=(0.3, 0.4, -4.0), res=(48, 36))
    a = oracle.render_gbuffer(u, normal_mode=0, shadows=False)
    nd.enable_accel(u, res=32, levels=5, safety=0.92)
    b = oracle.render_gbuffer(u, normal_mode=0, shadows=False)
    assert np.array_equal(a["hit"], b["hit"])
    h = a["hit"] > 0
    eps = float(u.epsilon_coefficient) * a["depth"][h]
    assert np.all(np.abs(a["depth"][h] - b["depth"][h]) <= 2 * eps + 1e-9)
    assert b["evals"].sum() < 0.8 * a["evals"].sum()


def test_shadow_skips_keep_visibility():
    """Unoccluded light stays 1 and a blocked one 0 with accelerator skips."""
    u, _ = box_scene(cam=(0.0, 0.0, -4.0))
    nd.enable_accel(u, res=32, levels=5, safety=0.92)
    L.ndo_shadow.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_int,
                             C.c_void_p, C.c_void_p]
    def vis(o, d, dist):
        o = np.array(o, np.float64); d = np.array(d, np.float64); d /= np.linalg.norm(d)
        v = np.zeros(1); ev = np.zeros(1, np.int32)
        L.ndo_shadow(C.addressof(u), None, o.ctypes.data, d.ctypes.data, dist, 16.0, 128, v.ctypes.data, ev.ctypes.data)
        return v[0], ev[0]
    v, ev = vis((0, 0, -1.01), (0, 0, -1), 20.0)        # from the front face straight away
    assert v == pytest.approx(1.0)
    v, _ = vis((0, 0, -3.0), (0, 0, 1), 20.0)           # towards the box
    assert v == 0.0
