"""Pins of the oracle's reference gradient (ndo_adjoint_pixels), CPU only.

  closed form   on an exact SDF (a box face) the hit depth is t = (1 - n.o) / (n.d - c), so
                dt/do = -n / (n.d - c) exactly
  convergence   implicit-function gradient = lim h->0 of central differences of the polished
                hit depth, on the differentiable scene (Mandelbulb + Mandelbox): agreement
                improves steadily as h shrinks and reaches 1e-3 relative on >= 95 % of pixels at
                h = 1e-7 for every parameter. (At h = 1e-3 the hit depth of a fractal at an
                epsilon of 1e-3 pixel is not smooth on the scale of h, so the secant and the
                tangent disagree there; that is a property of the surface, not of either method.)
"""
import numpy as np

import numericdream_amd as nd
import oracle
from tests.test_oracle_march import box_scene


def test_box_face_closed_form():
    u, R = box_scene(res=(32, 24), cam=(0.2, 0.1, -4.0))
    pix = [(x, y) for y in range(8, 16) for x in range(10, 22)]
    t, gi, gf, ok = oracle.adjoint_pixels(u, pix, h=1e-6)
    c = float(u.epsilon_coefficient)
    o = np.array(u.camera_pos)
    n = np.array([0.0, 0.0, -1.0])  # the z = -1 face seen from z = -4
    checked = 0
    for i, (x, y) in enumerate(pix):
        if t[i] <= 0:
            continue
        d = oracle.primary_dir(u, x, y)
        p = o + d * t[i]
        if np.max(np.abs(p[:2])) > 0.9:
            continue  # near an edge the nearest face changes
        exp = -n / (n @ d - c)
        assert np.allclose(gi[i, 3:], exp, rtol=1e-9, atol=1e-12)
        assert np.allclose(gf[i, 3:], exp, rtol=1e-5, atol=1e-9)
        checked += 1
    assert checked > 40


def test_ift_equals_fd_limit():
    u = nd.resize(nd.config_uniforms(5, 0), 48, 48)
    pix = [(x, y) for y in range(0, 48, 2) for x in range(0, 48, 2)]
    prev = None
    for h in (1e-5, 1e-7):
        t, gi, gf, ok = oracle.adjoint_pixels(u, pix, h)
        hit = t > 0
        fr = []
        for k in range(6):
            m = hit & ((ok >> k) & 1).astype(bool) & (np.abs(gi[:, k]) > 0)
            rel = np.abs(gi[m, k] - gf[m, k]) / np.abs(gi[m, k])
            fr.append((rel < 1e-3).mean())
        fr = np.array(fr)
        if prev is not None:
            assert np.all(fr >= prev - 0.02)
        prev = fr
    assert np.all(prev >= 0.95), prev
