"""Pins of the oracle's direct-lighting colour and its dual-number derivative
(ndo_radiance_pixels), on the radiance-adjoint scene (tests/radiance_scene.py: Mandelbulb +
Mandelbox union lit by a point, a global and a ray light, all shadowing):

- the colour direct_rgb gives at the march's own hit equals the oracle path trace's pixel
  (trace_path with Monte Carlo off and the reflection continuation cut), so the function
  differentiated is the renderer's;
- the derivative by dual numbers (the hit's implicit-function tangent, then duals through the
  normal, AO, lights and soft shadows) equals central differences of the colour at the
  perturbed scenes' own polished roots, in FP64 at h = 1e-8, for all six parameters. The
  colour is only piecewise smooth (the soft shadow's minimum over march steps switches step),
  so differences at h = 1e-6 straddle kinks on a few percent of entries; at 1e-8 they do not.
A dropped tangent (the hit depth's, the stencil width's, a light's distance) fails the second.
"""
import numpy as np

import oracle
from tests.radiance_scene import radiance_scene

W = 96


def sample_pixels(n=400, seed=5):
    rng = np.random.default_rng(seed)
    return np.stack([rng.integers(0, W, n), rng.integers(0, W, n)], 1)


def test_colour_is_the_path_traces():
    u = radiance_scene(size=(W, W))
    img = oracle.render_mc(u, 1, prec=0)
    pix = sample_pixels()
    dep = img["depth"][pix[:, 1], pix[:, 0]].astype(np.float64)
    hit = dep > 0
    r = oracle.radiance_pixels(u, pix[hit], t=dep[hit])
    ref = img["rgb"][pix[hit, 1], pix[hit, 0]]
    assert hit.sum() > 100
    assert np.abs(r["rgb"] - ref).max() < 1e-5  # FP32 storage of the image


def test_dual_equals_central_differences():
    u = radiance_scene(size=(W, W))
    r = oracle.radiance_pixels(u, sample_pixels(), h=1e-8)
    hit = r["t"] > 0
    assert hit.sum() > 100
    assert np.all(r["ok"][hit] == 63)
    du, fd = r["dual"][hit], r["fd"][hit]
    err = np.abs(du - fd) / np.maximum(np.abs(fd), 1e-2)
    print("\n[radiance oracle] entries %d, within 1e-3: %.4f, 1e-2: %.4f; mean |d/dtheta| %s"
          % (err.size, (err < 1e-3).mean(), (err < 1e-2).mean(), np.round(np.abs(fd).mean(axis=(0, 1)), 3)))
    assert (err < 1e-3).mean() >= 0.995
    assert (err < 1e-2).mean() >= 0.998
    # every parameter moves the colour somewhere
    assert np.all(np.abs(fd).max(axis=(0, 1)) > 1e-2)
