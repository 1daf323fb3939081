"""GPU radiance adjoint (nd_radiance_adjoint) on the radiance-adjoint scene
(tests/radiance_scene.py: Mandelbulb + Mandelbox union, a point, a global and a ray light,
all shadowing), through the C ABI:

- colour and d(colour)/d(theta) against the FP64 oracle (dual numbers at its polished root),
  pointwise on sampled pixels;
- central differences of the GPU's own forward colour at fixed seeds (camera +- h), pointwise,
  on a scene of exact distance fields (a sphere on a slab, same lights);
- the colour it differentiates is the path trace's at the polished root: on exact distance
  fields with a hundredth-of-a-pixel epsilon nd_path_trace's colour (Monte Carlo off,
  reflection continuation cut) equals the adjoint's forward colour;
- loss and gradient are the sum over pixels of |rgb - target|^2 and its derivative from the
  per-pixel outputs; descent on the colour loss recovers perturbed parameters (the Mandelbox's
  scale and min radius with views that frame it).
"""
import numpy as np
import pytest

import numericdream_amd as nd
import oracle
from tests.radiance_scene import radiance_scene
from tests.torch_adjoint import apply_params

pytestmark = pytest.mark.gpu

W = 96
TRUTH = np.array([8.0, -1.5, 0.5, 0.0, 0.0, 0.0])


def forward(r, views):
    n = sum(int(v.resolution[0]) * int(v.resolution[1]) for v in views)
    rgb = np.zeros(n * 3, np.float32)
    drgb = np.zeros(n * 18, np.float32)
    out = r.radiance_adjoint(views, rgb_out=rgb, drgb_out=drgb)
    return rgb.reshape(-1, 3), drgb.reshape(-1, 3, 6), out


def report(name, **kv):
    print("\n[radiance] %s %s" % (name, " ".join("%s=%.6g" % (k, v) for k, v in kv.items())), flush=True)


def test_against_fp64_oracle(renderer):
    u = radiance_scene(size=(W, W))
    print("\n[case] radiance adjoint vs FP64 oracle %dx%d" % (W, W), flush=True)
    rgb, drgb, (_, _, _, st) = forward(renderer, [u])
    rng = np.random.default_rng(11)
    pix = np.stack([rng.integers(0, W, 500), rng.integers(0, W, 500)], 1)
    ref = oracle.radiance_pixels(u, pix, h=1e-8)
    idx = pix[:, 1] * W + pix[:, 0]
    both = (ref["t"] > 0) & (np.abs(rgb[idx]).sum(1) > 0)
    g, gd = rgb[idx][both], drgb[idx][both]
    o, od = ref["rgb"][both], ref["dual"][both]
    val_ok = (np.abs(g - o) <= 2e-3 * np.maximum(np.abs(o), 1e-2)).all(1).mean()
    scale = np.abs(od).max(axis=(0, 1))  # per-parameter scale of the derivative
    der_err = np.abs(gd - od) / np.maximum(np.abs(od), 0.02 * scale)
    report("vs oracle", pixels=both.sum(), value_ok=val_ok, deriv_within_2pct=(der_err < 0.02).mean(),
           deriv_within_10pct=(der_err < 0.1).mean(), median=np.median(der_err))
    assert both.sum() > 150
    assert val_ok >= 0.98
    assert (der_err < 0.02).mean() >= 0.95
    assert (der_err < 0.1).mean() >= 0.98


def camera_moved(base, offset):
    out = []
    for b in base:
        u = nd.Uniforms.from_buffer_copy(b)
        for k in range(3):
            u.camera_pos[k] = b.camera_pos[k] + float(offset[k])
        out.append(u)
    return out


def test_against_central_differences_of_the_forward(renderer):
    """Central differences of the GPU's own forward colour (camera +- h, fixed seeds) against
    the adjoint, pixel by pixel, with shadows, on the sphere-and-slab scene: exact distance
    fields, so the colour is smooth at the scale of FP32 differences. A fractal's escape-time
    DE changes its iteration count across the surface and its colour is only piecewise smooth
    at much finer scales (FP64 differences leave the linear range below h = 1e-5 there); the
    fractal's derivative is checked against the FP64 oracle's at h = 1e-8 instead."""
    from tests.radiance_scene import SMOOTH_CODES, smooth_scene
    renderer.bind_formulas(SMOOTH_CODES)
    try:
        base = [smooth_scene(size=(W, W))]
        print("\n[case] radiance adjoint vs central differences, sphere and slab %dx%d" % (W, W), flush=True)
        rgb, drgb, _ = forward(renderer, base)
        # against the FP64 oracle too (all six derivatives; the formula ones are zero here)
        rng = np.random.default_rng(2)
        pix = np.stack([rng.integers(0, W, 400), rng.integers(0, W, 400)], 1)
        ref = oracle.radiance_pixels(base[0], pix, h=1e-8, code_map=SMOOTH_CODES)
        idx = pix[:, 1] * W + pix[:, 0]
        both = (ref["t"] > 0) & (np.abs(rgb[idx]).sum(1) > 0)
        od = ref["dual"][both]
        oe = np.abs(drgb[idx][both] - od) / np.maximum(np.abs(od), 0.02 * np.abs(od).max())
        report("smooth vs oracle", pixels=both.sum(), within_2pct=(oe < 0.02).mean())
        assert both.sum() > 100 and (oe < 0.02).mean() >= 0.98
        h = 1e-3
        for k in (3, 4, 5):
            dp = np.zeros(6)
            dp[k] = h
            cp, _, _ = forward(renderer, camera_moved(base, dp[3:]))
            cm, _, _ = forward(renderer, camera_moved(base, -dp[3:]))
            live = (np.abs(rgb).sum(1) > 0) & (np.abs(cp).sum(1) > 0) & (np.abs(cm).sum(1) > 0)
            fd = (cp - cm)[live] / (2 * h)
            ad = drgb[live][:, :, k]
            e = np.abs(ad - fd) / np.maximum(np.abs(fd), 0.02 * np.abs(fd).max())
            report("fd k=%d" % k, pixels=live.sum(), median=np.median(e), within_5pct=(e < 0.05).mean(),
                   within_20pct=(e < 0.2).mean(), mean_abs=np.abs(fd).mean())
            assert np.median(e) < 0.01, k
            assert (e < 0.05).mean() >= 0.95, k
    finally:
        renderer.bind_formulas([0, 1, 2, 3])


def test_forward_is_the_path_traces(renderer):
    """The colour the adjoint differentiates is nd_path_trace's (Monte Carlo off, reflection
    continuation cut) at the polished root instead of the march's hit, which its bisection
    leaves up to eps(t) past it. On exact distance fields with the epsilon at a hundredth of a
    pixel the two points coincide to far below a pixel, and so do the colours. (On a fractal a
    tiny epsilon makes the tetrahedral normal's stencil tiny, and FP32 rounding of the DE then
    decides the normal; the oracle shows that the same function evaluated at the march's hit is
    the path trace's pixel, tests/test_oracle_radiance_adjoint.py.)"""
    from tests.radiance_scene import SMOOTH_CODES, smooth_scene
    renderer.bind_formulas(SMOOTH_CODES)
    try:
        u = smooth_scene(size=(W, W))
        nd.set_pixel_epsilon(u, 60.0, 0.01)
        print("\n[case] radiance adjoint forward vs nd_path_trace, sphere and slab, eps 0.01 px", flush=True)
        rgb, _, _ = forward(renderer, [u])
        pt, depth, _, _ = renderer.path_trace(u, 1, depth=np.zeros((W, W), np.float32))
        hit = (depth.reshape(-1) > 0) & (np.abs(rgb).sum(1) > 0)
        a, b = rgb[hit], pt.reshape(-1, 3)[hit]
        ok = (np.abs(a - b) <= 1e-3 * np.maximum(np.abs(b), 1e-2)).all(1).mean()
        report("vs path trace", pixels=hit.sum(), within_1e3=ok, max_abs=np.abs(a - b).max())
        assert hit.sum() > 2000
        assert ok >= 0.97
    finally:
        renderer.bind_formulas([0, 1, 2, 3])


def test_loss_and_gradient_from_the_per_pixel_outputs(renderer):
    base = [radiance_scene(view=v, size=(64, 64)) for v in range(4)]
    target_rgb, _, _ = forward(renderer, apply_params(base, TRUTH))
    target = target_rgb.reshape(-1).copy()
    miss = np.abs(target_rgb).sum(1) == 0
    target.reshape(-1, 3)[miss, 0] = -1.0  # misses are left out of the loss
    # loss and gradient from the per-pixel outputs
    rng = np.random.default_rng(3)
    p0 = TRUTH.copy()
    p0[:3] *= 1 + rng.uniform(-0.03, 0.03, 3)
    p0[3:] = rng.uniform(-0.05, 0.05, 3)
    rgb, drgb, _ = forward(renderer, apply_params(base, p0))
    loss, grad, used, _ = renderer.radiance_adjoint(apply_params(base, p0), target=target)
    m = (np.abs(rgb).sum(1) > 0) & ~miss
    res = rgb[m].astype(np.float64) - target.reshape(-1, 3)[m]
    assert used == int(m.sum())
    assert loss == pytest.approx(float((res ** 2).sum()), rel=1e-5)
    exp = 2 * (res[:, :, None] * drgb[m]).sum(axis=(0, 1))
    assert np.allclose(grad, exp, rtol=1e-4, atol=1e-6 * np.abs(exp).max())


def adam(loss_grad, p0, lr, iters, decay=0.98):
    """Adam with a decaying step; returns (losses, best iterate)."""
    p, m, v = p0.copy(), np.zeros(6), np.zeros(6)
    losses, best = [], None
    for it in range(iters):
        l, g = loss_grad(p)
        losses.append(l)
        if l == min(losses):
            best = p.copy()
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        mh, vh = m / (1 - 0.9 ** (it + 1)), v / (1 - 0.999 ** (it + 1))
        p = p - lr * (decay ** it) * mh / (np.sqrt(vh) + 1e-12)
    return losses, best


def colour_target(renderer, views):
    rgb, _, _ = forward(renderer, apply_params(views, TRUTH))
    rgb[np.abs(rgb).sum(1) == 0, 0] = -1.0  # misses are left out of the loss
    return rgb.reshape(-1).copy()


def box_views(n=8, size=128):
    """Cameras at distance 1 around the Mandelbox (the bulb hidden), so that it fills the view."""
    out = []
    for v in range(n):
        u = radiance_scene(view=v, size=(size, size))
        u.instances[0].hidden = 1.0
        ang = 2 * np.pi * v / n
        el = 0.4 * np.sin(3 * ang)
        c = np.array([0.75, 0.0, 0.0])
        nd.camera_look_at(u, c + np.array([np.cos(ang) * np.cos(el), np.sin(el), np.sin(ang) * np.cos(el)]), c, 60.0)
        nd.set_pixel_epsilon(u, 60.0, 1.0)
        out.append(u)
    return out


def test_colour_descent_recovers_parameters(renderer):
    """Descent on the colour loss from a perturbed start: over the whole union (4 views), the
    Mandelbulb power and the camera come back; with views that frame the Mandelbox, its scale
    and min radius come back from a 1% error (in the union view the box covers few pixels)."""
    base = [radiance_scene(view=v, size=(64, 64)) for v in range(4)]
    target = colour_target(renderer, base)
    rng = np.random.default_rng(1)
    p0 = TRUTH.copy()
    p0[:3] *= 1 + rng.uniform(-0.03, 0.03, 3)
    p0[3:] = rng.uniform(-0.05, 0.05, 3)
    lg = lambda p: renderer.radiance_adjoint(apply_params(base, p), target=target)[:2]
    losses, best = adam(lg, p0, np.array([0.05, 0.01, 0.005, 0.01, 0.01, 0.01]), 60)
    norm = np.array([8, 1.5, 0.5, 1, 1, 1])
    e0, e1 = np.abs(p0 - TRUTH) / norm, np.abs(best - TRUTH) / norm
    print("\n[radiance recovery] union: loss %.4g -> %.4g; rel error %s -> %s" % (
        losses[0], min(losses), np.round(e0, 4), np.round(e1, 4)), flush=True)
    well = [0, 3, 4, 5]
    assert min(losses) < 0.2 * losses[0]
    assert e1[well].sum() < 0.5 * e0[well].sum()
    views = box_views()
    target = colour_target(renderer, views)
    p0 = TRUTH.copy()
    p0[1] *= 1.01
    p0[2] *= 0.99
    lg = lambda p: renderer.radiance_adjoint(apply_params(views, p), target=target)[:2]
    losses, best = adam(lg, p0, np.array([0, 0.001, 0.001, 0, 0, 0]), 300)
    e0, e1 = np.abs(p0 - TRUTH) / norm, np.abs(best - TRUTH) / norm
    print("[radiance recovery] framed Mandelbox: loss %.4g -> %.4g; scale %.4g -> %.4g, min radius %.4g -> %.4g"
          % (losses[0], min(losses), e0[1], e1[1], e0[2], e1[2]), flush=True)
    assert e1[1] < 0.2 * e0[1] and e1[2] < 0.2 * e0[2]


@pytest.mark.parametrize("edit,status", [
    (lambda u: setattr(u, "mc_enabled", 1.0), 1),                       # Monte Carlo on
    (lambda u: setattr(u, "debug_parts", 0.0), 1),                      # reflections on
    (lambda u: nd.enable_accel(u), 7),                                  # accelerator
    (lambda u: setattr(u.instances[0], "color_count", 2.0), 7),         # colour strips
])
def test_scope_errors(renderer, edit, status):
    u = radiance_scene(size=(32, 32))
    edit(u)
    with pytest.raises(nd.NdError) as e:
        renderer.radiance_adjoint([u])
    assert e.value.status == status
