"""GPU parity of depth of field in the wavefront path trace (cs_wf_generate's lens samples and
fs_wf_resolve's average, template.wgsl:5150-5182, 5373-5394) against the MC oracle.

The four-fractal scene focused at the median depth of its hits, with a 0.1 aperture and a
0.1 focus range, 8 lens samples per out-of-focus pixel (mc_inner_samples):
  direct only     Monte Carlo off, reflections off: the lens average per pixel against the
                  FP64 oracle's; pixels well inside the focus range equal the GPU's own
                  pinhole render bit for bit;
  full transport  MSE against the FP64 oracle at equal spp within twice the FP32 oracle's,
                  mean radiance within 3%, alone and with the primary hit cache and adaptive
                  sampling on;
  no aperture     DOF on with the aperture at DOF_MIN_APERTURE or the part bit off renders
                  the pinhole image bit for bit.
"""
import numpy as np
import pytest

import numericdream_amd as nd
import oracle

pytestmark = pytest.mark.gpu

PART_DOF = 1 << 13


def c3(w, h):
    return nd.resize(nd.config_uniforms(3, 0), w, h)


def trace(r, u, spp):
    W, H = int(u.resolution[0]), int(u.resolution[1])
    depth = np.zeros((H, W), np.float32)
    normal = np.zeros((H, W, 3), np.float32)
    color, depth, normal, st = r.path_trace(u, spp, depth=depth, normal=normal)
    return color, depth, st


def report(name, **kv):
    print("\n[dof] %s %s" % (name, " ".join("%s=%.6g" % (k, v) for k, v in kv.items())), flush=True)


def with_dof(u, focus, aperture=0.1, focus_range=0.1, inner=8):
    u.dof_enabled = 1.0
    u.aperture = aperture
    u.focus_distance = focus
    u.focus_range = focus_range
    u.mc_inner_samples = float(inner)
    return u


def focus_of(r, w, h):
    u = c3(w, h)
    u.mc_enabled = 0.0
    _, depth, _ = trace(r, u, 1)
    return float(np.median(depth[depth > 0])), depth


def test_dof_direct_only(renderer):
    w, h = 64, 36
    focus, depth = focus_of(renderer, w, h)
    base = c3(w, h)
    base.mc_enabled = 0.0
    base.debug_parts = float(int(base.debug_parts) | (1 << 6))  # reflections off
    pin, _, _ = trace(renderer, base, 1)
    u = with_dof(c3(w, h), focus)
    u.mc_enabled = 0.0
    u.debug_parts = base.debug_parts
    print("\n[case] dof direct only %dx%d focus %.4f" % (w, h, focus), flush=True)
    g, _, st = trace(renderer, u, 1)
    ref = oracle.render_mc(u, 1, prec=0)["rgb"]
    rel = np.abs(g - ref) / np.maximum(np.abs(ref), 1e-2)
    close = (rel <= 1e-2).all(-1).mean()
    # pixels whose centre ray lies well inside half the focus range keep the pinhole colour
    d = np.array([oracle.primary_dir(u, x, y) for y in range(h) for x in range(w)]).reshape(h, w, 3)
    ft = np.float32(focus) / np.maximum(d @ np.array(u.camera_forward[:]), 1e-4)
    probe = np.where(depth > 0, depth, float(u.max_dist))
    sharp = np.abs(probe - ft) < 0.05 - 2e-3
    report("direct", close_1pct=close, sharp=sharp.sum(), mean_gpu=g.mean(), mean_ref=ref.mean(),
           segments_per_sample=st.segments / st.paths)
    assert close >= 0.95
    assert sharp.sum() > 20 and np.array_equal(g[sharp], pin[sharp])
    assert st.segments > 1.5 * st.paths  # out-of-focus pixels trace several lens paths


@pytest.mark.parametrize("variant", ["plain", "cache", "adaptive"])
def test_dof_full_transport(renderer, variant):
    w, h, spp = 64, 36, 8
    focus, _ = focus_of(renderer, w, h)
    u = with_dof(c3(w, h), focus)
    if variant == "cache":
        u.primary_cache_key = 1.0
    if variant == "adaptive":
        u.adapt_enabled = 1.0
        u.adapt_row_width = float(w)
        u.adapt_min_samples = 4.0
        u.adapt_threshold = 0.05
    print("\n[case] dof full transport %s %dx%d spp %d" % (variant, w, h, spp), flush=True)
    g, _, st = trace(renderer, u, spp)
    r64 = oracle.render_mc(u, spp, prec=0)["rgb"]
    r32 = oracle.render_mc(u, spp, prec=1)["rgb"]
    mse = float(((g - r64) ** 2).mean())
    floor = float(((r32 - r64) ** 2).mean())
    report(variant, mse_gpu=mse, mse_fp32_oracle=floor, mean_gpu=g.mean(), mean_ref=r64.mean())
    assert mse <= 2.0 * floor + 1e-6
    assert abs(g.mean() - r64.mean()) <= 0.03 * r64.mean()


def test_dof_without_effective_aperture_is_pinhole(renderer):
    w, h = 48, 27
    u = c3(w, h)
    pin, _, _ = trace(renderer, u, 2)
    for case in ("min_aperture", "part_bit"):
        v = with_dof(c3(w, h), 4.0, aperture=0.0001 if case == "min_aperture" else 0.2)
        if case == "part_bit":
            v.debug_parts = float(int(v.debug_parts) | PART_DOF)
        g, _, _ = trace(renderer, v, 2)
        assert np.array_equal(g, pin), case
