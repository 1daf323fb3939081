"""GPU parity of the transmissive path trace (refraction, the inside march, dispersion,
shadows through transparent occluders, thin-film coats, the light pick) against the MC
oracle, through the C ABI.

Scenes: the four-fractal scene with its Mandelbox (the instance that fills the view; the
Mandelbulb lies inside the union) made of dispersive glass (the strips keep their colours,
transparency 0.9, index 1.5, Abbe number 40), with the default
soft dispersion weights and with the one-wavelength collapse; the same scene lit by two
shadowed lights with the Monte Carlo light pick; and a clear sphere of index 1, whose centre
pixel must show the sky behind it. Full transport is compared like the opaque path trace:
MSE against the FP64 oracle at equal spp within twice the FP32 oracle restatement's own MSE,
and the mean radiance per channel within 3%.
"""
import numpy as np
import pytest

import numericdream_amd as nd
import oracle
from tests import formula_table as ft

pytestmark = pytest.mark.gpu


def glass_c3(w=64, h=36, soft=1.0, abbe=40.0):
    u = nd.resize(nd.config_uniforms(3, 0), w, h)
    inst = u.instances[1]
    for s in range(int(inst.color_count)):
        c = inst.colors[s]
        c.transparency, c.ior, c.abbe, c.inner_max_steps, c.reflectiveness = 0.9, 1.5, abbe, 64.0, 0.08
    u.has_transparency = 1.0
    u.photon_dispersion_soft = soft
    return u


def two_lights(u):
    u.light_count = 2.0
    L0, L1 = u.lights[0], u.lights[1]
    for name, _ in L0._fields_:
        setattr(L1, name, getattr(L0, name))
    L1.light_type = 0.0
    for k, v in enumerate((-3.0, 2.5, -2.0)):
        L1.position_or_direction[k] = v
    for k, v in enumerate((0.4, 0.6, 1.0)):
        L1.color[k] = v
    L1.brightness = 12.0
    L1.cast_shadows = 1.0
    u.mc_light_pick = 1.0
    u.mc_enabled = 1.0
    return u


def compare_full(r, u, spp, name, code_map=None):
    print("\n[case] transmission %s %dx%d spp %d" % (name, int(u.resolution[0]), int(u.resolution[1]), spp), flush=True)
    g, depth, normal, st = r.path_trace(u, spp)
    r64 = oracle.render_mc(u, spp, prec=0, code_map=code_map)["rgb"]
    r32 = oracle.render_mc(u, spp, prec=1, code_map=code_map)["rgb"]
    mse_gpu = float(((g - r64) ** 2).mean())
    mse_floor = float(((r32 - r64) ** 2).mean())
    m_gpu, m_ref = g.reshape(-1, 3).mean(0), r64.reshape(-1, 3).mean(0)
    print("[mc] %s mse_gpu=%.4g mse_fp32_oracle=%.4g mean_gpu=%.5g mean_ref=%.5g segments_per_path=%.3f"
          % (name, mse_gpu, mse_floor, m_gpu.mean(), m_ref.mean(), st.segments / st.paths), flush=True)
    assert np.isfinite(g).all()
    assert mse_gpu <= 2.0 * mse_floor + 1e-6
    assert np.all(np.abs(m_gpu - m_ref) <= 0.03 * m_ref + 1e-4)
    return st


def test_dispersive_glass_soft_weights(renderer):
    st = compare_full(renderer, glass_c3(), 8, "glass soft")
    assert st.segments / st.paths > 1.5  # paths go through the glass


def test_dispersive_glass_collapse(renderer):
    compare_full(renderer, glass_c3(soft=0.0), 8, "glass collapse")


def filmed_c3(w=64, h=36, glass=False):
    """The Mandelbox with an iridescent coat: film index 1.8, 350 nm, strength 0.8, its
    normal bent by the film noise; reflective enough that the film tints the reflections."""
    u = glass_c3(w, h) if glass else nd.resize(nd.config_uniforms(3, 0), w, h)
    inst = u.instances[1]
    for s in range(int(inst.color_count)):
        c = inst.colors[s]
        c.film_strength, c.film_ior, c.film_thickness, c.film_angle_scale = 0.8, 1.8, 350.0, 1.0
        c.film_perturb, c.film_perturb_scale = 0.3, 0.25
        c.reflectiveness = max(c.reflectiveness, 0.3)
    return u


def test_thin_film(renderer):
    compare_full(renderer, filmed_c3(), 8, "thin film")


def test_thin_film_on_glass(renderer):
    compare_full(renderer, filmed_c3(glass=True), 8, "thin film on glass")


def test_light_pick(renderer):
    compare_full(renderer, two_lights(nd.resize(nd.config_uniforms(3, 0), 64, 36)), 8, "light pick")


def test_glass_and_light_pick(renderer):
    compare_full(renderer, two_lights(glass_c3()), 8, "glass + light pick")


@pytest.fixture
def sphere_bound(renderer):
    renderer.bind_formulas([ft.ID["sphere_primitive"]])
    yield renderer
    renderer.bind_formulas([0, 1, 2, 3])


def clear_sphere(hidden):
    u, cm = ft.scene("sphere_primitive", params=[1.0], cam=(0.0, 3.0, -2.5), size=(9, 9))
    inst = u.instances[0]
    inst.color_count = 1.0
    c = inst.colors[0]
    for k in range(3):
        c.color[k] = 1.0
    c.transparency, c.reflectiveness, c.ior, c.abbe, c.roughness, c.inner_max_steps = 1.0, 0.0, 1.0, 0.0, 0.0, 128.0
    u.has_transparency = 1.0
    u.light_count = 0.0
    u.mc_enabled = 0.0
    inst.hidden = 1.0 if hidden else 0.0
    return u


def test_index_one_clear_glass_is_invisible_on_axis(sphere_bound):
    """The centre ray enters and leaves a clear index-1 sphere at normal incidence: the
    centre pixel equals the render with the sphere hidden."""
    print("\n[case] transmission clear sphere 9x9 spp 16", flush=True)
    g = sphere_bound.path_trace(clear_sphere(False), 16)[0]
    h = sphere_bound.path_trace(clear_sphere(True), 16)[0]
    assert np.allclose(g[4, 4], h[4, 4], rtol=1e-4, atol=1e-6), (g[4, 4], h[4, 4])
    assert not np.allclose(g, h, rtol=1e-4)  # off axis the sphere shows


def test_primary_hit_cache_semantics(renderer):
    """With primary_cache_key the primary normals are octahedral-packed, as the oracle's."""
    u = nd.resize(nd.config_uniforms(3, 0), 64, 36)
    u.primary_cache_key = 1.0
    compare_full(renderer, u, 8, "primary cache")


def test_primary_hit_cache_reuse_is_exact(renderer):
    """At 1080p, 9 samples run as chunks of 8 + 1; the ninth sample reuses the primary
    segments stored by sample 0 instead of marching. It must equal that sample rendered on
    its own (mc_sample 8, one chunk, nothing reused): 9 c9 - 8 c8 = c(sample 8)."""
    import math
    u = nd.config_uniforms(3, 0)
    u.primary_cache_key = 1.0
    print("\n[case] transmission primary cache reuse 1080p spp 9", flush=True)
    c9, _, _, st9 = renderer.path_trace(u, 9)
    c8, _, _, _ = renderer.path_trace(u, 8)
    u1 = nd.config_uniforms(3, 0)
    u1.primary_cache_key = 1.0
    u1.mc_sample = 8.0
    c1, _, _, st1 = renderer.path_trace(u1, 1)
    last = c9.astype(np.float64) * 9 - c8.astype(np.float64) * 8
    err = np.abs(last - c1)
    print("[mc] cache reuse max_err=%.3g mean=%.5g de_per_path_9=%.1f de_per_path_1=%.1f"
          % (err.max(), c1.mean(), st9.de_evals / st9.paths, st1.de_evals / st1.paths), flush=True)
    assert err.max() <= 1e-5 * max(1.0, float(np.abs(c1).max())) * 20
    assert st9.de_evals / st9.paths < st1.de_evals / st1.paths  # the reused samples skip the primary march


def adaptive_c3(threshold, min_samples, w=64, h=36):
    u = nd.resize(nd.config_uniforms(3, 0), w, h)
    u.adapt_enabled = 1.0
    u.adapt_threshold = threshold
    u.adapt_min_samples = min_samples
    u.adapt_row_width = float(w)
    return u


def test_adaptive_sampling(renderer):
    """Adaptive sampling at threshold 0.05 from 4 samples, 16 spp: parity like the full
    transport (converged pixels keep the mean of the samples they took)."""
    compare_full(renderer, adaptive_c3(0.05, 4.0), 16, "adaptive")


def test_adaptive_limit_and_progressive_calls(renderer):
    """An unbounded threshold stops every pixel at its min-samples count (the plain 5-sample
    image); sixteen 1-sample calls with increasing mc_sample keep the per-pixel statistics
    in the context and end at the 16-sample call's image."""
    u = adaptive_c3(1e9, 5.0)
    a = renderer.path_trace(u, 12)[0]
    u.adapt_enabled = 0.0
    b = renderer.path_trace(u, 5)[0]
    assert np.allclose(a, b, rtol=2e-6, atol=1e-8)
    u = adaptive_c3(0.05, 4.0)
    one = renderer.path_trace(u, 16)[0]
    for i in range(16):
        u.mc_sample = float(i)
        prog = renderer.path_trace(u, 1)[0]
    print("\n[mc] adaptive progressive max_err=%.3g" % np.abs(prog - one).max(), flush=True)
    assert np.allclose(prog, one, rtol=2e-6, atol=1e-8)
