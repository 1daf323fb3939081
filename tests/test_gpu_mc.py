"""GPU parity of the wavefront Monte Carlo path trace against the MC oracle, through the C ABI.

The four-fractal smooth-union scene with its 16-stop strips, point light and procedural sky,
reduced in resolution. Three levels:
  sky only       every path misses: colour is sky -> spectrum -> RGB with the same hashed hero
                 wavelength as the oracle, so FP32 GPU and FP32 oracle agree per pixel;
  direct only    no reflection and no indirect bounce: colour is a smooth function of the
                 hit point, compared per pixel;
  full transport MSE against the FP64 oracle at equal spp, bounded by a multiple of the MSE
                 of the FP32 oracle restatement against the same FP64 oracle (the floor set by
                 FP32 geometry steering the bounce rays), and the mean radiance per channel.
"""
import numpy as np
import pytest

import numericdream_amd as nd
import oracle

pytestmark = pytest.mark.gpu


def c3(w, h):
    return nd.resize(nd.config_uniforms(3, 0), w, h)


def trace(r, u, spp, **kw):
    W, H = int(u.resolution[0]), int(u.resolution[1])
    depth = np.zeros((H, W), np.float32)
    normal = np.zeros((H, W, 3), np.float32)
    color, depth, normal, st = r.path_trace(u, spp, depth=depth, normal=normal, **kw)
    return color, depth, normal, st


def report(name, **kv):
    print("\n[mc] %s %s" % (name, " ".join("%s=%.6g" % (k, v) for k, v in kv.items())), flush=True)


def test_sky_only(renderer):
    u = c3(64, 36)
    for i in range(4):
        u.instances[i].hidden = 1.0
    print("\n[case] mc sky only 64x36 spp 4", flush=True)
    g, depth, _, st = trace(renderer, u, 4)
    ref = oracle.render_mc(u, 4, prec=1)["rgb"]
    err = np.abs(g - ref)
    report("sky", max_abs=err.max(), mean=g.mean())
    assert np.all(depth == 0)
    assert np.allclose(g, ref, rtol=1e-4, atol=2e-6)
    assert st.paths == 64 * 36 * 4
    assert st.segments == st.paths  # one segment per path: straight to the sky


def test_direct_only(renderer):
    u = c3(64, 36)
    u.mc_enabled = 0
    u.debug_parts = float(int(u.debug_parts) | (1 << 6))  # reflections off
    print("\n[case] mc direct only 64x36 spp 1", flush=True)
    g, depth, normal, st = trace(renderer, u, 1)
    ref = oracle.render_mc(u, 1, prec=0)
    close = (np.abs(g - ref["rgb"]) <= 1e-3 * np.maximum(np.abs(ref["rgb"]), 1e-2)).all(-1).mean()
    hit = ref["depth"] > 0
    d_ok = (np.abs(depth - ref["depth"]) <= 1e-4 * np.maximum(ref["depth"], 1e-6))[hit].mean()
    report("direct", close=close, depth_ok=d_ok, mean_gpu=g.mean(), mean_ref=ref["rgb"].mean())
    assert close >= 0.95
    assert d_ok >= 0.98
    assert ((depth > 0) != hit).mean() < 0.005


def test_full_transport_mse(renderer):
    u = c3(64, 36)
    spp = 8
    print("\n[case] mc full 64x36 spp 8", flush=True)
    g, depth, normal, st = trace(renderer, u, spp)
    r64 = oracle.render_mc(u, spp, prec=0)["rgb"]
    r32 = oracle.render_mc(u, spp, prec=1)["rgb"]
    mse_gpu = float(((g - r64) ** 2).mean())
    mse_floor = float(((r32 - r64) ** 2).mean())
    m_gpu, m_ref = g.reshape(-1, 3).mean(0), r64.reshape(-1, 3).mean(0)
    report("full", mse_gpu=mse_gpu, mse_fp32_oracle=mse_floor, mean_gpu=m_gpu.mean(), mean_ref=m_ref.mean(),
           segments_per_path=st.segments / st.paths, de_per_path=st.de_evals / st.paths)
    assert np.isfinite(g).all()
    assert mse_gpu <= 2.0 * mse_floor + 1e-6
    assert np.allclose(m_gpu, m_ref, rtol=0.05, atol=1e-4)
    assert st.paths == 64 * 36 * spp
    assert st.segments >= st.paths


def test_deterministic_and_device_outputs(renderer):
    import torch
    u = c3(48, 27)
    a, da, na, _ = trace(renderer, u, 2)
    b, db, nb, _ = trace(renderer, u, 2)
    assert np.array_equal(a, b) and np.array_equal(da, db) and np.array_equal(na, nb)
    col = torch.zeros((27, 48, 3), dtype=torch.float32, device="cuda:0")
    dep = torch.zeros((27, 48), dtype=torch.float32, device="cuda:0")
    nor = torch.zeros((27, 48, 3), dtype=torch.float32, device="cuda:0")
    renderer.path_trace(u, 2, color=col, depth=dep, normal=nor)
    renderer.sync()
    assert np.array_equal(col.cpu().numpy(), a)
    assert np.array_equal(dep.cpu().numpy(), da)
    assert np.array_equal(nor.cpu().numpy(), na)


def test_rect_matches_full_frame(renderer):
    u = c3(48, 27)
    full, _, _, _ = trace(renderer, u, 2)
    sub, _, _, _ = renderer.path_trace(u, 2, rect=(8, 5, 24, 13))
    assert np.array_equal(sub, full[5:18, 8:32])


def test_material_features_are_accepted(renderer):
    """Transmissive, thin-film and subsurface materials are accepted; fog is rejected by name."""
    u = c3(16, 9)
    u.fog_count = 1.0
    with pytest.raises(nd.NdError) as e:
        renderer.path_trace(u, 1)
    assert e.value.status == nd.abi.ND_E_UNSUPPORTED and "fog" in str(e.value)
    u = c3(16, 9)
    u.instances[1].colors[2].subsurface = 0.5
    u.instances[1].colors[3].transparency = 0.5
    u.instances[1].colors[4].film_strength = 0.5
    u.has_transparency = 1.0
    assert np.isfinite(renderer.path_trace(u, 1)[0]).all()


def test_graph_replays_match_eager_launches(renderer):
    """Each chunk's passes are replayed from a captured hipGraph. Every replay, with other
    eager work and synchronisations in between, equals the launch-by-launch result; the
    second case has two chunks of unequal size (8 + 1 samples at 1080p)."""
    import torch
    t = renderer.tuning()
    cases = [(c3(64, 36), 3), (nd.config_uniforms(3, 0), 9)]
    try:
        for u, spp in cases:
            renderer.set_tuning(t.waves_per_cu, t.refill_min_idle, t.chunk_rays, graphs=False)
            eager, _, _, st_e = renderer.path_trace(u, spp)
            renderer.set_tuning(t.waves_per_cu, t.refill_min_idle, t.chunk_rays, graphs=True)
            for rep in range(3):
                g, _, _, st_g = renderer.path_trace(u, spp)
                assert np.array_equal(g, eager), (spp, rep)
                assert st_g.segments == st_e.segments and st_g.de_evals == st_e.de_evals
                # unrelated eager work on the same context and on torch's stream
                renderer.render_gbuffer(nd.resize(nd.config_uniforms(2, rep), 64, 36), nd.GBuffer(64, 36))
                x = torch.ones(1 << 20, device="cuda:0")
                torch.cuda.synchronize()
                assert float(x.sum()) == float(1 << 20)
    finally:
        renderer.set_tuning(t.waves_per_cu, t.refill_min_idle, t.chunk_rays)


def test_diffuse_sphere_closed_form(renderer):
    """The diffuse unit sphere under one point light with Monte Carlo indirect light off: the
    path trace equals (1 - F) albedo (0.15 AO + L n.l shadow) per pixel, with AO = shadow = 1
    on the convex sphere, the hit point and normal from the ray-sphere intersection, and
    Schlick's F at index 1 (tests/test_oracle_mc_pins.py pins the oracle the same way)."""
    from tests.test_oracle_mc_pins import diffuse_sphere_scene, sphere_closed_form
    u, cm = diffuse_sphere_scene()
    hit, depth, ndl, rad, cosi = sphere_closed_form(u)
    renderer.bind_formulas(cm)
    try:
        g, gdepth, _, st = trace(renderer, u, 4)
    finally:
        renderer.bind_formulas([0, 1, 2, 3])
    lit = hit & (ndl > 0.2) & (cosi > 0.3)
    dark = hit & (ndl == 0.0) & (cosi > 0.3)
    got = g.astype(np.float64)
    err = max(np.abs(got[..., k][lit | dark] / rad[lit | dark] - 1).max() for k in range(3))
    report("diffuse sphere", lit=lit.sum(), dark=dark.sum(), max_rel=err)
    assert lit.sum() > 100 and dark.sum() > 20
    eps = float(u.epsilon_coefficient) * depth[hit] + float(u.epsilon_floor)
    assert np.all(np.abs(gdepth[hit] - depth[hit]) <= 2 * eps / cosi[hit] + 1e-5)
    for k in range(3):
        assert np.allclose(got[..., k][lit], rad[lit], rtol=2e-3), k
        assert np.allclose(got[..., k][dark], rad[dark], rtol=2e-3), k
