"""GPU parity of the empty-space accelerator (grids built on the device, skips in the march
and shadow passes of both schedules and of the path trace) against the oracle's FP64
restatement of the same algorithm, which builds its own FP64 grid.

With the accelerator the march takes the reference's skip steps, so the surface it finds can
differ from the unaccelerated one (the grid's bound is not exact for a fractal DE); parity
is against the oracle with the accelerator on. FP32 and FP64 grids can round a skip across
eps(t) differently, which moves a few hits; the bars are stated per test.
"""
import numpy as np
import pytest

import numericdream_amd as nd
from numericdream_amd import abi
import oracle
from tests.parity import compare, pointwise

pytestmark = pytest.mark.gpu


def gpu_render(r, u, kernel, shadows=True, shadow_max_steps=128):
    W, H = int(u.resolution[0]), int(u.resolution[1])
    g = nd.GBuffer(W, H)
    st = r.render_gbuffer(u, g, normal_mode=abi.NORMAL_ANALYTIC, shadows=shadows, shadow_max_steps=shadow_max_steps,
                          kernel=kernel)
    return g.numpy(), st


def report(name, st):
    print("\n[accel] %s %s" % (name, " ".join("%s=%.6g" % (k, v) for k, v in st.items())), flush=True)


@pytest.mark.parametrize("cfg,cam", [(2, 0), (1, 0)])
def test_gbuffer_with_accelerator(renderer, cfg, cam):
    u = nd.resize(nd.config_uniforms(cfg, cam), 240, 135)
    nd.enable_accel(u, res=64, levels=6, safety=0.92)
    shadows = cfg == 2
    print("\n[case] accel config %d camera %d 240x135" % (cfg, cam), flush=True)
    ref = oracle.render_gbuffer(u, normal_mode=0, shadows=shadows, shadow_max_steps=128)
    for kernel in (abi.KERNEL_PERSISTENT, abi.KERNEL_PLAIN):
        g, _ = gpu_render(renderer, u, kernel, shadows=shadows)
        st = compare(g, ref)
        st.update(pointwise(g, u, 0))
        report("cfg%d kernel=%d" % (cfg, kernel), st)
        assert st["hit_agree_inner"] >= 0.99, st
        assert st["depth_rel_le_1e-4"] >= 0.95, st
        assert st["depth_rel_median"] <= 1e-5, st
        assert st["pw_trap_equal"] >= 0.99, st
        if shadows:
            assert st["shadow_abs_le_0.02"] >= 0.95, st


def test_schedules_identical_and_fewer_evaluations(renderer):
    u = nd.resize(nd.config_uniforms(2, 3), 200, 120)
    base, st0 = gpu_render(renderer, u, abi.KERNEL_PERSISTENT)
    nd.enable_accel(u, res=64, levels=6, safety=0.92)
    a, st1 = gpu_render(renderer, u, abi.KERNEL_PLAIN)
    b, st2 = gpu_render(renderer, u, abi.KERNEL_PERSISTENT)
    for k in ("depth", "normal", "position", "trap", "hit", "shadow", "evals"):
        assert np.array_equal(a[k], b[k]), k
    print("\n[accel] DE evaluations %d -> %d" % (st0.de_evals, st2.de_evals), flush=True)
    assert st2.de_evals < st0.de_evals
    # a second batch of the same scene reuses the cached grid and gives the same image
    c, _ = gpu_render(renderer, u, abi.KERNEL_PERSISTENT)
    assert np.array_equal(b["depth"], c["depth"])


def test_path_trace_with_accelerator(renderer):
    u = nd.resize(nd.config_uniforms(3, 0), 64, 36)
    nd.enable_accel(u, res=32, levels=5, safety=0.92)
    spp = 8
    g, _, _, st = renderer.path_trace(u, spp)
    r64 = oracle.render_mc(u, spp, prec=0)["rgb"]
    r32 = oracle.render_mc(u, spp, prec=1)["rgb"]
    mse_gpu = float(((g - r64) ** 2).mean())
    mse_floor = float(((r32 - r64) ** 2).mean())
    print("\n[accel] mc mse_gpu=%.4g floor=%.4g mean gpu %.4f ref %.4f" % (mse_gpu, mse_floor, g.mean(), r64.mean()),
          flush=True)
    assert mse_gpu <= 2.0 * mse_floor + 1e-6
    assert np.allclose(g.reshape(-1, 3).mean(0), r64.reshape(-1, 3).mean(0), rtol=0.05, atol=1e-4)
