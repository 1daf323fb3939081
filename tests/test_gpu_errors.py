"""Error paths of the C ABI that a host can force without any kernel misbehaving: argument
checks before launch, the cancel flag, the round cap of the persistent passes (every wave
stops after a few rounds and exits normally) and the features the library rejects so the
host can fall back to its WGSL pass (nd_status, include/nd_hip.h; the reference's error
sink is src/gpu/webgpu_context.zig:131-146)."""
import ctypes as C

import numpy as np
import pytest

import numericdream_amd as nd
from numericdream_amd import abi

pytestmark = pytest.mark.gpu


def small_c2():
    return nd.resize(nd.config_uniforms(2, 0), 96, 54)


def raw_gbuffer(r, u, size, out):
    opts = abi.MarchOpts(abi.NORMAL_ANALYTIC, 0, 1, abi.KERNEL_PERSISTENT)
    st = abi.BatchStats()
    return nd.lib.nd_render_gbuffer(r._h, C.byref(u), size, C.byref(opts), abi.Rect(0, 0, 0, 0),
                                    C.byref(out) if out is not None else None, C.byref(st))


def test_argument_checks(renderer):
    u = small_c2()
    g = nd.GBuffer(96, 54)
    out = g.out()
    assert raw_gbuffer(renderer, u, abi.UNIFORMS_SIZE - 4, out) == abi.ND_E_INVALID
    assert "uniforms_size" in nd.lib.nd_last_error(renderer._h).decode()
    assert raw_gbuffer(renderer, u, abi.UNIFORMS_SIZE, None) == abi.ND_E_INVALID
    # the path trace needs its colour output
    img = abi.ImageOut(None, None, None, 0, 0)
    ts = abi.TraceStats()
    st = nd.lib.nd_path_trace(renderer._h, C.byref(u), abi.UNIFORMS_SIZE, 1, abi.Rect(0, 0, 0, 0), C.byref(img),
                              C.byref(ts))
    assert st == abi.ND_E_INVALID
    assert "colour" in nd.lib.nd_last_error(renderer._h).decode()
    # a rect outside the target
    with pytest.raises(nd.NdError) as e:
        renderer.render_gbuffer(u, nd.GBuffer(8, 8), rect=(90, 50, 8, 8))
    assert e.value.status == abi.ND_E_INVALID
    # a device index that does not exist
    h = C.c_void_p()
    assert nd.lib.nd_ctx_create(1 << 20, C.byref(h)) == abi.ND_E_INVALID
    # tuning ranges
    with pytest.raises(nd.NdError) as e:
        renderer.set_tuning(refill_min_idle=0)
    assert e.value.status == abi.ND_E_INVALID
    for k in (8, 28):
        with pytest.raises(nd.NdError) as e:
            renderer.set_tuning(mc_pool_log2=k)
        assert e.value.status == abi.ND_E_INVALID
    renderer.set_tuning()
    # after every rejected call the context still renders
    st = renderer.render_gbuffer(u, g)
    assert st.rays == 96 * 54 and g.hit.sum() > 0


def test_unsupported_features_and_formulas(renderer):
    u = small_c2()
    u.warp_count = 1.0
    with pytest.raises(nd.NdError) as e:
        renderer.render_gbuffer(u, nd.GBuffer(96, 54))
    assert e.value.status == abi.ND_E_UNSUPPORTED and "warp" in str(e.value)
    u = nd.resize(nd.config_uniforms(3, 0), 32, 18)
    u.fog_count = 1.0
    with pytest.raises(nd.NdError) as e:
        renderer.path_trace(u, 1)
    assert e.value.status == abi.ND_E_UNSUPPORTED
    u = small_c2()
    u.instances[0].code = 9.0  # no built-in bound to code 9
    with pytest.raises(nd.NdError) as e:
        renderer.render_gbuffer(u, nd.GBuffer(96, 54))
    assert e.value.status == abi.ND_E_UNSUPPORTED_FORMULA


def test_cancel_flag(renderer):
    u = small_c2()
    g = nd.GBuffer(96, 54)
    flag = C.c_int(1)
    renderer.set_cancel(flag)
    try:
        with pytest.raises(nd.NdError) as e:
            renderer.render_gbuffer(u, g)
        assert e.value.status == abi.ND_E_CANCELLED
        with pytest.raises(nd.NdError) as e:
            renderer.path_trace(nd.resize(nd.config_uniforms(3, 0), 32, 18), 4)
        assert e.value.status == abi.ND_E_CANCELLED
        flag.value = 0
        renderer.render_gbuffer(u, g)
        assert g.hit.sum() > 0
    finally:
        renderer.set_cancel(None)


def test_round_cap_times_out_and_recovers(renderer):
    """A round cap of 3 stops every wave of the march and shadow passes early: the call
    reports ND_E_TIMEOUT, and with the cap lifted the next call gives the full result."""
    u = small_c2()
    ref = nd.GBuffer(96, 54)
    renderer.render_gbuffer(u, ref, shadows=True, shadow_max_steps=128)
    g = nd.GBuffer(96, 54)
    renderer.set_tuning(round_cap=3)
    try:
        with pytest.raises(nd.NdError) as e:
            renderer.render_gbuffer(u, g, shadows=True, shadow_max_steps=128)
        assert e.value.status == abi.ND_E_TIMEOUT and "round cap" in str(e.value)
    finally:
        renderer.set_tuning()
    g = nd.GBuffer(96, 54)
    renderer.render_gbuffer(u, g, shadows=True, shadow_max_steps=128)
    a, b = g.numpy(), ref.numpy()
    for k in ("hit", "depth", "normal", "trap", "shadow"):
        assert np.array_equal(a[k], b[k]), k
