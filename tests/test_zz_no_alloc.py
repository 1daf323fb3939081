"""nd_render_gbuffer_noalloc: the G-buffer pass with no device allocation, checked bitwise
against nd_render_gbuffer on the same inputs (two calls in a row, the second smaller, so the
shared scratch array carries the first call's contents into the second), at its ray-slot
limit, and for its rejections. Device memory reported free by the runtime does not change
across the calls of a context that never allocated."""
import numpy as np
import pytest

import numericdream_amd as nd
from numericdream_amd import abi

pytestmark = pytest.mark.gpu

FIELDS = ("depth", "normal", "position", "trap", "hit", "shadow", "evals")


def dev_gbuffer(u):
    return nd.GBuffer(int(u.resolution[0]), int(u.resolution[1]), device="cuda:0")


def same(a, b):
    x, y = a.numpy(), b.numpy()
    for k in FIELDS:
        assert np.array_equal(x[k], y[k]), k


def test_noalloc_equals_allocating_entry():
    import torch
    r = nd.Renderer(0)  # a context of its own: no allocating call has grown its buffers
    try:
        cases = [
            # 1080p Mandelbox camera 0 at 320x180, shadows of 128 steps, persistent passes
            (nd.resize(nd.config_uniforms(2, 0), 320, 180), dict(shadows=True, shadow_max_steps=128)),
            # a smaller frame of another camera next: the scratch still holds the first call's data
            (nd.resize(nd.config_uniforms(2, 5), 96, 54), dict(shadows=True, shadow_max_steps=0,
                                                                 normal_mode=abi.NORMAL_TETRA)),
            # the plain kernel through the same scratch
            (nd.resize(nd.config_uniforms(1, 0), 160, 90), dict(shadows=False, kernel=abi.KERNEL_PLAIN)),
        ]
        # The HIP runtime reserves private-segment (spill) memory at the first launch of a
        # kernel that spills, in this process, whichever entry point launches it: one warm-up
        # call per case comes before the free-memory check.
        for u, kw in cases:
            r.render_gbuffer_noalloc(u, dev_gbuffer(u), **kw)
        torch.cuda.synchronize()
        outs = []
        for i, (u, kw) in enumerate(cases):
            print("\n[case] noalloc call %d %dx%d %s" % (i, int(u.resolution[0]), int(u.resolution[1]), kw), flush=True)
            b = dev_gbuffer(u)
            torch.cuda.synchronize()
            free0 = torch.cuda.mem_get_info()[0]
            sb = r.render_gbuffer_noalloc(u, b, **kw)
            torch.cuda.synchronize()
            free1 = torch.cuda.mem_get_info()[0]
            assert free1 == free0, (free0, free1)
            outs.append((u, kw, b, sb))
        for u, kw, b, sb in outs:
            a = dev_gbuffer(u)
            sa = r.render_gbuffer(u, a, **kw)
            same(a, b)
            assert (sa.rays, sa.hits, sa.de_evals) == (sb.rays, sb.hits, sb.de_evals)
            assert sb.rays == int(u.resolution[0]) * int(u.resolution[1]) and a.numpy()["hit"].sum() > 0
    finally:
        r.close()


def test_noalloc_limit_and_rejections(renderer):
    # exactly ND_NOALLOC_MAX_RAYS ray slots: 128 x 64 tiles of 64 slots
    u = nd.resize(nd.config_uniforms(2, 3), 1024, 512)
    a, b = dev_gbuffer(u), dev_gbuffer(u)
    renderer.render_gbuffer(u, a, shadows=True, shadow_max_steps=128)
    renderer.render_gbuffer_noalloc(u, b, shadows=True, shadow_max_steps=128)
    same(a, b)
    # one tile column more is rejected before any launch
    big = nd.resize(nd.config_uniforms(2, 3), 1032, 512)
    with pytest.raises(nd.NdError) as e:
        renderer.render_gbuffer_noalloc(big, dev_gbuffer(big))
    assert e.value.status == abi.ND_E_INVALID and "ray slots" in str(e.value)
    # host outputs would need a staging allocation
    small = nd.resize(nd.config_uniforms(2, 0), 64, 32)
    with pytest.raises(nd.NdError) as e:
        renderer.render_gbuffer_noalloc(small, nd.GBuffer(64, 32))
    assert e.value.status == abi.ND_E_INVALID and "device pointers" in str(e.value)
    # the accelerator's grids are allocated
    acc = nd.enable_accel(nd.resize(nd.config_uniforms(2, 0), 64, 32), res=16, levels=3)
    with pytest.raises(nd.NdError) as e:
        renderer.render_gbuffer_noalloc(acc, dev_gbuffer(acc))
    assert e.value.status == abi.ND_E_UNSUPPORTED
    # after the rejections the entry point still renders, equal to the allocating one
    a, b = dev_gbuffer(small), dev_gbuffer(small)
    renderer.render_gbuffer(small, a)
    renderer.render_gbuffer_noalloc(small, b)
    same(a, b)
