"""GPU parity of the @image formula (formulas/heightmap_primitive.wgsl) against the FP64 oracle.

The image is bound to the library (nd_bind_image) and to the oracle (the same packed words,
from nd_image_pack). The test image has smooth colour ramps, a bump, and a transparent disk
(premultiplied: its texels read 0, so the height drops there, with steep slopes at its rim).
Compared like the other formulas: hit mask away from silhouettes, depth within 1e-4
relative, and the oracle's normal and trap at the GPU's own hit points, for the analytic
normal (dual numbers through the bilinear lookup) and the reference's tetrahedral one.
"""
import numpy as np
import pytest

import numericdream_amd as nd
from numericdream_amd import abi
import oracle
from tests import formula_table as ft
from tests.parity import compare, pointwise

pytestmark = pytest.mark.gpu

HM = ft.ID["heightmap_primitive"]


def height_image(w=64, h=48):
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    u, v = x / (w - 1) * 2 - 1, y / (h - 1) * 2 - 1
    bump = np.clip(1.0 - 1.6 * (u * u + v * v), 0, 1)
    img = np.zeros((h, w, 4))
    img[..., 0] = 0.3 + 0.7 * bump
    img[..., 1] = 0.5 + 0.4 * u
    img[..., 2] = 0.5 + 0.5 * np.sin(3.0 * v)
    img[..., 3] = 1.0
    hole = (u - 0.55) ** 2 + (v + 0.3) ** 2 < 0.06
    img[hole, 3] = 0.0
    return np.round(np.clip(img, 0, 1) * 65535).astype(np.uint16)


@pytest.fixture
def hm(renderer):
    """The renderer with code 0 (and 1) bound to the heightmap formula and the test image in
    slot 0 of both the library and the oracle; everything unbound after."""
    renderer.bind_formulas([HM, HM])
    img = height_image()
    renderer.bind_image(0, img)
    words, w, h, slope = nd.image_pack(img)
    oracle.bind_image(0, words, w, h, slope.astype(np.float64))
    yield renderer
    renderer.bind_image(0, None)
    renderer.bind_image(4, None)
    oracle.bind_image(0, None)
    oracle.bind_image(4, None)
    renderer.bind_formulas([0, 1, 2, 3])


def scene(size=(96, 72), params=None):
    return ft.scene("heightmap_primitive", params=params, cam=(0.0, 1.5, -1.6), size=size)


def gpu_render(r, u, normal_mode=abi.NORMAL_ANALYTIC, kernel=abi.KERNEL_PERSISTENT):
    W, H = int(u.resolution[0]), int(u.resolution[1])
    g = nd.GBuffer(W, H)
    r.render_gbuffer(u, g, normal_mode=normal_mode, shadows=True, shadow_max_steps=0, kernel=kernel)
    return g.numpy()


def report(name, st):
    print("\n[parity] %s %s" % (name, " ".join("%s=%.6g" % (k, v) for k, v in st.items())), flush=True)


@pytest.mark.parametrize("mode", [abi.NORMAL_ANALYTIC, abi.NORMAL_TETRA])
def test_heightmap_gbuffer(hm, mode):
    print("\n[case] heightmap normal mode %d" % mode, flush=True)
    u, cm = scene()
    ref = oracle.render_gbuffer(u, normal_mode=mode, shadows=True, code_map=cm)
    g = gpu_render(hm, u, normal_mode=mode)
    st = compare(g, ref)
    st.update(pointwise(g, u, mode, code_map=cm))
    report("heightmap mode %d" % mode, st)
    assert st["hits_ref"] > 0.3 * g["hit"].size, st
    assert st["hit_agree_inner"] >= 0.999, st
    assert st["depth_rel_le_1e-4"] >= 0.99, st
    assert st["pw_normal_deg_le_0.5"] >= 0.97, st
    assert st["pw_trap_equal"] >= 0.99, st
    assert st["shadow_abs_le_0.02"] >= 0.97, st
    b = gpu_render(hm, u, normal_mode=mode, kernel=abi.KERNEL_PLAIN)
    for k in ("hit", "depth", "normal", "trap", "shadow"):
        assert np.array_equal(g[k], b[k]), k


def test_two_instances_read_their_own_slots(hm):
    """Instance 1 samples slot 4 (instance * 4): a second image there, against the oracle."""
    u, cm = scene()
    u.instance_count = 2.0
    u.instances[1] = u.instances[0]
    u.instances[1].code = 1.0
    u.instances[0].offset[0] = -0.6
    u.instances[1].offset[0] = 0.6
    for i in range(2):
        for k in range(3):
            u.instances[i].scale[k] = 0.5
    img2 = np.flip(height_image(40, 40), axis=1).copy()
    hm.bind_image(4, img2)
    words, w, h, slope = nd.image_pack(img2)
    oracle.bind_image(4, words, w, h, slope.astype(np.float64))
    ref = oracle.render_gbuffer(u, normal_mode=0, shadows=True, code_map=[HM, HM])
    g = gpu_render(hm, u)
    st = compare(g, ref)
    report("heightmap two slots", st)
    assert st["hit_agree_inner"] >= 0.999 and st["depth_rel_le_1e-4"] >= 0.99, st
    # without the second image the right instance is a flat slab: the renders differ there
    hm.bind_image(4, None)
    g2 = gpu_render(hm, u)
    assert not np.array_equal(g["depth"], g2["depth"])


def test_unbound_slot_is_the_flat_base(hm):
    hm.bind_image(0, None)
    oracle.bind_image(0, None)
    u, cm = scene()
    ref = oracle.render_gbuffer(u, normal_mode=0, shadows=False, code_map=cm)
    g = gpu_render(hm, u)
    st = compare(g, ref)
    report("heightmap unbound", st)
    assert st["hits_ref"] > 0 and st["hit_agree_inner"] >= 0.999 and st["depth_rel_le_1e-4"] >= 0.99, st


def test_heightmap_path_trace(hm):
    """The wavefront path trace of the height field: MSE against the FP64 oracle at equal spp
    within twice the FP32 oracle's own, mean radiance within 3%."""
    u = nd.resize(nd.config_uniforms(3, 0), 48, 36)  # the four-fractal scene's camera, lights and sky
    for i in range(1, 4):
        u.instances[i].hidden = 1.0
    inst = u.instances[0]
    inst.code = 0.0
    inst.mixin_count = 0.0
    for k, v in enumerate(ft.defaults("heightmap_primitive")):
        inst.params[k] = v
    for k in range(3):
        inst.scale[k] = 1.5
    cm = [HM]
    spp = 8
    col, depth, normal, st = hm.path_trace(u, spp)
    r64 = oracle.render_mc(u, spp, prec=0, code_map=cm)["rgb"]
    r32 = oracle.render_mc(u, spp, prec=1, code_map=cm)["rgb"]
    mse = float(((col - r64) ** 2).mean())
    floor = float(((r32 - r64) ** 2).mean())
    print("\n[mc] heightmap mse_gpu=%.4g mse_fp32_oracle=%.4g mean_gpu=%.5g mean_ref=%.5g"
          % (mse, floor, col.mean(), r64.mean()), flush=True)
    assert mse <= 2.0 * floor + 1e-6
    assert abs(col.mean() - r64.mean()) <= 0.03 * r64.mean()
