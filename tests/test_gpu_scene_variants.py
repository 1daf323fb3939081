"""GPU parity of the scene features the five workloads leave at their defaults, against the
FP64 oracle through the C ABI:

- combine modes 2-6 of combine_de (template.wgsl:1708-1730), smooth and hard;
- orbit-trap shapes 0-4 under the max, average and last modes (template.wgsl:1227-1289);
- instances with an offset, a rotated frame, a non-uniform signed scale and a step safety
  below 1 (instance_local / instance_de_factor, template.wgsl:1600-1630);
- the central-difference normal (h = eps / 2) on the Mandelbulb workload.

Tolerances as in the G-buffer tests: hit mask away from silhouettes, depth within 1e-4
relative, and the oracle's normal and trap evaluated at the GPU's own hit point.
"""
import math

import numpy as np
import pytest

import numericdream_amd as nd
from numericdream_amd import abi
import oracle
from tests.parity import compare, meets, pointwise

pytestmark = pytest.mark.gpu


def render(r, u, kernel=abi.KERNEL_PERSISTENT, normal_mode=abi.NORMAL_ANALYTIC, shadows=False):
    g = nd.GBuffer(int(u.resolution[0]), int(u.resolution[1]))
    r.render_gbuffer(u, g, normal_mode=normal_mode, shadows=shadows, shadow_max_steps=0, kernel=kernel)
    return g.numpy()


def report(name, st):
    print("\n[parity] %s %s" % (name, " ".join("%s=%.6g" % (k, v) for k, v in st.items())), flush=True)


def rotation(axis, deg):
    """Columns of the rotation by deg about axis (Rodrigues)."""
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    t = math.radians(deg)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(t) * K + (1 - math.cos(t)) * K @ K


def set_frame(inst, R):
    for k in range(3):  # frame[k] = column k of R
        for j in range(3):
            inst.frame[k][j] = float(R[j, k])


def two_instances(mode, k, size=(128, 72)):
    """The Mandelbulb workload's camera over a Mandelbulb (offset -0.3, scale 0.8) and the
    Mandelbox of the 1080p workload (offset (0.3, 0.1, 0), scale 0.4) combined by `mode`."""
    u = nd.resize(nd.config_uniforms(1, 0), *size)
    box = nd.config_uniforms(2, 0).instances[0]
    u.instance_count = 2.0
    a = u.instances[0]
    a.offset[0] = -0.3
    for i in range(3):
        a.scale[i] = 0.8
    u.instances[1] = box
    b = u.instances[1]
    for i, v in enumerate((0.3, 0.1, 0.0)):
        b.offset[i] = v
    for i in range(3):
        b.scale[i] = 0.4
    b.combine_mode = float(mode)
    b.blend_k = float(k)
    return u


def check_against_oracle(r, u, name, depth_frac=0.98, normal_frac=0.9, trap_frac=0.98, min_hits=100,
                         normal_mode=abi.NORMAL_ANALYTIC):
    ref = oracle.render_gbuffer(u, normal_mode=normal_mode, shadows=False)
    out = {}
    for kernel in (abi.KERNEL_PERSISTENT, abi.KERNEL_PLAIN):
        g = render(r, u, kernel, normal_mode=normal_mode)
        st = compare(g, ref)
        st.update(pointwise(g, u, normal_mode, floor=True))
        report("%s kernel=%d" % (name, kernel), st)
        assert st["hits_ref"] >= min_hits, st
        assert st["hit_agree_inner"] >= 0.999, st
        assert st["depth_rel_le_1e-4"] >= depth_frac, st
        assert st["depth_rel_median"] <= 1e-5, st
        assert meets(st["pw_normal_deg_le_0.5"], normal_frac, st["floor_normal_deg_le_0.5"]), st
        assert meets(st["pw_trap_equal"], trap_frac, st["floor_trap_equal"]), st
        out[kernel] = g
    for key in ("hit", "depth", "normal", "trap"):
        assert np.array_equal(out[abi.KERNEL_PERSISTENT][key], out[abi.KERNEL_PLAIN][key]), key
    return out


@pytest.mark.parametrize("mode,k", [(2, 0.05), (3, 0.2), (4, 0.2), (5, 0.2), (6, 0.2), (2, 0.0), (3, 0.0)])
def test_combine_modes(renderer, mode, k):
    """Smooth max, smooth subtract, the linear and non-linear chamfers and the blend, and
    the hard forms of max and subtract (k <= 1e-4). The surfaces of max and subtract run
    through the inside of one fractal, where its DE follows a bounded orbit whose gradient
    FP32 cannot resolve as well; the floor of pointwise() states how well it can."""
    print("\n[case] combine mode %d k %g" % (mode, k), flush=True)
    u = two_instances(mode, k)
    check_against_oracle(renderer, u, "combine %d k=%g" % (mode, k), min_hits=150)


@pytest.mark.parametrize("shape", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("mode", [1, 2, 3])
def test_trap_shapes_and_modes(renderer, shape, mode):
    """Point, sphere, box, cross and torus traps under the max, average and last modes on
    the Mandelbox, with an off-centre trap."""
    print("\n[case] trap shape %d mode %d" % (shape, mode), flush=True)
    u = nd.resize(nd.config_uniforms(2, 0), 96, 54)
    t = u.instances[0]
    t.trap_shape = float(shape)
    t.trap_mode = float(mode)
    for i, v in enumerate((0.1, 0.2, -0.1)):
        t.trap_center[i] = v
    for i, v in enumerate((0.5, 0.7, 0.3)):
        t.trap_box[i] = v
    t.trap_radius = 0.8
    t.trap_tube = 0.25
    # The max and last modes follow the orbit to its last iterate, where the FP32 growth of z
    # is largest: there the bar is the FP32 restatement's own agreement with FP64.
    check_against_oracle(renderer, u, "trap %d/%d" % (shape, mode))


def test_transformed_instances(renderer):
    """Both instances rotated, the Mandelbox with a non-uniform scale of mixed sign and a step
    safety of 0.8, the Mandelbulb with a uniform scale: instance_local = R^T (p - offset) /
    scale, the DE scaled by min |scale| * step safety."""
    u = two_instances(1, 0.1)
    a, b = u.instances[0], u.instances[1]
    set_frame(a, rotation((0.2, 1.0, 0.3), 35.0))
    set_frame(b, rotation((1.0, 1.0, 0.0), -50.0))
    for i, v in enumerate((0.45, -0.35, 0.4)):
        b.scale[i] = v
    b.step_safety = 0.8
    out = check_against_oracle(renderer, u, "transformed")
    g = out[abi.KERNEL_PERSISTENT]
    # the transform moved the surface: the untransformed scene differs on many pixels
    plain = render(renderer, two_instances(1, 0.1))
    assert (plain["hit"] != g["hit"]).mean() > 0.02


def test_central_normal_c1(renderer):
    """The Mandelbulb workload's central-difference normal with h = eps(t) / 2."""
    u = nd.resize(nd.config_uniforms(1, 0), 160, 90)
    check_against_oracle(renderer, u, "c1 central", normal_frac=0.75, trap_frac=0.995,
                         normal_mode=abi.NORMAL_CENTRAL)
