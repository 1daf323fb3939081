"""GPU parity of every formula plug-in beyond the four tuned ones, against the FP64 oracle.

Each formula renders its test scene (tests/formula_table.py: @param defaults, a camera on
-z, epsilon 0.05 pixel) at 64x48 with the analytic normal and the light-0 soft shadow,
through the kernels that carry the extended formula set. Compared like the tuned formulas:
hit mask away from silhouettes, depth within 1e-4 relative, and the oracle's normal and
trap at the GPU's own hit points. Then the same kernels under the other schedules: the
plain kernel, a batch that mixes extended and tuned formulas, the accelerator and the path
trace.
"""
import numpy as np
import pytest

import numericdream_amd as nd
from numericdream_amd import abi
import oracle
from tests import formula_table as ft
from tests.parity import compare, pointwise

pytestmark = pytest.mark.gpu

# Formulas whose iteration amplifies FP32 rounding of z (chaotic folds, lattice wraps,
# spherical power maps) or whose fine structure sits at the hit epsilon: the share of
# pixels whose depth meets 1e-4 is lower there (hit decisions near thin features flip).
DEPTH_FRAC = {n: 0.95 for n in ft.FILES}
# The lattice wrap (round half to even of z / cell) makes the DE jump at cell faces, so a
# march that differs by an FP32 rounding can land on another face, as for the built-in
# Kleinian; the DE itself agrees pointwise (the normal and trap checks below).
DEPTH_FRAC["pseudokleinian_kleinian"] = 0.75
NORMAL_FRAC = {n: 0.85 for n in ft.FILES}


@pytest.fixture
def bound(renderer):
    """The renderer with code 0 bound to a formula; the default binding restored after."""
    def bind(ids):
        renderer.bind_formulas(ids)
        return renderer
    yield bind
    renderer.bind_formulas([0, 1, 2, 3])


def gpu_render(r, u, kernel=abi.KERNEL_PERSISTENT, shadows=True, normal_mode=abi.NORMAL_ANALYTIC):
    W, H = int(u.resolution[0]), int(u.resolution[1])
    g = nd.GBuffer(W, H)
    r.render_gbuffer(u, g, normal_mode=normal_mode, shadows=shadows, shadow_max_steps=0, kernel=kernel)
    return g.numpy()


def report(name, st):
    print("\n[parity] %s %s" % (name, " ".join("%s=%.6g" % (k, v) for k, v in st.items())), flush=True)


@pytest.mark.parametrize("name", ft.FILES)
def test_formula_gbuffer(bound, name):
    print("\n[case] formula %s" % name, flush=True)
    u, cm = ft.test_scene(name)
    r = bound(cm)
    ref = oracle.render_gbuffer(u, normal_mode=0, shadows=True, code_map=cm)
    g = gpu_render(r, u)
    st = compare(g, ref)
    if st["hits_ref"] and st["hits_gpu"]:
        st.update(pointwise(g, u, 0, code_map=cm))
    report(name, st)
    assert st["hits_ref"] > 0 and st["hits_gpu"] > 0, st
    assert st["hit_agree_inner"] >= 0.99, st
    assert st["depth_rel_le_1e-4"] >= DEPTH_FRAC[name], st
    assert st["pw_trap_equal"] >= 0.9, st
    assert st["pw_normal_deg_le_0.5"] >= NORMAL_FRAC[name], st
    assert st["shadow_abs_le_0.02"] >= 0.9, st


@pytest.mark.parametrize("name", ["cube_primitive", "juliabox_box_julia", "quatjulia_bulb_julia", "tree_ifs"])
def test_plain_equals_persistent(bound, name):
    """The plain kernel and the persistent passes run the same arithmetic per ray."""
    u, cm = ft.test_scene(name)
    r = bound(cm)
    a = gpu_render(r, u, abi.KERNEL_PERSISTENT)
    b = gpu_render(r, u, abi.KERNEL_PLAIN)
    for k in ("hit", "depth", "normal", "trap", "shadow"):
        assert np.array_equal(a[k], b[k]), k


def test_batch_mixes_extended_and_tuned(bound):
    """One persistent batch of jobs whose scenes use extended and tuned formulas: every job
    equals its own single-job render."""
    names = ["sphere_primitive", "mandelboxsort_box_fold", "juliabulb_bulb_julia", "kleinianlimit_kleinian"]
    ids = [ft.ID[n] for n in names] + [0]
    r = bound(ids)
    jobs, singles = [], []
    for code, n in enumerate(names):
        u, _ = ft.test_scene(n, size=(40, 24))
        u.instances[0].code = float(code)
        jobs.append(u)
    mb = nd.resize(nd.config_uniforms(1, 0), 40, 24)
    mb.instances[0].code = 4.0  # the Mandelbulb through code 4
    jobs.append(mb)
    gb = [nd.GBuffer(int(u.resolution[0]), int(u.resolution[1])) for u in jobs]
    r.render_batch(list(zip(jobs, gb)), normal_mode=abi.NORMAL_ANALYTIC, shadows=True)
    for u, g in zip(jobs, gb):
        one = gpu_render(r, u)
        got = g.numpy()
        for k in ("hit", "depth", "normal", "trap", "shadow"):
            assert np.array_equal(got[k], one[k]), k


@pytest.mark.parametrize("name", ["cube_primitive", "mandelbulbtwist_bulb_twist", "juliabox_box_julia"])
def test_extended_formula_with_accelerator(bound, name):
    """The empty-space accelerator over an extended formula, against the accelerated oracle."""
    u, cm = ft.test_scene(name)
    nd.enable_accel(u, res=32, levels=4, safety=0.92)
    r = bound(cm)
    ref = oracle.render_gbuffer(u, normal_mode=0, shadows=True, code_map=cm)
    g = gpu_render(r, u)
    st = compare(g, ref)
    report(name + " accel", st)
    assert st["hit_agree_inner"] >= 0.99, st
    assert st["depth_rel_le_1e-4"] >= 0.95, st


def test_extended_formula_path_trace(bound):
    """The wavefront path trace of the four-fractal scene with its first instance replaced by
    an extended formula (the others hidden, the instance scaled by 3): MSE against the FP64
    oracle at equal spp within twice the FP32 oracle's own MSE, mean radiance within 3%."""
    u = nd.resize(nd.config_uniforms(3, 0), 48, 36)
    for i in range(1, 4):
        u.instances[i].hidden = 1.0
    inst = u.instances[0]
    inst.code = 0.0
    for k, v in enumerate(ft.defaults("hybridkifs_hybrid_ifs")):
        inst.params[k] = v
    for k in range(3):
        inst.scale[k] = 3.0
    cm = [ft.ID["hybridkifs_hybrid_ifs"]]
    r = bound(cm)
    spp = 8
    g, depth, normal, st = r.path_trace(u, spp)
    r64 = oracle.render_mc(u, spp, prec=0, code_map=cm)["rgb"]
    r32 = oracle.render_mc(u, spp, prec=1, code_map=cm)["rgb"]
    hit = oracle.render_gbuffer(u, shadows=False, code_map=cm)["hit"] > 0
    mse_gpu = float(((g - r64) ** 2).mean())
    mse_floor = float(((r32 - r64) ** 2).mean())
    m_gpu, m_ref = float(g.mean()), float(r64.mean())
    print("\n[mc] hybridkifs hit=%.3f mse_gpu=%.4g mse_fp32_oracle=%.4g mean_gpu=%.5g mean_ref=%.5g"
          % (hit.mean(), mse_gpu, mse_floor, m_gpu, m_ref), flush=True)
    assert hit.mean() > 0.1
    assert mse_gpu <= 2.0 * mse_floor + 1e-6
    assert abs(m_gpu - m_ref) <= 0.03 * m_ref


# The library rejects cycles of two or more mix-ins: a render of this scene did not finish
# on the GPU and no cause was found (a host run of the same per-ray code finished and
# matched the oracle).
STALLED = "two mix-ins switched off: a GPU render of this scene did not finish, cause not found"
HYBRID_CASES = [pytest.param(n, marks=pytest.mark.skip(reason=STALLED)) if len(ft.HYBRIDS[n][2]) > 1 else n
                for n in sorted(ft.HYBRIDS)]


@pytest.mark.parametrize("name", HYBRID_CASES)
def test_hybrid_gbuffer(bound, name):
    """Hybrid mix-in cycles (eval_mixed) against the oracle's restatement."""
    print("\n[case] hybrid %s" % name, flush=True)
    u, cm = ft.hybrid_test_scene(name)
    r = bound(cm)
    ref = oracle.render_gbuffer(u, normal_mode=0, shadows=True, code_map=cm)
    g = gpu_render(r, u)
    st = compare(g, ref)
    st.update(pointwise(g, u, 0, code_map=cm))
    report("hybrid " + name, st)
    assert st["hits_ref"] > 0 and st["hit_agree_inner"] >= 0.99, st
    assert st["depth_rel_le_1e-4"] >= 0.95, st
    assert st["pw_trap_equal"] >= 0.9, st
    assert st["pw_normal_deg_le_0.5"] >= 0.85, st
    b = gpu_render(r, u, abi.KERNEL_PLAIN)
    for k in ("hit", "depth", "normal", "trap", "shadow"):
        assert np.array_equal(g[k], b[k]), k


def test_hybrid_of_two_mixins_is_rejected(bound):
    """Two mix-ins in one cycle: the scene is rejected at packing, before any launch."""
    u, cm = ft.hybrid_test_scene("juliabulb_sierpinski_bulb3")
    r = bound(cm)
    with pytest.raises(nd.NdError) as e:
        gpu_render(r, u)
    assert e.value.status == abi.ND_E_UNSUPPORTED_FORMULA and "mix-ins" in str(e.value)
