"""GPU parity of the mesher's distance passes (nd_mesh_points = cs_mesh_points,
nd_mesh_probe = cs_mesh_probe; mesh.wgsl:49-92) through the C ABI.

The grid's index map (block * 8 + local + local_offset, x fastest) is pinned on the exact
sphere SDF, whose values are known in closed form; the scene DE on the four-fractal scene is
held to the FP64 oracle as closely as the FP32 oracle restatement is; the probe's gradient is
the textbook tetrahedral stencil of the sphere's distance, and h < 0 drops the carves."""
import numpy as np
import pytest

import numericdream_amd as nd
import oracle
from tests import formula_table as ft

pytestmark = pytest.mark.gpu


def grid_points(origin, cell, span, offset, blocks):
    """Point coordinates of every block's span^3 points (x fastest), as the pass lays them out."""
    l = np.arange(span ** 3)
    local = np.stack([l % span, (l // span) % span, l // (span * span)], 1)
    idx = blocks[:, None, :3].astype(np.int64) * 8 + local[None] + offset
    return (np.asarray(origin, np.float64) + idx * cell).reshape(-1, 3)


@pytest.fixture
def sphere_bound(renderer):
    renderer.bind_formulas([ft.ID["sphere_primitive"]])
    yield renderer
    renderer.bind_formulas([0, 1, 2, 3])


def test_points_index_map_on_the_exact_sphere(sphere_bound):
    u, _ = ft.scene("sphere_primitive", params=[1.0], size=(16, 16))
    origin, cell, iso = (-1.5, -1.5, -1.5), 3.0 / 32, 0.05
    blocks = np.array([[x, y, z, 0] for z in range(4) for y in range(4) for x in range(4)], np.uint32)
    print("\n[case] mesh points sphere 32^3 span 8 and 4^3 coarse", flush=True)
    vals, ms = sphere_bound.mesh_points(u, origin, cell, iso, 8, 0, blocks)
    p = grid_points(origin, cell, 8, 0, blocks)
    assert np.allclose(vals, np.linalg.norm(p, axis=1) - 1.0 - iso, atol=2e-6)
    # the coarse pass: one point per block at its centre (local_offset = block / 2)
    vals, _ = sphere_bound.mesh_points(u, origin, cell, iso, 1, 4, blocks)
    p = grid_points(origin, cell, 1, 4, blocks)
    assert np.allclose(vals, np.linalg.norm(p, axis=1) - 1.0 - iso, atol=2e-6)


def test_points_match_the_oracle_on_four_fractals(renderer):
    u = nd.config_uniforms(3, 0)
    rng = np.random.default_rng(4)
    origin, cell = (-1.6, -1.6, -1.6), 3.2 / 64
    blocks = np.concatenate([rng.integers(0, 8, (40, 3)), np.zeros((40, 1), np.int64)], 1).astype(np.uint32)
    print("\n[case] mesh points four fractals, 40 blocks of 8^3", flush=True)
    vals, _ = renderer.mesh_points(u, origin, cell, 0.0, 8, 0, blocks)
    p = grid_points(origin, cell, 8, 0, blocks).astype(np.float32).astype(np.float64)
    r64 = np.ravel(oracle.scene_de(u, p, prec=0))
    r32 = np.ravel(oracle.scene_de(u, p, prec=1))
    # most of this cube lies within 1e-4 of the fractals (median |DE| ~ 1e-4), so values are
    # compared in absolute terms: the FP32 oracle restatement is within 1e-6 of FP64 on
    # 99.99% of these points
    err, err32 = np.abs(vals - r64), np.abs(r32 - r64)
    print("[mesh] abs error <= 1e-6: gpu %.5f fp32 oracle %.5f; <= 1e-5: gpu %.5f; max gpu %.3g"
          % ((err <= 1e-6).mean(), (err32 <= 1e-6).mean(), (err <= 1e-5).mean(), err.max()), flush=True)
    assert (err <= 1e-5).mean() >= 0.999
    assert (err <= 1e-6).mean() >= (err32 <= 1e-6).mean() - 0.01
    assert ((vals >= 0) == (r64 >= 0))[np.abs(r64) > 1e-5].all()


def tetra(f, q, h):
    K = np.array([[1, -1, -1], [-1, -1, 1], [-1, 1, -1], [1, 1, 1]], np.float64)
    return sum(k * f(q + k * h) for k in K) / (4 * h)


def test_probe_on_the_exact_sphere_and_carves(sphere_bound):
    u, _ = ft.scene("sphere_primitive", params=[1.0], size=(16, 16))
    rng = np.random.default_rng(7)
    q = rng.uniform(-2, 2, (500, 3))
    h = rng.uniform(1e-3, 0.05, 500)
    pts = np.concatenate([q, h[:, None]], 1).astype(np.float32)
    print("\n[case] mesh probe sphere 500 points", flush=True)
    out = sphere_bound.mesh_probe(u, pts)
    qf, hf = pts[:, :3].astype(np.float64), pts[:, 3].astype(np.float64)
    sdf = lambda x: np.linalg.norm(x) - 1.0
    assert np.allclose(out[:, 0], np.linalg.norm(qf, axis=1) - 1.0, atol=2e-6)
    ref = np.array([tetra(sdf, qf[i], hf[i]) for i in range(len(qf))])
    assert np.allclose(out[:, 1:], ref, atol=5e-4)
    # one carve (a sphere of radius 0.5 at (1, 0, 0)): h > 0 sees it, h < 0 does not
    u.carve_count = 1.0
    u.carves[0][0], u.carves[0][1], u.carves[0][2], u.carves[0][3] = 1.0, 0.0, 0.0, 0.5
    near = np.array([[0.9, 0.05, 0.0, 0.01], [0.9, 0.05, 0.0, -0.01]], np.float32)
    o = sphere_bound.mesh_probe(u, near)
    assert np.isclose(o[1, 0], np.linalg.norm(near[1, :3].astype(np.float64)) - 1.0, atol=2e-6)
    carved = np.ravel(oracle.scene_de(u, near[:1, :3].astype(np.float64), code_map=[ft.ID["sphere_primitive"]]))[0]
    assert abs(o[0, 0] - o[1, 0]) > 1e-3 and np.isclose(o[0, 0], carved, atol=1e-5)
