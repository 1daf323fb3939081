"""Height-map images for the @image formula (formulas/heightmap_primitive.wgsl), on the CPU.

- nd_image_pack (the library's host code) against a numpy restatement of the reference's
  post-decode steps (height_maps.zig: decode, quantize, slopeCode, dilate): premultiplied
  16-bit texels, box downsampling to 2048 texels per edge, slopes dilated over 8 + 2 and
  64 + 2 texels (dilateLine's block runs, themselves held to scipy's window maximum away
  from the line ends), the 8-bit slope codes and the header's slope bound. Bit for bit.
- The oracle's image_sample (template.wgsl:1180-1193): texel centres give the texel, the
  bilinear weights between them, clamping at the edges.
- The oracle's heightmap DE (heightmap_primitive.wgsl:10-64): with no image, a slab of
  thickness base below y = 0; with a flat image, the exact distance above the top face and
  below the base face; on a bumpy image, never longer than the brute-force distance to the
  height field (the slope terms make it a bound), and the Invert flag's closed forms.
"""
import json
import os

import numpy as np
import pytest
from scipy.ndimage import maximum_filter1d

import numericdream_amd as nd
import oracle
from tests import formula_table as ft

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
F32 = np.float32


def ref_pack(img):
    """numpy restatement of height_maps.zig decode() after the pixel read, in float32."""
    h, w, _ = img.shape
    if img.dtype == np.uint16:
        inv = F32(1.0) / F32(65535.0)
        px = img.astype(F32) * inv
        a = px[..., 3:4]
        px = np.concatenate([px[..., :3] * a, a], axis=2).astype(F32)
    else:
        rgb = img[..., :3]
        fin = np.isfinite(rgb)
        lo, hi = rgb[fin].min(), rgb[fin].max()
        span = F32(hi - lo) if hi > lo else F32(1.0)
        c = np.where(fin, np.clip((rgb - F32(lo)) / span, 0, 1), 0).astype(F32)
        px = np.concatenate([c, np.ones((h, w, 1), F32)], axis=2)
    f = 1
    while w // f > 2048 or h // f > 2048:
        f += 1
    dw, dh = max(w // f, 1), max(h // f, 1)
    acc = np.zeros((dh, dw, 4), F32)
    for sy in range(f):
        rows = np.minimum(np.arange(dh) * f + sy, h - 1)
        for sx in range(f):
            cols = np.minimum(np.arange(dw) * f + sx, w - 1)
            acc = (acc + px[rows][:, cols]).astype(F32)
    v = (acc * (F32(1.0) / F32(f * f))).astype(F32)
    v = (np.floor(np.clip(v, 0, 1) * F32(65535.0) + F32(0.5)) / F32(65535.0)).astype(F32)  # half away from 0
    x1 = np.minimum(np.arange(dw) + 1, dw - 1)
    y1 = np.minimum(np.arange(dh) + 1, dh - 1)
    a, b, c, d = v, v[:, x1], v[y1], v[y1][:, x1]
    dx = np.maximum(np.abs(b - a), np.abs(d - c))
    dy = np.maximum(np.abs(c - a), np.abs(d - b))
    g = np.sqrt(dx * dx + dy * dy).astype(F32)
    glob = g.reshape(-1, 4).max(axis=0)

    def dil(x, r):
        return dilate_ref(dilate_ref(x, r, axis=1), r, axis=0)

    near, far = dil(g, 10), dil(g, 66)

    def code(s):
        with np.errstate(divide="ignore"):
            st = np.ceil(F32(16.0) * (np.log2(s * F32(1.001)) - F32(1.0)) + F32(255.0))
        return np.where(s > 0, np.clip(st, 1, 255), 0).astype(np.uint32)

    q = np.round(v * F32(65535.0)).astype(np.uint32)
    words = np.zeros((dh, dw, 4), np.uint32)
    words[..., 0] = q[..., 0] | (q[..., 1] << 16)
    words[..., 1] = q[..., 2] | (q[..., 3] << 16)
    cn, cf = code(near), code(far)
    words[..., 2] = cn[..., 0] | (cn[..., 1] << 8) | (cn[..., 2] << 16) | (cn[..., 3] << 24)
    words[..., 3] = cf[..., 0] | (cf[..., 1] << 8) | (cf[..., 2] << 16) | (cf[..., 3] << 24)
    return words.reshape(-1), dw, dh, (glob * F32(1.001)).astype(F32)


def dilate_ref(x, r, axis):
    """dilateLine (height_maps.zig): per line, the forward running maximum from each block of
    2r + 1 start and the backward one to each block end; out[j] = max(bwd[j - r], fwd[j + r])
    with both indexes clipped to the line. In the interior that is the window maximum over
    [j - r, j + r]; a window clipped by an end takes in the rest of its first block too."""
    x = np.moveaxis(x, axis, 0)
    n, k = x.shape[0], 2 * r + 1
    fwd, bwd = x.copy(), x.copy()
    for i in range(1, n):
        if i % k:
            fwd[i] = np.maximum(fwd[i - 1], x[i])
    for i in range(n - 2, -1, -1):
        if (i + 1) % k:
            bwd[i] = np.maximum(bwd[i + 1], x[i])
    j = np.arange(n)
    out = np.maximum(bwd[np.maximum(j - r, 0)], fwd[np.minimum(j + r, n - 1)])
    return np.moveaxis(out, 0, axis)


def test_dilation_is_a_window_maximum():
    """dilate_ref against scipy's window maximum: equal wherever the window lies inside the
    line, never smaller near its ends (so the slope stays a bound)."""
    rng = np.random.default_rng(1)
    x = rng.random((300, 4)).astype(F32)
    for r in (1, 10, 66):
        got = dilate_ref(x, r, axis=0)
        exact = maximum_filter1d(x, 2 * r + 1, axis=0, mode="nearest")
        inner = slice(r, 300 - r)
        assert np.array_equal(got[inner], exact[inner])
        assert np.all(got >= exact)


def rgba16(h, w, seed):
    rng = np.random.default_rng(seed)
    img = rng.integers(0, 65536, size=(h, w, 4)).astype(np.uint16)
    img[0, 0, 3] = 0          # a transparent texel reads 0 in every channel
    img[-1, -1, 3] = 65535
    return img


@pytest.mark.parametrize("shape", [(5, 3), (40, 70), (2, 4100), (1, 1)])
def test_pack_matches_reference_steps(shape):
    img = rgba16(*shape, seed=shape[0] * 7 + shape[1])
    words, w, h, slope = nd.image_pack(img)
    rw, rwd, rh, rs = ref_pack(img)
    assert (w, h) == (rwd, rh)
    if shape == (2, 4100):
        assert (w, h) == (1366, 1)  # factor 3: 4100 / 2 = 2050 still exceeds 2048
    assert np.array_equal(words, rw)
    assert np.array_equal(slope, rs)


def test_pack_hdr_normalises_finite_range():
    rng = np.random.default_rng(3)
    img = rng.normal(5.0, 3.0, size=(17, 23, 4)).astype(np.float32)
    img[2, 3, 1] = np.nan
    img[4, 4, 0] = np.inf
    words, w, h, slope = nd.image_pack(img)
    rw, _, _, rs = ref_pack(img)
    assert np.array_equal(words, rw) and np.array_equal(slope, rs)
    alpha = words.reshape(-1, 4)[:, 1] >> 16
    assert np.all(alpha == 65535)  # HDR alpha is 1


def test_pack_rejects_bad_input():
    with pytest.raises(ValueError):
        nd.image_pack(np.zeros((4, 4, 3), np.uint16))
    with pytest.raises(ValueError):
        nd.image_pack(np.zeros((4, 4, 4), np.uint8))


def bind_both(slot, img):
    words, w, h, slope = nd.image_pack(img)
    oracle.bind_image(slot, words, w, h, slope.astype(np.float64))
    return w, h


def test_image_sample_bilinear_and_clamped():
    img = np.zeros((2, 3, 4), np.uint16)
    img[..., 3] = 65535
    img[..., 0] = np.array([[0, 30000, 60000], [6000, 9000, 12000]], np.uint16)
    w, h = bind_both(0, img)
    tex = img[..., 0].astype(np.float64) / 65535.0
    try:
        # texel centres
        for y in range(h):
            for x in range(w):
                v, _, _ = oracle.image_sample(0, [(x + 0.5) / w, (y + 0.5) / h])
                assert abs(v[0, 0] - tex[y, x]) < 1e-12 and abs(v[0, 3] - 1.0) < 1e-12
        # between four texel centres: the bilinear weights
        fx, fy = 0.3, 0.8
        v, _, _ = oracle.image_sample(0, [(0.5 + fx) / w, (0.5 + fy) / h])  # t = (fx, fy) from texel (0, 0)
        want = (tex[0, 0] * (1 - fx) + tex[0, 1] * fx) * (1 - fy) + (tex[1, 0] * (1 - fx) + tex[1, 1] * fx) * fy
        assert abs(v[0, 0] - want) < 1e-12
        # outside the image: the edge texel
        v, _, _ = oracle.image_sample(0, [(-3.0, -2.0), (5.0, 9.0)])
        assert abs(v[0, 0] - tex[0, 0]) < 1e-12 and abs(v[1, 0] - tex[1, 2]) < 1e-12
    finally:
        oracle.bind_image(0, None)


def heightmap_scene(params=None):
    u, cm = ft.scene("heightmap_primitive", params=params, cam=(0.0, 2.0, -2.0))
    return u, cm


def test_no_image_is_a_slab():
    oracle.bind_image(0, None)
    u, cm = heightmap_scene()
    # Size 1, Height 0.2, Base 0.02: with no image the top is y = 0 and the slab is [-0.02, 0]
    ys = np.array([0.3, 0.05, -0.5])
    pts = np.stack([np.zeros(3), ys, np.zeros(3)], 1)
    d = oracle.scene_de(u, pts, code_map=cm)
    assert np.allclose(d, [0.3, 0.05, 0.5 - float(F32(0.02))], atol=1e-12)


def test_flat_image_is_a_box():
    img = np.zeros((16, 16, 4), np.uint16)
    img[..., :3] = 32768
    img[..., 3] = 65535
    bind_both(0, img)
    try:
        u, cm = heightmap_scene()
        # the block stores the @param values as float32
        lum = float(F32(0.2126)) + float(F32(0.7152)) + float(F32(0.0722))
        top = float(F32(0.2)) * lum * 32768 / 65535.0
        rng = np.random.default_rng(5)
        # above the top face and below the base face, inside the footprint: exact distances
        xz = rng.uniform(-0.9, 0.9, size=(50, 2))
        up = rng.uniform(top + 0.01, top + 1.0, size=50)
        down = rng.uniform(-1.0, -0.03, size=50)
        d_up = oracle.scene_de(u, np.stack([xz[:, 0], up, xz[:, 1]], 1), code_map=cm)
        d_dn = oracle.scene_de(u, np.stack([xz[:, 0], down, xz[:, 1]], 1), code_map=cm)
        assert np.allclose(d_up, up - top, rtol=0, atol=1e-9)
        assert np.allclose(d_dn, -float(F32(0.02)) - down, rtol=0, atol=1e-9)
        # anywhere: never longer than the distance to the box [-1, 1] x [-0.02, top] x [-1, 1]
        p = rng.uniform(-2.0, 2.0, size=(400, 3))
        lo, hi = np.array([-1.0, -float(F32(0.02)), -1.0]), np.array([1.0, top, 1.0])
        out = np.maximum(np.maximum(lo - p, p - hi), 0.0)
        dist = np.linalg.norm(out, axis=1)
        inside = np.all((p > lo) & (p < hi), axis=1)
        d = oracle.scene_de(u, p, code_map=cm)
        assert np.all(d[~inside] <= dist[~inside] + 1e-9)
        assert np.all(d[inside] <= 0.0)
    finally:
        oracle.bind_image(0, None)


def bump_image(n=24):
    y, x = np.mgrid[0:n, 0:n] / (n - 1.0) * 2.0 - 1.0
    hgt = np.clip(1.0 - (x * x + y * y), 0.0, 1.0) * 0.8 + 0.1 * (x + 1.0)
    img = np.zeros((n, n, 4), np.uint16)
    img[..., 0] = img[..., 1] = img[..., 2] = np.round(np.clip(hgt, 0, 1) * 65535).astype(np.uint16)
    img[..., 3] = 65535
    return img


def test_bumpy_image_is_a_bound():
    img = bump_image()
    w, h = bind_both(0, img)
    try:
        u, cm = heightmap_scene()
        # the height field of the oracle's own sampling, densely: y = 0.2 * luminance(x, z)
        g = np.linspace(-1.0, 1.0, 301)
        X, Z = np.meshgrid(g, g)
        uv = np.stack([X.ravel() * 0.5 + 0.5, Z.ravel() * 0.5 + 0.5], 1)
        v, _, _ = oracle.image_sample(0, uv)
        Y = float(F32(0.2)) * (v[:, :3] @ np.array([0.2126, 0.7152, 0.0722], F32).astype(np.float64))
        surf = np.stack([X.ravel(), Y, Z.ravel()], 1)
        rng = np.random.default_rng(11)
        p = np.stack([rng.uniform(-0.8, 0.8, 300), rng.uniform(0.25, 0.8, 300), rng.uniform(-0.8, 0.8, 300)], 1)
        d = oracle.scene_de(u, p, code_map=cm)
        brute = np.array([np.min(np.linalg.norm(surf - q, axis=1)) for q in p])
        spacing = 2.0 / 300 * np.sqrt(2)
        assert np.all(d <= brute + spacing)  # a bound (the grid's own spacing aside)
        assert np.median(d / brute) > 0.3   # and not a loose one
    finally:
        oracle.bind_image(0, None)


def test_invert_closed_forms():
    for rgb, want in ((65535, 0.0), (0, float(F32(0.2)) * float(F32(0.2126) + F32(0.7152) + F32(0.0722)))):
        img = np.zeros((8, 8, 4), np.uint16)
        img[..., :3] = rgb
        img[..., 3] = 65535
        bind_both(0, img)
        try:
            p = ft.defaults("heightmap_primitive")
            p[7] = 1.0  # Invert: k = (-w.rgb, w.r + w.g + w.b + w.a)
            u, cm = heightmap_scene(p)
            d = oracle.scene_de(u, [[0.1, 0.6, -0.2]], code_map=cm)
            assert abs(d[0] - (0.6 - want)) < 1e-9, (rgb, d)
        finally:
            oracle.bind_image(0, None)


def test_dream_reports_image_paths():
    src = open(os.path.join(ROOT, "tests", "golden", "scenes", "wavefront_opaque.dream")).read()
    scene = json.loads(src)
    inst = scene["instances"][0]
    inst["formula"]["path"] = "formulas/heightmap_primitive.wgsl"
    inst["formula"]["image"] = "maps/terrain.png"
    inst["formula"]["params"] = []
    u, info = nd.load_dream(json.dumps(scene))
    assert info.image_path[0].value == b"maps/terrain.png"
    assert all(info.image_path[k].value == b"" for k in range(1, 16))
    assert list(info.code_to_builtin[:info.n_codes]).count(ft.ID["heightmap_primitive"]) == 1
