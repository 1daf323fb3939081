"""Pins of the oracle's restatement of the formula plug-ins beyond the four tuned ones
(oracle/nd_oracle_formulas.hpp), each chosen so that a plausible slip (a swapped @param
slot, a dropped term, a wrong sign or mirror) fails:

  primitives        against brute-force distances to densely sampled surfaces;
  identities        a formula equals an already pinned one under the parameter map that the
                    two files imply (a Julia variant = its Mandelbrot sibling with CScale 0
                    and the Julia offset; a twist of angle a = a z rotation of a; a fold
                    that is the identity on part of space; a power-1 map that is the
                    identity), compared at random points and parameters;
  symmetries        the mirrors and wedges a formula starts with make its DE invariant
                    under them;
  closed forms      radius recurrences of power maps, the full quaternion iteration, and
                    far-field values (radioactive's floor plane, recursivesphere's outer
                    sphere, kleinianlimit's ceiling cut, fbmsdf's box);
  escape            a bulb step outside the bail-out radius leaves the IFS or box step it
                    is fused with.

The packed-carry formulas (radioactive, recursivesphere, tree, kleinianlimit, fbmsdf) are
pinned only by their far-field and trunk closed forms: their interiors are pinned by
nothing but the GPU parity against this restatement. Seascape is pinned by its calm-sea
closed form (the wave sum is not), dragon by brute force over its 2^n curve segments, and
romanesco is parity unpinned.
"""
import numpy as np
import pytest

import oracle
from tests import formula_table as ft

RNG = np.random.default_rng(20261203)


def f32(v):
    """Parameters as the FP32 values the Uniforms block carries."""
    return [float(np.float32(x)) for x in v]


def de(name, params, pts):
    """FP64 DE of one instance of the formula (identity transform) at pts."""
    u, cm = ft.scene(name, params=params)
    return oracle.scene_de(u, np.asarray(pts, np.float64), code_map=cm)


def step(name, z, dr, pos, params):
    return oracle.formula_step(ft.ID[name], z, dr, pos, params)


def rand_params(name, keep=(), fixed=None):
    """Random values inside each @param's [min, max]; integer @params and the slots in keep
    stay at their defaults; fixed overrides."""
    rows = ft.param_table()[name]
    p = ft.defaults(name)
    for i, r in enumerate(rows):
        if r["int"] or i in keep:
            continue
        p[i] = float(RNG.uniform(r["min"], r["max"]))
    for k, v in (fixed or {}).items():
        p[k] = v
    return f32(p)


def pts_box(n, half=2.0):
    return RNG.uniform(-half, half, (n, 3))


def close(a, b, rtol=1e-9, atol=1e-12):
    a, b = np.asarray(a), np.asarray(b)
    return np.allclose(a, b, rtol=rtol, atol=atol)


# ============================================================================= primitives
def sample_surface(kind, n=400):
    """Points on the unit-size primitive surfaces (size 1), dense enough for 1e-2 checks."""
    g = np.linspace(-1, 1, n)
    if kind == "cube":
        a, b = np.meshgrid(g, g)
        a, b = a.ravel(), b.ravel()
        one = np.ones_like(a)
        faces = [np.stack(v, 1) for v in ((one, a, b), (-one, a, b), (a, one, b), (a, -one, b), (a, b, one), (a, b, -one))]
        return np.concatenate(faces)
    if kind == "cylinder":
        th = np.linspace(0, 2 * np.pi, 4 * n, endpoint=False)
        T, Y = np.meshgrid(th, g)
        side = np.stack([np.sin(T).ravel(), Y.ravel(), np.cos(T).ravel()], 1)
        R, T2 = np.meshgrid(np.linspace(0, 1, n // 2), th)
        cap = np.stack([(R * np.sin(T2)).ravel(), np.ones(R.size), (R * np.cos(T2)).ravel()], 1)
        cap2 = cap * np.array([1, -1, 1])
        return np.concatenate([side, cap, cap2])
    if kind == "pyramid":
        # apex (0, 1, 0), square base |x|, |z| <= 1 at y = -1
        s = np.linspace(0, 1, n)
        S, A = np.meshgrid(s, g)
        S, A = S.ravel(), A.ravel()
        y = 1 - 2 * S
        faces = [np.stack([S, y, S * A], 1), np.stack([-S, y, S * A], 1), np.stack([S * A, y, S], 1),
                 np.stack([S * A, y, -S], 1)]
        a, b = np.meshgrid(g, g)
        base = np.stack([a.ravel(), -np.ones(a.size), b.ravel()], 1)
        return np.concatenate(faces + [base])
    raise ValueError(kind)


def brute(pts, surf):
    d = np.empty(len(pts))
    for i, p in enumerate(pts):
        d[i] = np.sqrt(((surf - p) ** 2).sum(1)).min()
    return d


def outside_points(n, lo=1.3, hi=2.5):
    v = RNG.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    return v * RNG.uniform(lo, hi, (n, 1))


def test_sphere_is_the_exact_sdf():
    for R in f32([0.3, 1.0, 2.5]):
        p = pts_box(200, 4.0)
        assert close(de("sphere_primitive", [R], p), np.linalg.norm(p, axis=1) - R, rtol=1e-12, atol=1e-12)


def test_cube_matches_brute_force():
    size = f32([1.7])[0]
    p = outside_points(60, 1.8, 4.0) * size
    d = de("cube_primitive", [size], p)
    bf = brute(p / size, sample_surface("cube")) * size
    assert np.abs(d - bf).max() < 2e-2 * size
    inside = RNG.uniform(-0.9, 0.9, (50, 3)) * size
    assert close(de("cube_primitive", [size], inside), -(size - np.abs(inside).max(1)), atol=1e-12)


def test_cylinder_matches_brute_force():
    size = 0.8
    p = outside_points(60, 1.6, 3.0) * size
    d = de("cylinder_primitive", [size], p)
    bf = brute(p / size, sample_surface("cylinder", 200)) * size
    assert np.abs(d - bf).max() < 2e-2 * size


def test_torus_is_the_exact_sdf():
    R, r = f32([1.3, 0.4])
    p = pts_box(300, 3.0)
    q = np.sqrt(p[:, 0] ** 2 + p[:, 2] ** 2) - R
    assert close(de("torus_primitive", [R, r], p), np.sqrt(q ** 2 + p[:, 1] ** 2) - r, rtol=1e-9, atol=1e-9)


def test_pyramid_is_a_lower_bound_and_exact_over_faces():
    size = 1.0
    surf = sample_surface("pyramid", 300)
    p = outside_points(60, 1.6, 3.0)
    d = de("pyramid_primitive", [size], p)
    bf = brute(p, surf)
    assert np.all(d <= bf + 1e-2)
    # just outside the middle of a side face (normal (2, 1, 0) / sqrt 5) and below the base
    n = np.array([2.0, 1.0, 0.0]) / np.sqrt(5)
    face_mid = np.array([0.5, 0.0, 0.0])
    q = np.array([face_mid + 0.2 * n, [0.1, -1.3, 0.2]])
    assert close(de("pyramid_primitive", [size], q), [0.2, 0.3], rtol=1e-9)


@pytest.mark.parametrize("sides", [3, 5, 6, 8])
def test_prism_matches_the_polygon_prism(sides):
    size, height = 1.2, 0.7
    # the n-gon of circumradius size in xz with vertices at angle 2 pi k / n from +z toward +x
    ang = 2 * np.pi * np.arange(sides) / sides
    vx, vz = size * np.sin(ang), size * np.cos(ang)
    pts = []
    for k in range(sides):
        a = np.array([vx[k], vz[k]]); b = np.array([vx[(k + 1) % sides], vz[(k + 1) % sides]])
        for t in np.linspace(0, 1, 120):
            e = a + (b - a) * t
            for y in np.linspace(-height, height, 60):
                pts.append([e[0], y, e[1]])
    # caps: fill the polygon by radial interpolation to the boundary
    for k in range(sides):
        a = np.array([vx[k], vz[k]]); b = np.array([vx[(k + 1) % sides], vz[(k + 1) % sides]])
        for t in np.linspace(0, 1, 40):
            e = a + (b - a) * t
            for s in np.linspace(0, 1, 40):
                pts.append([s * e[0], height, s * e[1]])
                pts.append([s * e[0], -height, s * e[1]])
    surf = np.array(pts)
    p = outside_points(40, 1.6, 3.0)
    d = de("prism_primitive", [size, height, float(sides)], p)
    bf = brute(p, surf)
    assert np.abs(d - bf).max() < 2.5e-2


# ============================================================================= identities
def check_identity(a, pa, b, pb, n=300, half=2.0, rtol=1e-9, atol=1e-12):
    p = pts_box(n, half)
    da, db = de(a, pa, p), de(b, pb, p)
    assert close(da, db, rtol=rtol, atol=atol), (a, b, np.abs(da - db).max())


def test_juliakifs_is_sierpinski_with_the_julia_offset():
    for _ in range(3):
        j = rand_params("juliakifs_ifs_julia")
        s = ft.defaults("sierpinski3_ifs")
        s[0] = j[0]; s[1] = j[4]; s[2:5] = j[1:4]; s[5:8] = j[5:8]
        check_identity("juliakifs_ifs_julia", j, "sierpinski3_ifs", s)


def test_juliabox_is_amazingbox_with_cscale_zero():
    for _ in range(3):
        j = rand_params("juliabox_box_julia")
        a = [j[0], j[1], j[2], j[3], j[7], j[8], j[9], j[10], 0.0, j[4], j[5], j[6]]
        check_identity("juliabox_box_julia", j, "amazingbox_box", a)


def test_mandelboxtwist_is_amazingbox_rotated_about_z():
    for _ in range(3):
        t = rand_params("mandelboxtwist_box_twist", fixed={6: 0.0, 7: 0.0})
        a = [t[0], t[1], t[2], t[3], t[5], 0.0, 0.0, t[4], t[8], t[9], t[10], t[11]]
        check_identity("mandelboxtwist_box_twist", t, "amazingbox_box", a)


def test_buffalo_family():
    for _ in range(3):
        t = rand_params("buffalotwist_box_twist", fixed={5: 0.0, 6: 0.0})
        b = [t[0], t[1], t[2], t[4], 0.0, 0.0, t[3], t[7], t[8], t[9], t[10]]
        check_identity("buffalotwist_box_twist", t, "buffalo_box", b)
        j = rand_params("buffalojulia_box_julia")
        b = [j[0], j[1], j[2], j[6], j[7], j[8], j[9], 0.0, j[3], j[4], j[5]]
        check_identity("buffalojulia_box_julia", j, "buffalo_box", b)


def test_apollonian_family():
    for _ in range(3):
        a = rand_params("apollonian_kleinian")
        # no box fold: amazingbox with a folding limit no orbit reaches
        m = [a[0], a[1], a[2], 1e9, a[3], a[4], a[5], a[6], a[7], a[8], a[9], a[10]]
        check_identity("apollonian_kleinian", a, "amazingbox_box", m)
        j = rand_params("apollonianjulia_kleinian_julia")
        b = [j[0], j[1], j[2], j[6], j[7], j[8], j[9], 0.0, j[3], j[4], j[5]]
        check_identity("apollonianjulia_kleinian_julia", j, "apollonian_kleinian", b)


def test_kleinian_relatives_are_kleinian_without_the_box_fold():
    for _ in range(3):
        j = rand_params("kleinianjulia_kleinian_julia")
        k = [1e9, j[0], j[1], j[2], j[3], j[8], j[7], 0.0, j[4], j[5], j[6]]
        check_identity("kleinianjulia_kleinian_julia", j, "kleinian_kleinian", k)
        p = rand_params("pseudokleinian_kleinian")
        k = [1e9, p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9]]
        check_identity("pseudokleinian_kleinian", p, "kleinian_kleinian", k)


def test_one_iteration_of_surf_and_foam_is_amazingbox():
    # the |z.x|, |z.z| after the step do not change the length the first finalize reads
    for name in ("amazingsurf_box_fold", "mandelboxfoam_box_fold"):
        s = rand_params(name, fixed={4: 1.0})
        check_identity(name, s, "amazingbox_box", s[:12])


def test_mandelboxinv_on_its_inversion_sphere_is_amazingbox():
    for _ in range(3):
        v = rand_params("mandelboxinv_box_fold", fixed={5: 1.0})
        dirs = RNG.normal(size=(200, 3))
        p = dirs / np.linalg.norm(dirs, axis=1)[:, None] * v[0]
        a = [v[1], v[2], v[3], v[4], 1.0, v[6], v[7], v[8], v[9], v[10], v[11], v[12]]
        assert close(de("mandelboxinv_box_fold", v, p), de("amazingbox_box", a, p), rtol=1e-9)


def test_bulb_variants_reduce_to_the_mandelbulb():
    for power in (3.0, 8.0):
        m = [power, 6.0]
        t = [power, 6.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0]
        check_identity("mandelbulbtwist_bulb_twist", t, "mandelbrot", m, half=1.3, rtol=1e-8)
        # a second power of 1 is the identity map
        d = [power, 6.0, 1.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0]
        check_identity("mandelbulbdual_hybrid_bulb", d, "mandelbrot", m, half=1.3, rtol=1e-8)


def test_menger_relatives():
    for _ in range(3):
        m = rand_params("menger3_ifs", fixed={5: 0.0, 6: 0.0, 7: 0.0})
        k = [m[4], 0.0, m[0], m[1], m[2], m[3], 0.0, 0.0]
        check_identity("kaleidoscopicifs_ifs_twist", k, "menger3_ifs", m)


def test_icosahedral_relatives():
    for _ in range(3):
        i = rand_params("icosahedralifs_ifs", fixed={5: 0.0, 6: 0.0, 7: 0.0})
        j = [i[0], i[1], i[2], 0.0, 0.0, 0.0, i[3], i[4], 0.0, 0.0, 0.0]
        check_identity("icosahedraljulia_ifs_julia", j, "icosahedralifs_ifs", i)
        t = [i[0], i[1], i[2], 0.0, i[3], i[4], 0.0, 0.0]
        check_identity("icosahedraltwist_ifs_twist", t, "icosahedralifs_ifs", i)


def test_fused_steps_outside_the_bailout():
    """|z| > 2 skips the bulb part: the rest is the IFS / box step it is fused with."""
    for _ in range(20):
        pos = RNG.normal(size=3)
        pos *= RNG.uniform(2.2, 4.0) / np.linalg.norm(pos)
        b = rand_params("bulbtetra_hybrid_bulb_ifs")
        s = [b[1], b[2], b[3], b[4], b[5], 0.0, 0.0, 0.0]
        z1, d1 = step("bulbtetra_hybrid_bulb_ifs", pos, 1.0, pos, b)
        z2, d2 = step("sierpinski3_ifs", pos, 1.0, pos, s)
        assert close(z1, z2) and close(d1, d2)
        bb = rand_params("bulbbox_hybrid_bulb_box")
        a = [bb[1], bb[2], bb[3], bb[4], bb[5], bb[8], bb[9], bb[10], bb[11], bb[12], bb[13], bb[14]]
        z1, d1 = step("bulbbox_hybrid_bulb_box", pos, 1.0, pos, bb)
        z2, d2 = step("amazingbox_box", pos, 1.0, pos, a)
        assert close(z1, z2) and close(d1, d2)


def test_hybridkifs_in_the_positive_octant_is_a_menger_step():
    """tetra_fold is the identity where x + y, x + z, y + z >= 0."""
    for _ in range(20):
        pos = RNG.uniform(0.05, 2.0, 3)
        h = rand_params("hybridkifs_hybrid_ifs", fixed={5: 0.0, 6: 0.0, 7: 0.0})
        m = [h[1], h[2], h[3], h[4], h[0], 0.0, 0.0, 0.0]
        z1, d1 = step("hybridkifs_hybrid_ifs", pos, 1.0, pos, h)
        z2, d2 = step("menger3_ifs", pos, 1.0, pos, m)
        assert close(z1, z2) and close(d1, d2)


# ============================================================================= symmetries
def check_invariant(name, params, maps, n=300, half=2.0, rtol=1e-9):
    p = pts_box(n, half)
    d0 = de(name, params, p)
    for f in maps:
        d1 = de(name, params, f(p))
        assert close(d0, d1, rtol=rtol, atol=1e-12), (name, np.abs(d0 - d1).max())


SIGNS = [lambda p, s=s: p * np.array(s) for s in ((-1, 1, 1), (1, -1, 1), (1, 1, -1))]
PERMS = [lambda p: p[:, [1, 0, 2]], lambda p: p[:, [2, 1, 0]], lambda p: p[:, [0, 2, 1]]]
SWAP_XY = [lambda p: p[:, [1, 0, 2]]]
# tetra_fold's first conditional mirror (across x + y = 0) maps p and its image to the same
# point; the later two only act on what the first leaves, so they give no invariance alone
TETRA = [lambda p: np.stack([-p[:, 1], -p[:, 0], p[:, 2]], 1)]
# The packed-carry formulas keep their running minimum as FP32 with its low mantissa bits
# replaced by a counter: 5 bits are 31 ulp (3.7e-6 relative), 7 bits 127 ulp (1.5e-5).
PACK5, PACK7 = 4e-6, 1.6e-5


def wedge_maps(degree):
    a = 2 * np.pi / degree
    rot = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    return [lambda p: p @ rot.T, lambda p: p * np.array([1, -1, 1])]


def test_cubic_folds_make_cubic_symmetry():
    check_invariant("cubejulia_ifs_julia", rand_params("cubejulia_ifs_julia"), SIGNS + PERMS)
    check_invariant("mandelboxsort_box_fold", rand_params("mandelboxsort_box_fold", fixed={8: 0.0}), SIGNS + PERMS)


def test_tetrahedral_folds_make_tetrahedral_symmetry():
    check_invariant("sierpinski3_ifs", rand_params("sierpinski3_ifs"), TETRA)
    check_invariant("juliakifs_ifs_julia", rand_params("juliakifs_ifs_julia"), TETRA)
    check_invariant("diagonaltetra_ifs_fold", rand_params("diagonaltetra_ifs_fold"), SWAP_XY)
    check_invariant("mandelboxdiag_box_fold", rand_params("mandelboxdiag_box_fold", fixed={8: 0.0}), SWAP_XY)


def test_abs_folds_make_mirror_symmetry():
    for name in ("icosahedralifs_ifs", "icosahedraljulia_ifs_julia", "icosahedraltwist_ifs_twist"):
        check_invariant(name, rand_params(name), SIGNS)
    check_invariant("diagonalicosa_ifs_fold", rand_params("diagonalicosa_ifs_fold"), SIGNS + SWAP_XY)
    check_invariant("buffalojulia_box_julia", rand_params("buffalojulia_box_julia"), SIGNS)
    w = ft.defaults("wobblekifs_ifs_twist_julia")
    for k in (7, 8, 9, 10, 11, 14, 15):
        w[k] = 0.0
    check_invariant("wobblekifs_ifs_twist_julia", w, SIGNS)


@pytest.mark.parametrize("degree", [3.0, 5.0, 6.0])
def test_wedge_folds_make_dihedral_symmetry(degree):
    s = rand_params("starkifs_ifs_fold", fixed={0: degree})
    check_invariant("starkifs_ifs_fold", s, wedge_maps(degree), rtol=1e-8)
    k = rand_params("kaleidojulia_ifs_fold_julia", fixed={0: degree})
    check_invariant("kaleidojulia_ifs_fold_julia", k, wedge_maps(degree), rtol=1e-8)
    p = rand_params("prismkifs_box_fold", fixed={0: degree, 9: 0.0})
    check_invariant("prismkifs_box_fold", p, wedge_maps(degree), rtol=1e-8)


def test_mandalay_is_odd_symmetric():
    m = ft.defaults("mandalay_box_fold")
    check_invariant("mandalay_box_fold", m, [lambda p: -p], half=1.0)


# ============================================================================= closed forms
def bulb_log_escape(r, dr):
    return np.minimum(0.5 * np.log(np.maximum(r, 1e-6)) * r / np.maximum(dr, 1e-6), r)


@pytest.mark.parametrize("name", ["juliabulb_bulb_julia", "mandelbulbmirror_bulb_fold"])
def test_power_map_radius_recurrence(name):
    """With no additive term, |z^p| = |z|^p whatever the angles and phases: the DE follows
    r <- r^p, dr <- p r^(p-1) dr + 1 until r > 2."""
    for power in (2.0, 5.0, 8.0):
        p = ft.defaults(name)
        p[0] = power
        iters = int(p[1])
        if name == "juliabulb_bulb_julia":
            p[2:5] = [0.0, 0.0, 0.0]; p[5], p[6] = 17.0, -33.0
        else:
            p[2], p[3] = 21.0, -8.0; p[4] = 0.0; p[5:8] = [0.0, 0.0, 0.0]
        pts = pts_box(200, 1.5)
        r = np.linalg.norm(pts, axis=1); dr = np.ones(len(r))
        for _ in range(iters):
            live = r <= 2.0
            dr = np.where(live, power * r ** (power - 1) * dr + 1.0, dr)
            r = np.where(live, r ** power, r)
        assert close(de(name, p, pts), bulb_log_escape(r, dr), rtol=1e-8)


def quat_mul(a, b):
    return np.stack([a[..., 0] * b[..., 0] - a[..., 1] * b[..., 1] - a[..., 2] * b[..., 2] - a[..., 3] * b[..., 3],
                     a[..., 0] * b[..., 1] + a[..., 1] * b[..., 0] + a[..., 2] * b[..., 3] - a[..., 3] * b[..., 2],
                     a[..., 0] * b[..., 2] - a[..., 1] * b[..., 3] + a[..., 2] * b[..., 0] + a[..., 3] * b[..., 1],
                     a[..., 0] * b[..., 3] + a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1] + a[..., 3] * b[..., 0]], -1)


def test_quatjulia_is_the_quaternion_julia_set():
    """q <- q^2 + c in full quaternions from q = (x, y, z, slice), |q'| <- 2 |q| |q'|,
    bail-out at |q|^2 > 256; DE = 0.5 ln|q| |q| / |q'|."""
    for _ in range(4):
        p = rand_params("quatjulia_bulb_julia")
        iters = int(p[5])
        pts = pts_box(200, 1.5)
        q = np.concatenate([pts, np.full((len(pts), 1), p[4])], 1)
        c = np.array(p[0:4])
        dq = np.ones(len(pts))
        for _ in range(iters):
            n2 = (q * q).sum(1)
            live = n2 <= 256.0
            dq = np.where(live, 2.0 * np.sqrt(n2) * dq, dq)
            q = np.where(live[:, None], quat_mul(q, q) + c, q)
        r = np.linalg.norm(q, axis=1)
        want = 0.5 * np.log(np.maximum(r, 1e-6)) * r / np.maximum(dq, 1e-6)
        assert close(de("quatjulia_bulb_julia", p, pts), want, rtol=1e-7, atol=1e-10)


def test_radioactive_floor_plane():
    """DE >= y - floor always (the running max starts there), and equals it once y - floor
    exceeds every level's cube half-size (<= Size * Cube)."""
    p = f32(ft.defaults("radioactive_ifs_carve_twist"))
    pts = pts_box(300, 12.0)
    d = de("radioactive_ifs_carve_twist", p, pts)
    plane = pts[:, 1] - p[5]
    assert np.all(d >= plane - PACK5 * np.maximum(np.abs(plane), 1))
    high = plane > p[1] * p[3] + 1e-3
    assert high.any() and close(d[high], plane[high], rtol=PACK5)


def test_recursivesphere_far_field_is_its_outer_sphere():
    """Far from the set and near an axis the outer sphere is the nearest (the child spheres
    sit toward the corners, so along a diagonal they come closer)."""
    p = f32(ft.defaults("recursivesphere_ifs_carve_twist"))
    pts = outside_points(2000, 3.0, 8.0) * p[1]
    pts = pts[np.abs(pts).max(1) >= 0.95 * np.linalg.norm(pts, axis=1)]
    assert len(pts) >= 50
    assert close(de("recursivesphere_ifs_carve_twist", p, pts), np.linalg.norm(pts, axis=1) - p[1], rtol=PACK5)


def test_kleinianlimit_ceiling_cut():
    """Above the ceiling the DE is the cut plane pos.y + 1 - Ceiling * KleinR."""
    p = f32(ft.defaults("kleinianlimit_kleinian"))
    pts = pts_box(300, 3.0)
    pts[:, 1] = RNG.uniform(2.0, 5.0, len(pts))
    cut = pts[:, 1] + 1.0 - p[4] * p[0]
    assert close(de("kleinianlimit_kleinian", p, pts), cut, rtol=PACK7)


def test_fbmsdf_far_field_is_its_box():
    p = f32(ft.defaults("fbmsdf_primitive_carve"))
    pts = outside_points(200, 4.0, 8.0)
    q = np.abs(pts) - p[1]
    box = np.minimum(q.max(1), 0) + np.linalg.norm(np.maximum(q, 0), axis=1)
    assert close(de("fbmsdf_primitive_carve", p, pts), box, rtol=PACK5)


def test_tree_trunk_is_a_round_cone():
    """Near the foot of the trunk the nearest limb is the trunk, a round cone (radii
    Thickness and Thickness Ratio, height Trunk, in the root frame scaled by 0.6 about
    (0, -1, 0)), checked against its sampled surface."""
    p = f32(ft.defaults("tree_ifs"))
    r1, r2, h = p[5], p[5] * p[3], p[6]
    # sample the round cone: the convex hull of two spheres on the y axis
    th = np.linspace(0, 2 * np.pi, 180, endpoint=False)
    ph = np.linspace(0, np.pi, 90)
    T, F = np.meshgrid(th, ph)
    sph = np.stack([np.sin(F) * np.sin(T), np.cos(F), np.sin(F) * np.cos(T)], -1).reshape(-1, 3)
    s = np.linspace(0, 1, 200)
    ring = []
    for k in s:
        # tangent line between the two circles in the (radial, y) plane
        b = (r1 - r2) / h
        a = np.sqrt(1 - b * b)
        y0, y1 = r1 * b, h + r2 * b
        rad0, rad1 = r1 * a, r2 * a
        ring.append(np.stack([(rad0 + (rad1 - rad0) * k) * np.sin(th), np.full_like(th, y0 + (y1 - y0) * k),
                              (rad0 + (rad1 - rad0) * k) * np.cos(th)], 1))
    surf = np.concatenate([sph * r1, sph * r2 + np.array([0, h, 0])] + ring)
    pts = np.stack([RNG.uniform(0.15, 0.35, 40), RNG.uniform(-0.98, -0.9, 40), RNG.uniform(-0.05, 0.05, 40)], 1)
    local = (pts + np.array([0, 1, 0])) / 0.6
    inside = np.linalg.norm(local[:, [0, 2]], axis=1) < r1
    bf = brute(local, surf) * 0.6
    bf[inside] = -bf[inside]
    assert np.abs(de("tree_ifs", p, pts) - bf).max() < 3e-3


def test_seascape_calm_sea_is_its_slab():
    """Height 0: no waves, the DE is the box [-Extent, Extent] x [-Depth, 0] x [-Extent,
    Extent] (in units of Size)."""
    p = f32(ft.defaults("seascape_primitive"))
    p[2] = 0.0
    size, ext, depth = p[1], p[6], p[7]
    pts = pts_box(300, 3.0)
    q = pts / size
    e = np.stack([np.abs(q[:, 0]) - ext, np.maximum(-depth - q[:, 1], q[:, 1]), np.abs(q[:, 2]) - ext], 1)
    box = (np.linalg.norm(np.maximum(e, 0), axis=1) + np.minimum(e.max(1), 0)) * size
    assert close(de("seascape_primitive", p, pts), box, rtol=1e-9, atol=1e-12)


def rot_z(a):
    return np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1.0]])


def rot_x(a):
    return np.array([[1.0, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])


@pytest.mark.parametrize("levy", [0.0, 1.0])
def test_dragon_is_the_distance_to_its_curve(levy):
    """The depth-first descent finds the nearest of the 2^n segments of the curve: the images
    of [0, 1] x (0, 0) under every n-fold composition of f_a(w) = s qa w and
    f_b(w) = s qb w + ob, in the frame where the world point is z / 2 + centre."""
    n, fold, ta, tb, thick = 6, 41.0, 70.0, 110.0, 0.2
    p = f32([n, fold, ta, tb, levy, thick])
    th = np.radians(p[1])
    s = 0.5 / np.cos(th)
    qa = rot_z(th) @ rot_x(np.radians(p[2]))
    qb = rot_z(-th if levy else np.pi - th) @ rot_x(np.radians(p[3]))
    ob = np.array([0.5, 0.5 * np.tan(th), 0.0]) if levy else np.array([1.0, 0.0, 0.0])
    fa = lambda w: s * (qa @ w)
    fb = lambda w: s * (qb @ w) + ob
    segs = []
    for word in range(2 ** n):
        a, b = np.zeros(3), np.array([1.0, 0, 0])
        for k in range(n - 1, -1, -1):
            f = fb if (word >> k) & 1 else fa
            a, b = f(a), f(b)
        segs.append((a, b))
    A = np.array([x[0] for x in segs]); B = np.array([x[1] for x in segs])
    # the centre: whichever of the fixed points of the mean map and of f_b gives the tighter
    # bounding sphere
    I = np.eye(3)
    cen = np.linalg.solve(I - (qa + qb) * (0.5 * s), 0.5 * ob)
    mid = 0.5 * np.linalg.solve(I - qb * s, ob)
    rad = lambda c: max(np.linalg.norm(s * (qa @ c) - c), np.linalg.norm(ob + s * (qb @ c) - c)) / (1 - s)
    c = mid if rad(mid) < rad(cen) else cen
    pts = RNG.uniform(-1.2, 1.2, (60, 3))
    local = pts / 2 + c
    ab = B - A
    want = np.empty(len(pts))
    for i, q in enumerate(local):
        t = np.clip(((q - A) * ab).sum(1) / (ab * ab).sum(1), 0, 1)
        want[i] = np.linalg.norm(A + ab * t[:, None] - q, axis=1).min()
    want = 2 * (want - thick * s ** n)
    assert close(de("dragon_ifs_twist", p, pts), want, rtol=1e-7, atol=1e-9)


# ============================================================================= hybrids
def de_u(u, cm, pts):
    return oracle.scene_de(u, np.asarray(pts, np.float64), code_map=cm)


def test_hybrid_of_base_steps_only_is_the_plain_formula():
    p = f32(ft.defaults("amazingbox_box"))
    n = int(p[4])
    u, cm = ft.hybrid_scene("amazingbox_box", [("juliabox_box_julia", None, 0)], base_n=n, total=n, base_params=p)
    pts = pts_box(200, 3.0)
    assert close(de_u(u, cm, pts), de("amazingbox_box", p, pts), rtol=1e-12)


def test_hybrid_of_mixin_steps_only_is_the_mixin_formula():
    """base_n = 0: every step is the mix-in's; the base only finalizes (both scaled)."""
    j = rand_params("juliabox_box_julia")
    k = 9
    u, cm = ft.hybrid_scene("amazingbox_box", [("juliabox_box_julia", j, 1)], base_n=0, total=k)
    j2 = list(j)
    j2[7] = float(k)
    pts = pts_box(200, 3.0)
    assert close(de_u(u, cm, pts), de("juliabox_box_julia", j2, pts), rtol=1e-12)


def test_hybrid_cycle_schedule():
    """Mandelbulb x2, Mandelbox x1, Menger x0, Kleinian x1 per cycle: the DE equals the
    steps composed in that order by hand, finalized by the Mandelbulb."""
    bulb = f32([8.0, 6.0])
    box = rand_params("amazingbox_box")
    kle = rand_params("kleinian_kleinian")
    u, cm = ft.hybrid_scene("mandelbrot", [("amazingbox_box", box, 1), ("menger3_ifs", None, 0),
                                            ("kleinian_kleinian", kle, 1)], base_n=2, total=11, base_params=bulb)
    sched = [("mandelbrot", bulb), ("mandelbrot", bulb), ("amazingbox_box", box), ("kleinian_kleinian", kle)]
    pts = pts_box(60, 1.5)
    for q in pts:
        z, dr = q.copy(), 1.0
        for it in range(11):
            name, prm = sched[it % 4]
            z, dr = step(name, z, dr, q, prm)
        want = oracle.helper(16, [ft.ID["mandelbrot"], z[0], z[1], z[2], dr], 1)[0]
        got = de_u(u, cm, [q])[0]
        assert np.isclose(got, want, rtol=1e-12, atol=1e-15), (q, got, want)


@pytest.mark.parametrize("name", sorted(ft.HYBRIDS))
def test_hybrid_analytic_normal_is_the_de_gradient(name):
    """The analytic normal of a hybrid (forward-mode tangents through the mix-in cycle) equals
    the central-difference gradient of the hybrid DE at its own hit points."""
    u, cm = ft.hybrid_test_scene(name, size=(24, 18))
    g = oracle.render_gbuffer(u, normal_mode=0, shadows=False, code_map=cm)
    h = g["hit"] > 0
    assert h.sum() > 20
    py, px = np.nonzero(h)
    o = np.array([u.camera_pos[k] for k in range(3)], np.float64)
    d = np.array([oracle.primary_dir(u, int(x), int(y)) for x, y in zip(px, py)])
    pos = o + d * g["depth"][h].astype(np.float64)[:, None]
    n = oracle.normal(u, pos, 0, None, code_map=cm)
    fd = np.zeros_like(pos)
    for k in range(3):
        e = np.zeros(3)
        e[k] = 1e-6
        fd[:, k] = de_u(u, cm, pos + e) - de_u(u, cm, pos - e)
    fd /= np.linalg.norm(fd, axis=1, keepdims=True)
    cosang = (n * fd).sum(1) / np.linalg.norm(n, axis=1)
    assert np.median(cosang) > 1 - 1e-8 and (cosang > 0.999).mean() > 0.95, np.sort(cosang)[:5]
