"""The formula plug-ins beyond the four tuned ones: builtin id (nd_builtin_formula), the
reference formula file, and a single-instance test scene for each (its @param defaults
from tests/golden/formula_params.json and a camera on -z looking at the origin)."""
import json
import os

import numericdream_amd as nd

HERE = os.path.dirname(os.path.abspath(__file__))

FILES = [
    "sphere_primitive", "cube_primitive", "cylinder_primitive", "torus_primitive", "pyramid_primitive",
    "prism_primitive", "sierpinski3_ifs", "cubejulia_ifs_julia", "juliakifs_ifs_julia", "diagonaltetra_ifs_fold",
    "starkifs_ifs_fold", "kaleidojulia_ifs_fold_julia", "bulbtetra_hybrid_bulb_ifs", "icosahedralifs_ifs",
    "icosahedraljulia_ifs_julia", "icosahedraltwist_ifs_twist", "diagonalicosa_ifs_fold", "hybridkifs_hybrid_ifs",
    "kaleidoscopicifs_ifs_twist", "juliabox_box_julia", "amazingsurf_box_fold", "mandelboxfoam_box_fold",
    "mandelboxsort_box_fold", "mandelboxdiag_box_fold", "mandelboxinv_box_fold", "mandelboxtwist_box_twist",
    "prismkifs_box_fold", "buffalo_box", "buffalojulia_box_julia", "buffalotwist_box_twist", "apollonian_kleinian",
    "apollonianjulia_kleinian_julia", "kleinianjulia_kleinian_julia", "pseudokleinian_kleinian",
    "juliabulb_bulb_julia", "mandelbulbmirror_bulb_fold", "mandelbulbtwist_bulb_twist", "mandelbulbdual_hybrid_bulb",
    "quatjulia_bulb_julia", "bulbbox_hybrid_bulb_box", "mandalay_box_fold", "wobblekifs_ifs_twist_julia",
    "radioactive_ifs_carve_twist", "recursivesphere_ifs_carve_twist", "tree_ifs", "kleinianlimit_kleinian",
    "fbmsdf_primitive_carve", "seascape_primitive", "dragon_ifs_twist", "romanesco_ifs",
]
ID = {f: 4 + i for i, f in enumerate(FILES)}
BUILTIN_FILES = {0: "mandelbrot", 1: "amazingbox_box", 2: "menger3_ifs", 3: "kleinian_kleinian"}
for k, v in BUILTIN_FILES.items():
    ID[v] = k
# the image formula: its tests bind an image (tests/test_image.py, tests/test_gpu_heightmap.py)
ID["heightmap_primitive"] = 4 + len(FILES)

_PARAMS = None


def param_table():
    global _PARAMS
    if _PARAMS is None:
        with open(os.path.join(HERE, "golden", "formula_params.json")) as fh:
            _PARAMS = json.load(fh)
    return _PARAMS


def defaults(name):
    """The 16 parameter slots of a formula file at its @param defaults (0 past the last)."""
    p = [r["default"] for r in param_table()[name]]
    return p + [0.0] * (16 - len(p))


def scene(name, params=None, cam=(0.0, 0.0, -3.5), size=(64, 48), max_steps=256, max_dist=20.0, tan_half_fov=0.5,
          eps_px=0.05):
    """One instance of the formula (code 0, bound to its builtin id through code_map=[id])
    seen from cam looking at the origin; vertical field of view 2 atan(tan_half_fov); hit
    epsilon eps_px pixels."""
    u = nd.config_uniforms(1, 0)
    u.resolution[0], u.resolution[1] = float(size[0]), float(size[1])
    for k in range(3):
        u.camera_pos[k] = cam[k]
    # camera basis toward the origin (world up +y), |up| = tan(fov / 2)
    fx, fy, fz = -cam[0], -cam[1], -cam[2]
    fl = (fx * fx + fy * fy + fz * fz) ** 0.5
    fx, fy, fz = fx / fl, fy / fl, fz / fl
    # right = normalize(forward x up_world), up = right x forward
    rx, ry, rz = fy * 0 - fz * 1, fz * 0 - fx * 0, fx * 1 - fy * 0
    rl = (rx * rx + ry * ry + rz * rz) ** 0.5
    rx, ry, rz = rx / rl, ry / rl, rz / rl
    ux, uy, uz = ry * fz - rz * fy, rz * fx - rx * fz, rx * fy - ry * fx
    for k, v in enumerate((fx, fy, fz)):
        u.camera_forward[k] = v
    # the ray model applies the aspect ratio itself: d = F + uv.x R + uv.y U
    for k, v in enumerate((rx, ry, rz)):
        u.camera_right[k] = v * tan_half_fov
    for k, v in enumerate((ux, uy, uz)):
        u.camera_up[k] = v * tan_half_fov
    u.epsilon_coefficient = eps_px * 2.0 * tan_half_fov / size[1]
    u.epsilon_floor = 0.0
    u.max_steps = float(max_steps)
    u.max_dist = float(max_dist)
    u.instance_count = 1.0
    inst = u.instances[0]
    inst.code = 0.0
    p = defaults(name) if params is None else list(params) + [0.0] * (16 - len(params))
    for k in range(16):
        inst.params[k] = float(p[k])
    return u, [ID[name]]


# Test scene of each formula: camera distance (on (0, 0.35 d, -d)), and @param overrides
# (slot -> value) where the defaults show little from outside the set. Chosen so that
# a fair share of the pixels of a 53-degree view hits.
SCENES = {
    "sphere_primitive": (3.0, {}), "cube_primitive": (3.0, {}), "cylinder_primitive": (3.0, {}),
    "torus_primitive": (2.0, {}), "pyramid_primitive": (2.0, {}), "prism_primitive": (3.0, {}),
    "sierpinski3_ifs": (2.0, {}), "cubejulia_ifs_julia": (2.0, {}), "juliakifs_ifs_julia": (2.0, {}),
    "diagonaltetra_ifs_fold": (2.0, {}), "starkifs_ifs_fold": (2.0, {}), "kaleidojulia_ifs_fold_julia": (2.0, {}),
    "bulbtetra_hybrid_bulb_ifs": (2.0, {}), "icosahedralifs_ifs": (2.0, {}), "icosahedraljulia_ifs_julia": (2.0, {}),
    "icosahedraltwist_ifs_twist": (2.0, {}), "diagonalicosa_ifs_fold": (2.0, {}), "hybridkifs_hybrid_ifs": (3.0, {}),
    "kaleidoscopicifs_ifs_twist": (3.0, {}), "juliabox_box_julia": (13.0, {}), "amazingsurf_box_fold": (20.0, {}),
    "mandelboxfoam_box_fold": (20.0, {}), "mandelboxsort_box_fold": (13.0, {}), "mandelboxdiag_box_fold": (20.0, {}),
    "mandelboxinv_box_fold": (9.0, {1: -2.0}), "mandelboxtwist_box_twist": (20.0, {}),
    "prismkifs_box_fold": (20.0, {}), "buffalo_box": (13.0, {}),
    "buffalojulia_box_julia": (9.0, {0: 1.537, 1: 0.926, 2: 1.537, 3: -1.863, 4: -1.276, 5: -0.849}),
    "buffalotwist_box_twist": (13.0, {}),
    "apollonian_kleinian": (2.0, {0: -0.137, 1: 0.417, 2: 2.035, 7: -0.999, 8: -1.598, 9: -0.094, 10: 0.558}),
    "apollonianjulia_kleinian_julia": (4.5, {0: -0.518, 1: 0.702, 2: 1.323, 3: 0.428, 4: 0.914, 5: -1.477}),
    "kleinianjulia_kleinian_julia": (3.0, {7: 8.0}), "pseudokleinian_kleinian": (20.0, {}),
    "juliabulb_bulb_julia": (3.0, {}), "mandelbulbmirror_bulb_fold": (3.0, {}), "mandelbulbtwist_bulb_twist": (3.0, {}),
    "mandelbulbdual_hybrid_bulb": (3.0, {}), "quatjulia_bulb_julia": (2.0, {}), "bulbbox_hybrid_bulb_box": (20.0, {1: 2.0}),
    "mandalay_box_fold": (3.0, {}), "wobblekifs_ifs_twist_julia": (20.0, {1: 8.0}),
    "radioactive_ifs_carve_twist": (9.0, {}), "recursivesphere_ifs_carve_twist": (4.5, {}), "tree_ifs": (2.0, {}),
    "kleinianlimit_kleinian": (6.5, {5: 8.0}), "fbmsdf_primitive_carve": (3.0, {}),
    "seascape_primitive": (3.0, {}), "dragon_ifs_twist": (1.0, {}), "romanesco_ifs": (2.0, {}),
}


def test_scene(name, size=(64, 48), **kw):
    """The test scene of a formula (see SCENES)."""
    d, over = SCENES[name]
    p = defaults(name)
    for k, v in over.items():
        p[k] = v
    return scene(name, params=p, cam=(0.0, 0.35 * d, -d), size=size, max_dist=max(20.0, 3.0 * d), **kw)


def hybrid_scene(base, mixins, base_n, total, base_params=None, cam=(0.0, 0.0, -3.5), size=(64, 48), **kw):
    """One hybrid instance: the base formula (code 0) for base_n steps of each cycle, then each
    mix-in (name, params or None, iterations; codes 1, 2, 3) for its iterations, total steps
    in all, finalized by the base formula. Returns (uniforms, code_map)."""
    u, _ = scene(base, params=base_params, cam=cam, size=size, **kw)
    inst = u.instances[0]
    inst.mixin_count = float(len(mixins))
    inst.hybrid_base_iters = float(base_n)
    inst.hybrid_total_iters = float(total)
    cm = [ID[base]]
    for j, (name, params, iters) in enumerate(mixins):
        m = inst.mixins[j]
        m.code = float(1 + j)
        m.iterations = float(iters)
        p = defaults(name) if params is None else list(params) + [0.0] * (16 - len(params))
        for k in range(16):
            m.params[k] = float(p[k])
        cm.append(ID[name])
    return u, cm


# Hybrid test scenes: name -> (base, base params or None, [(mix-in, params or None,
# iterations)], base_n, total, camera distance d). The camera sits at (0, 0.35 d, -d).
HYBRIDS = {
    "bulb2_box1": ("mandelbrot", [8.0, 6.0], [("amazingbox_box", None, 1)], 2, 9, 20.0),
    "box1_juliabox1": ("amazingbox_box", None, [("juliabox_box_julia", None, 1)], 1, 12, 35.0),
    "juliabulb_sierpinski_bulb3": ("juliabulb_bulb_julia", None,
                                   [("sierpinski3_ifs", [2.0, 12.0, 1, 1, 1], 1), ("mandelbrot", [3.0, 4.0], 1)],
                                   1, 8, 2.0),
}


def hybrid_test_scene(name, size=(64, 48)):
    base, bp, mix, bn, tot, d = HYBRIDS[name]
    return hybrid_scene(base, mix, bn, tot, base_params=bp, cam=(0.0, 0.35 * d, -d), size=size,
                        max_dist=max(20.0, 3.0 * d))
