"""TEST INFRASTRUCTURE ONLY: ctypes binding of the scalar CPU oracle (oracle/libndoracle.so).

The oracle restates the reference's WGSL march, G-buffer and soft shadow in C++ templated on
the scalar type (FP64 oracle, FP32 CPU restatement, FLOP counter, dual numbers). Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it, as the checker,
never as the thing measured or shipped.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libndoracle.so")


def build():
    r = subprocess.run(["make", "-C", _HERE, "-s"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, i32, dp = C.c_void_p, C.c_int, C.POINTER(C.c_double)
        L.ndo_abi_size.restype = i32
        L.ndo_scene_de.argtypes = [vp, vp, i32, vp, i32, vp, vp]
        L.ndo_march.argtypes = [vp, vp, i32, vp, vp, i32, vp, vp, vp, vp]
        L.ndo_normal.argtypes = [vp, vp, i32, i32, vp, vp, i32, vp]
        L.ndo_shadow.argtypes = [vp, vp, vp, vp, C.c_double, C.c_double, i32, vp, vp]
        L.ndo_primary_dir.argtypes = [vp, i32, i32, vp]
        L.ndo_march_inside.argtypes = [vp, vp, vp, vp, i32, vp]
        L.ndo_visibility.argtypes = [vp, vp, C.c_double, vp, vp, C.c_double, C.c_double, i32, vp]
        L.ndo_subsurface_path.argtypes = [vp, vp, C.c_double, vp, vp, i32, i32, i32, C.c_double, C.c_double, vp, vp]
        L.ndo_render_gbuffer.argtypes = [vp, vp, i32, i32, i32, i32, i32, i32, i32, i32, vp, vp, vp, vp, vp, vp,
                                         i32, vp]
        L.ndo_flops_per_de.argtypes = [vp, vp, vp, i32, vp]
        L.ndo_flops_per_de.restype = C.c_double
        L.ndo_helper.argtypes = [i32, vp, vp]
        L.ndo_formula_step.argtypes = [i32, vp, C.c_double, vp, vp, vp]
        L.ndo_bind_image.argtypes = [i32, vp, i32, i32, i32, vp]
        L.ndo_image_sample.argtypes = [i32, vp, i32, vp]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _cm(code_map):
    if code_map is None:
        return None, None
    a = np.asarray(list(code_map) + [-1] * (16 - len(code_map)), np.int32)
    return a, a.ctypes.data


def _u(u):
    return C.addressof(u)


def scene_de(u, pts, prec=0, trap=False, code_map=None):
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    d = np.zeros(len(pts))
    t = np.zeros(len(pts)) if trap else None
    cm, cmp_ = _cm(code_map)
    lib().ndo_scene_de(_u(u), cmp_, prec, _p(pts), len(pts), _p(d), _p(t))
    return (d, t) if trap else d


def march(u, origins, dirs, prec=0, code_map=None):
    o = np.ascontiguousarray(origins, np.float64).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
    n = len(o)
    hit = np.zeros(n, np.int32); t = np.zeros(n); steps = np.zeros(n, np.int32); ev = np.zeros(n, np.int32)
    cm, cmp_ = _cm(code_map)
    lib().ndo_march(_u(u), cmp_, prec, _p(o), _p(d), n, _p(hit), _p(t), _p(steps), _p(ev))
    return hit.astype(bool), t, steps, ev


def normal(u, pts, mode, e=None, code_map=None, prec=0):
    """Normals at pts: mode 0 analytic (FP64 only), 1 tetrahedral, 2 central with tap e;
    prec 1 gives the FP32 restatement of the stencils."""
    p = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    e = np.zeros(len(p)) if e is None else np.ascontiguousarray(np.broadcast_to(e, (len(p),)), np.float64)
    out = np.zeros((len(p), 3))
    cm, cmp_ = _cm(code_map)
    if lib().ndo_normal(_u(u), cmp_, prec, mode, _p(p), _p(e), len(p), _p(out)) != 0:
        raise ValueError("normal mode %d has no prec %d restatement" % (mode, prec))
    return out


def shadow(u, origin, direction, light_dist, k, steps, code_map=None):
    o = np.asarray(origin, np.float64); d = np.asarray(direction, np.float64)
    vis = np.zeros(1); ev = np.zeros(1, np.int32)
    cm, cmp_ = _cm(code_map)
    lib().ndo_shadow(_u(u), cmp_, _p(o), _p(d), float(light_dist), float(k), int(steps), _p(vis), _p(ev))
    return float(vis[0]), int(ev[0])


def spectral_basis(hero, rgb):
    """The 8-wavelength spectrum of an RGB colour in the hero wavelength's basis."""
    return helper(29, [hero] + list(rgb), 8)


def march_inside(u, origin, direction, steps, code_map=None):
    """(exited, t, DE evaluations) of the inside march from origin."""
    out = np.zeros(3)
    o = np.ascontiguousarray(origin, np.float64)
    d = np.ascontiguousarray(direction, np.float64)
    cm, cmp_ = _cm(code_map)
    lib().ndo_march_inside(_u(u), cmp_, _p(o), _p(d), int(steps), _p(out))
    return bool(out[0]), float(out[1]), int(out[2])


def visibility(u, hero, origin, direction, light_dist, k, steps, code_map=None):
    """The path trace's spectral shadow visibility: 8 lanes for the hero wavelength."""
    out = np.zeros(8)
    o = np.ascontiguousarray(origin, np.float64)
    d = np.ascontiguousarray(direction, np.float64)
    cm, cmp_ = _cm(code_map)
    lib().ndo_visibility(_u(u), cmp_, float(hero), _p(o), _p(d), float(light_dist), float(k), int(steps), _p(out))
    return out


def subsurface_path(u, hero, pos, normal, start_inside, light, occlusion, amount, radius, albedo, code_map=None):
    """subsurface_light_path toward light `light` from pos: (8-lane transmission, DE
    evaluations, blurred normal, 8-lane reach) for a material of colour albedo."""
    out = np.zeros(20)
    p = np.ascontiguousarray(pos, np.float64)
    n = np.ascontiguousarray(normal, np.float64)
    a = np.ascontiguousarray(albedo, np.float64)
    cm, cmp_ = _cm(code_map)
    lib().ndo_subsurface_path(_u(u), cmp_, float(hero), _p(p), _p(n), int(start_inside), int(light), int(occlusion),
                              float(amount), float(radius), _p(a), _p(out))
    return out[:8].copy(), int(out[8]), out[9:12].copy(), out[12:20].copy()


def sky(u, dirs):
    """sky_color for directions (n, 3): RGB (n, 3)."""
    L = lib()
    L.ndo_sky.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
    d = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
    out = np.zeros_like(d)
    L.ndo_sky(_u(u), _p(d), len(d), _p(out))
    return out


def calc_ao(u, pos, nrm, t, code_map=None):
    """calc_ao at points (n, 3) with normals (n, 3) and ray depths t (n,)."""
    L = lib()
    L.ndo_calc_ao.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
    p = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    nn = np.ascontiguousarray(nrm, np.float64).reshape(-1, 3)
    tt = np.ascontiguousarray(np.broadcast_to(t, (len(p),)), np.float64)
    out = np.zeros(len(p))
    cm, cmp_ = _cm(code_map)
    L.ndo_calc_ao(_u(u), cmp_, _p(p), _p(nn), _p(tt), len(p), _p(out))
    return out


def primary_dir(u, px, py):
    out = np.zeros(3)
    lib().ndo_primary_dir(_u(u), px, py, _p(out))
    return out


def render_gbuffer(u, rect=None, prec=0, normal_mode=0, shadows=True, shadow_max_steps=0, threads=None,
                   code_map=None, cam_offset=None):
    """G-buffer of a rect (x, y, w, h) as numpy arrays (same layout as the HIP outputs)."""
    W, H = int(u.resolution[0]), int(u.resolution[1])
    x, y, w, h = rect if rect else (0, 0, W, H)
    out = {"depth": np.zeros((h, w), np.float32), "normal": np.zeros((h, w, 3), np.float32),
           "trap": np.zeros((h, w), np.float32), "hit": np.zeros((h, w), np.uint8),
           "shadow": np.zeros((h, w), np.float32), "evals": np.zeros((h, w), np.uint32)}
    cm, cmp_ = _cm(code_map)
    co = None if cam_offset is None else np.asarray(cam_offset, np.float64)
    threads = threads or min(16, os.cpu_count() or 1)
    lib().ndo_render_gbuffer(_u(u), cmp_, prec, normal_mode, shadow_max_steps, 1 if shadows else 0, x, y, w, h,
                             _p(out["depth"]), _p(out["normal"]), _p(out["trap"]), _p(out["hit"]),
                             _p(out["shadow"]), _p(out["evals"]), threads, _p(co))
    return out


def gbuffer_color(u, rect=None, prec=0, threads=None, code_map=None):
    """fs_simple's colour attachment and fs_aov's albedo of a rect: (colour, albedo), each
    (h, w, 3) float32."""
    L = lib()
    vp, i32 = C.c_void_p, C.c_int
    L.ndo_gbuffer_color.argtypes = [vp, vp, i32, i32, i32, i32, i32, vp, vp, i32]
    W, H = int(u.resolution[0]), int(u.resolution[1])
    x, y, w, h = rect if rect else (0, 0, W, H)
    col = np.zeros((h, w, 3), np.float32)
    alb = np.zeros((h, w, 3), np.float32)
    cm, cmp_ = _cm(code_map)
    L.ndo_gbuffer_color(_u(u), cmp_, prec, x, y, w, h, _p(col), _p(alb), threads or min(16, os.cpu_count() or 1))
    return col, alb


def flops_per_de(u, pts, code_map=None):
    p = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    tops = np.zeros(1)
    cm, cmp_ = _cm(code_map)
    f = lib().ndo_flops_per_de(_u(u), cmp_, _p(p), len(p), _p(tops))
    return f, float(tops[0])


def light_sample(light_type, pos_or_dir, color, brightness, pos, ray_direction=(0.0, -1.0, 0.0), spread=0.0,
                 waist=0.05, max_dist=10.0):
    """light_sample at a surface point (FP64): (direction, colour, distance)."""
    a = [light_type, *pos_or_dir, *color, brightness, *ray_direction, spread, waist, max_dist, *pos]
    o = helper(40, a, 8)
    return o[0:3], o[3:6], o[6]


def helper(op, args, nout=4):
    a = np.zeros(64); a[:len(args)] = args
    out = np.zeros(nout)
    r = lib().ndo_helper(op, _p(a), _p(out))
    if r != 0:
        raise ValueError("unknown helper op %d" % op)
    return out


def formula_step(f, z, dr, pos, params):
    z = np.asarray(z, np.float64); pos = np.asarray(pos, np.float64)
    p = np.zeros(16); p[:len(params)] = params
    out = np.zeros(4)
    lib().ndo_formula_step(f, _p(z), float(dr), _p(pos), _p(p), _p(out))
    return out[:3], out[3]


def adjoint_at(u, pixels, t, code_map=None):
    """FP64 implicit-function gradient (6 parameters) at p = o + d t for the given depths."""
    L = lib()
    L.ndo_adjoint_at.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
    pix = np.ascontiguousarray(pixels, np.int32).reshape(-1, 2)
    tt = np.ascontiguousarray(t, np.float64).reshape(-1)
    g = np.zeros((len(pix), 6))
    cm, cmp_ = _cm(code_map)
    L.ndo_adjoint_at(_u(u), cmp_, _p(pix), _p(tt), len(pix), _p(g))
    return g


def adjoint_pixels(u, pixels, h=1e-3, code_map=None):
    """FP64 reference gradients of hit depth for pixels [(x, y), ...]: polished root depth,
    implicit-function-theorem gradient and central finite differences (6 parameters:
    bulb power, box scale, box min radius, camera x/y/z) and a per-parameter FD-valid mask."""
    L = lib()
    L.ndo_adjoint_pixels.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_void_p]
    pix = np.ascontiguousarray(pixels, np.int32).reshape(-1, 2)
    n = len(pix)
    t = np.zeros(n); gi = np.zeros((n, 6)); gf = np.zeros((n, 6)); ok = np.zeros(n, np.int32)
    cm, cmp_ = _cm(code_map)
    L.ndo_adjoint_pixels(_u(u), cmp_, _p(pix), n, float(h), _p(t), _p(gi), _p(gf), _p(ok))
    return t, gi, gf, ok


def radiance_pixels(u, pixels, t=None, h=1e-6, code_map=None):
    """FP64 direct-lighting colour of pixels [(x, y), ...] (Monte Carlo off, the reflection
    continuation cut) and its derivative with respect to (bulb power, box scale, box min radius,
    camera x/y/z): dual numbers at the hit, and central differences with step h at the
    perturbed scenes' own polished roots. t: per-pixel hit depths to take the dual at (no
    differences then). Returns dict t, rgb (n, 3), dual and fd (n, 3, 6), ok (bit k: FD valid)."""
    L = lib()
    L.ndo_radiance_pixels.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    pix = np.ascontiguousarray(pixels, np.int32).reshape(-1, 2)
    n = len(pix)
    tin = None if t is None else np.ascontiguousarray(t, np.float64)
    out = {"t": np.zeros(n), "rgb": np.zeros((n, 3)), "dual": np.zeros((n, 3, 6)), "fd": np.zeros((n, 3, 6)),
           "ok": np.zeros(n, np.int32)}
    cm, cmp_ = _cm(code_map)
    L.ndo_radiance_pixels(_u(u), cmp_, _p(pix), _p(tin) if tin is not None else None, n, float(h), _p(out["t"]),
                          _p(out["rgb"]), _p(out["dual"]), _p(out["fd"]), _p(out["ok"]))
    return out


def render_mc(u, spp=1, rect=None, prec=0, threads=None, code_map=None, per_sample=False):
    """Monte Carlo path trace (opaque scenes) of a rect: mean colour over spp samples
    (sample i uses mc_sample = u.mc_sample + i), depth and normal of sample 0, DE evaluations,
    and optionally the per-sample colours (spp, h, w, 3)."""
    L = lib()
    vp, i32 = C.c_void_p, C.c_int
    L.ndo_render_mc.argtypes = [vp, vp, i32, i32, i32, i32, i32, i32, vp, vp, vp, vp, vp, i32]
    W, H = int(u.resolution[0]), int(u.resolution[1])
    x, y, w, h = rect if rect else (0, 0, W, H)
    out = {"rgb": np.zeros((h, w, 3), np.float32), "depth": np.zeros((h, w), np.float32),
           "normal": np.zeros((h, w, 3), np.float32), "evals": np.zeros((h, w), np.uint32)}
    samp = np.zeros((spp, h, w, 3), np.float32) if per_sample else None
    cm, cmp_ = _cm(code_map)
    threads = threads or min(16, os.cpu_count() or 1)
    L.ndo_render_mc(_u(u), cmp_, prec, x, y, w, h, spp, _p(out["rgb"]), _p(out["depth"]), _p(out["normal"]),
                    _p(samp), _p(out["evals"]), threads)
    if per_sample:
        out["samples"] = samp
    return out


def bind_image(slot, words=None, w=0, h=0, slope=(0.0, 0.0, 0.0, 0.0)):
    """The oracle's copy of a height-map image slot: the packed texel words (4 per texel),
    size and per-channel slope bound; words None empties the slot."""
    L = lib()
    if words is None:
        assert L.ndo_bind_image(slot, None, 0, 0, 0, None) == 0
        return
    wd = np.ascontiguousarray(words, dtype=np.uint32)
    sl = np.ascontiguousarray(slope, dtype=np.float64)
    assert L.ndo_bind_image(slot, wd.ctypes.data, wd.size, int(w), int(h), sl.ctypes.data) == 0


def image_sample(slot, uv):
    """Bilinear image_sample of slot at points uv (n x 2): (value, near, far), each n x 4."""
    uv = np.ascontiguousarray(uv, dtype=np.float64).reshape(-1, 2)
    out = np.zeros((len(uv), 12))
    assert lib().ndo_image_sample(slot, uv.ctypes.data, len(uv), out.ctypes.data) == 0
    return out[:, 0:4], out[:, 4:8], out[:, 8:12]
