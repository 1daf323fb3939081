"""numericdream_amd - MI355X (gfx950) distance-estimator ray march and path trace.

Host-side Python binding of the C ABI in include/nd_hip.h (libndhip.so). It mirrors the
reference's renderer entry points for this path (src/gpu/fractal_renderer.zig:
renderToImage / drawSlot(.simple) / encodeWavefront) with the same argument meaning:
a 12,016-byte Uniforms block per sample, a pixel rect, and G-buffer outputs. Device
memory for outputs may be torch tensors on the GPU (their data pointers are passed with
device_ptrs = 1) or numpy arrays on the host.

The library must be built (numericdream_amd/build.py); importing fails loudly otherwise.
"""
import ctypes as C
import os

import numpy as np

from . import abi
from .abi import (BatchStats, GBufferOut, Job, MarchOpts, Rect, Tuning, Uniforms, AdjointIO, ImageOut,
                  TraceStats)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libndhip.so")


class NdError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("%s: %s" % (abi.STATUS_NAMES.get(status, status), message))
        self.status = status


def _load():
    # PyTorch-ROCm ships its own libamdhip64.so.7 under the same soname as the system one;
    # the first one loaded serves the whole process. Loading torch first makes both this
    # library and torch's tensors use one HIP runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError("numericdream_amd: %s is missing; run numericdream_amd/build.py "
                          "(or __graft_entry__.build()) first" % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    vp, u32, i32 = C.c_void_p, C.c_uint32, C.c_int
    P = C.POINTER
    lib.nd_ctx_create.argtypes = [i32, P(vp)]
    lib.nd_ctx_destroy.argtypes = [vp]
    lib.nd_ctx_destroy.restype = None
    lib.nd_last_error.argtypes = [vp]
    lib.nd_last_error.restype = C.c_char_p
    lib.nd_sync.argtypes = [vp]
    lib.nd_set_cancel.argtypes = [vp, P(C.c_int)]
    lib.nd_ctx_stream.argtypes = [vp]
    lib.nd_ctx_stream.restype = vp
    lib.nd_bind_formulas.argtypes = [vp, P(C.c_int32), u32]
    lib.nd_bind_image.argtypes = [vp, u32, i32, vp, u32, u32]
    lib.nd_image_pack.argtypes = [i32, vp, u32, u32, vp, C.c_size_t, P(abi.ImageDesc)]
    lib.nd_render_gbuffer.argtypes = [vp, vp, C.c_size_t, P(MarchOpts), Rect, P(GBufferOut), P(BatchStats)]
    lib.nd_render_gbuffer_noalloc.argtypes = [vp, vp, C.c_size_t, P(MarchOpts), Rect, P(GBufferOut), P(BatchStats)]
    lib.nd_render_batch.argtypes = [vp, P(Job), u32, P(MarchOpts), P(BatchStats)]
    lib.nd_batch_job_latency.argtypes = [vp, P(C.c_float), u32]
    lib.nd_get_tuning.argtypes = [vp, P(Tuning)]
    lib.nd_set_tuning.argtypes = [vp, P(Tuning)]
    lib.nd_autotune.argtypes = [vp, P(Job), u32, P(MarchOpts), P(Tuning)]
    lib.nd_adjoint.argtypes = [vp, P(Uniforms), u32, P(AdjointIO), P(BatchStats)]
    lib.nd_radiance_adjoint.argtypes = [vp, P(Uniforms), u32, P(abi.RadianceIO), P(BatchStats)]
    lib.nd_path_trace.argtypes = [vp, vp, C.c_size_t, u32, Rect, P(ImageOut), P(TraceStats)]
    lib.nd_accel_probe.argtypes = [vp, vp, C.c_size_t, vp, u32, vp]
    lib.nd_mesh_points.argtypes = [vp, vp, C.c_size_t, P(abi.MeshGrid), vp, u32, vp, vp]
    lib.nd_mesh_probe.argtypes = [vp, vp, C.c_size_t, vp, u32, vp, u32]
    lib.nd_config_job_count.argtypes = [u32]
    lib.nd_config_job_count.restype = u32
    lib.nd_uniforms_defaults.argtypes = [P(Uniforms)]
    lib.nd_uniforms_defaults.restype = None
    lib.nd_config_build.argtypes = [u32, u32, C.c_uint64, P(Uniforms)]
    lib.nd_camera_look_at.argtypes = [P(Uniforms), P(C.c_float), P(C.c_float), C.c_float]
    lib.nd_camera_look_at.restype = None
    lib.nd_set_pixel_epsilon.argtypes = [P(Uniforms), C.c_float, C.c_float]
    lib.nd_set_pixel_epsilon.restype = None
    lib.nd_accel_setup.argtypes = [P(Uniforms), u32, u32, C.c_float]
    lib.nd_accel_setup.restype = None
    lib.nd_scene_from_dream.argtypes = [C.c_char_p, C.c_size_t, u32, u32, P(Uniforms), P(abi.DreamInfo), C.c_char_p,
                                        C.c_size_t]
    return lib


lib = _load()


# ----------------------------------------------------------------------------- configs
def config_job_count(config):
    return int(lib.nd_config_job_count(config))


def config_uniforms(config, index=0, seed=0):
    """Uniforms block `index` of workload `config` (1..5), as include/nd_config.h builds it."""
    u = Uniforms()
    st = lib.nd_config_build(config, index, seed, C.byref(u))
    if st != 0:
        raise NdError(st, "nd_config_build(%d, %d)" % (config, index))
    return u


def load_dream(scene, width=0, height=0):
    """The Uniforms block an export of a .dream scene renders, and what else the file asks of
    the export (abi.DreamInfo), through nd_scene_from_dream. scene: a path, or the file's
    text (str or bytes). width / height 0 take the file's export size. Bind the scene's
    formulas with Renderer.bind_formulas(info.code_to_builtin[:info.n_codes])."""
    if isinstance(scene, str) and not scene.lstrip().startswith("{"):
        with open(scene, "rb") as fh:
            scene = fh.read()
    text = scene.encode() if isinstance(scene, str) else bytes(scene)
    u = Uniforms()
    info = abi.DreamInfo()
    err = C.create_string_buffer(512)
    st = lib.nd_scene_from_dream(text, len(text), width, height, C.byref(u), C.byref(info), err, len(err))
    if st != 0:
        raise NdError(st, "nd_scene_from_dream: %s" % err.value.decode(errors="replace"))
    return u, info


def default_uniforms():
    u = Uniforms()
    lib.nd_uniforms_defaults(C.byref(u))
    return u


def camera_look_at(u, pos, target, fov_deg=60.0):
    p = (C.c_float * 3)(*pos)
    t = (C.c_float * 3)(*target)
    lib.nd_camera_look_at(C.byref(u), p, t, fov_deg)


def set_pixel_epsilon(u, fov_deg=60.0, fraction=1e-3):
    lib.nd_set_pixel_epsilon(C.byref(u), fov_deg, fraction)


def enable_accel(u, res=64, levels=6, safety=0.92):
    """Turns on the empty-space accelerator with cascades around u's camera (the app's
    defaults: 64^3 cells, 6 levels, safety 0.92)."""
    lib.nd_accel_setup(C.byref(u), res, levels, safety)
    return u


def pixel_fraction(u, fov_deg=60.0):
    """The epsilon of u as a fraction of its pixel footprint."""
    import math
    theta = 2.0 * math.tan(math.radians(fov_deg) / 2) / float(u.resolution[1])
    return float(u.epsilon_coefficient) / theta


def resize(u, w, h, fov_deg=60.0):
    """Same camera at another resolution; the epsilon keeps its fraction of a pixel."""
    frac = pixel_fraction(u, fov_deg)
    u.resolution[0] = float(w)
    u.resolution[1] = float(h)
    set_pixel_epsilon(u, fov_deg, frac)
    return u


def uniforms_bytes(u):
    return bytes(C.string_at(C.addressof(u), abi.UNIFORMS_SIZE))


# ----------------------------------------------------------------------------- outputs
def _ptr(a):
    if a is None:
        return None
    if hasattr(a, "data_ptr"):  # torch tensor
        return a.data_ptr()
    return a.ctypes.data


def _image_array(pixels):
    a = np.ascontiguousarray(pixels)
    if a.ndim != 3 or a.shape[2] != 4:
        raise ValueError("an image is H x W x 4")
    if a.dtype == np.uint16:
        return a, abi.IMAGE_RGBA16
    if a.dtype == np.float32:
        return a, abi.IMAGE_RGBA32F
    raise ValueError("image pixels are uint16 or float32")


def image_pack(pixels):
    """The texel words nd_bind_image makes of an image (host only): (words as uint32 array,
    width, height, per-channel slope bound)."""
    a, fmt = _image_array(pixels)
    d = abi.ImageDesc()
    st = lib.nd_image_pack(fmt, a.ctypes.data, a.shape[1], a.shape[0], None, 0, C.byref(d))
    if st != abi.ND_OK:
        raise NdError(st, "nd_image_pack")
    words = np.zeros(d.words, dtype=np.uint32)
    st = lib.nd_image_pack(fmt, a.ctypes.data, a.shape[1], a.shape[0], words.ctypes.data, words.size, C.byref(d))
    if st != abi.ND_OK:
        raise NdError(st, "nd_image_pack")
    return words, d.width, d.height, np.array(d.slope[:], dtype=np.float32)


class GBuffer:
    """Host (numpy) or device (torch) G-buffer arrays for a w x h rect."""

    FIELDS = ("depth", "normal", "position", "trap", "hit", "shadow", "evals", "color", "albedo")

    def __init__(self, w, h, device=None, fields=("depth", "normal", "position", "trap", "hit", "shadow", "evals")):
        """fields: the outputs to allocate; color and albedo (fs_simple's colour attachment and
        fs_aov's albedo, 3 floats per pixel each) are asked for by naming them."""
        self.w, self.h = w, h
        self.device = device
        shapes = {"depth": ((h, w), "f4"), "normal": ((h, w, 3), "f4"), "position": ((h, w, 3), "f4"),
                  "trap": ((h, w), "f4"), "hit": ((h, w), "u1"), "shadow": ((h, w), "f4"), "evals": ((h, w), "u4"),
                  "color": ((h, w, 3), "f4"), "albedo": ((h, w, 3), "f4")}
        for k in self.FIELDS:
            if k not in fields:
                setattr(self, k, None)
                continue
            shape, dt = shapes[k]
            if device is None:
                setattr(self, k, np.zeros(shape, dt))
            else:
                import torch
                tdt = {"f4": torch.float32, "u1": torch.uint8, "u4": torch.int32}[dt]
                setattr(self, k, torch.zeros(shape, dtype=tdt, device=device))

    def out(self):
        return GBufferOut(_ptr(self.depth), _ptr(self.normal), _ptr(self.position), _ptr(self.trap), _ptr(self.hit),
                          _ptr(self.shadow), _ptr(self.evals), _ptr(self.color), _ptr(self.albedo),
                          0 if self.device is None else 1, 0)

    def numpy(self):
        r = {}
        for k in self.FIELDS:
            a = getattr(self, k)
            if a is None:
                continue
            r[k] = a if isinstance(a, np.ndarray) else a.cpu().numpy()
        if "evals" in r:
            r["evals"] = r["evals"].view(np.uint32)
        return r


# ----------------------------------------------------------------------------- renderer
class Renderer:
    """One HIP context (device + stream), like the reference's FractalRenderer."""

    def __init__(self, device=0):
        h = C.c_void_p()
        st = lib.nd_ctx_create(device, C.byref(h))
        if st != 0:
            raise NdError(st, "nd_ctx_create(%d)" % device)
        self._h = h

    def close(self):
        if self._h:
            lib.nd_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, what):
        if st != 0:
            raise NdError(st, "%s: %s" % (what, lib.nd_last_error(self._h).decode()))

    @property
    def stream(self):
        return lib.nd_ctx_stream(self._h)

    def bind_formulas(self, code_to_builtin):
        arr = (C.c_int32 * len(code_to_builtin))(*code_to_builtin)
        self._check(lib.nd_bind_formulas(self._h, arr, len(code_to_builtin)), "nd_bind_formulas")

    def bind_image(self, slot, pixels):
        """Binds a decoded image (H x W x 4: uint16 straight-alpha RGBA, or float32 HDR RGBA) to
        image slot `slot` (instance * 4, + 1 + j for mix-in j); None unbinds the slot."""
        if pixels is None:
            self._check(lib.nd_bind_image(self._h, slot, 0, None, 0, 0), "nd_bind_image")
            return
        a, fmt = _image_array(pixels)
        self._check(lib.nd_bind_image(self._h, slot, fmt, a.ctypes.data, a.shape[1], a.shape[0]), "nd_bind_image")

    def render_gbuffer(self, u, gbuf, rect=None, normal_mode=abi.NORMAL_ANALYTIC, shadows=True, shadow_max_steps=0,
                       kernel=abi.KERNEL_PERSISTENT):
        opts = MarchOpts(normal_mode, shadow_max_steps, 1 if shadows else 0, kernel)
        r = Rect(*(rect or (0, 0, 0, 0)))
        st = BatchStats()
        out = gbuf.out()
        self._check(lib.nd_render_gbuffer(self._h, C.byref(u), abi.UNIFORMS_SIZE, C.byref(opts), r, C.byref(out),
                                          C.byref(st)), "nd_render_gbuffer")
        return st

    def render_gbuffer_noalloc(self, u, gbuf, rect=None, normal_mode=abi.NORMAL_ANALYTIC, shadows=True,
                               shadow_max_steps=0, kernel=abi.KERNEL_PERSISTENT):
        """render_gbuffer without any device allocation: gbuf must hold device tensors, the rect
        at most abi.NOALLOC_MAX_RAYS ray slots (64 per 8x8 tile), no accelerator."""
        opts = MarchOpts(normal_mode, shadow_max_steps, 1 if shadows else 0, kernel)
        r = Rect(*(rect or (0, 0, 0, 0)))
        st = BatchStats()
        out = gbuf.out()
        self._check(lib.nd_render_gbuffer_noalloc(self._h, C.byref(u), abi.UNIFORMS_SIZE, C.byref(opts), r,
                                                  C.byref(out), C.byref(st)), "nd_render_gbuffer_noalloc")
        return st

    def render_batch(self, jobs, normal_mode=abi.NORMAL_ANALYTIC, shadows=True, shadow_max_steps=0,
                     kernel=abi.KERNEL_PERSISTENT):
        """jobs: list of (Uniforms, GBuffer[, rect])."""
        arr = (Job * len(jobs))()
        keep = []
        for i, jb in enumerate(jobs):
            u, g = jb[0], jb[1]
            rect = jb[2] if len(jb) > 2 else (0, 0, 0, 0)
            keep.append(u)
            arr[i].uniforms = C.pointer(u)
            arr[i].rect = Rect(*rect)
            arr[i].out = g.out()
        opts = MarchOpts(normal_mode, shadow_max_steps, 1 if shadows else 0, kernel)
        st = BatchStats()
        self._check(lib.nd_render_batch(self._h, arr, len(jobs), C.byref(opts), C.byref(st)), "nd_render_batch")
        return st

    def job_latency_ms(self, n):
        a = (C.c_float * n)()
        self._check(lib.nd_batch_job_latency(self._h, a, n), "nd_batch_job_latency")
        return np.array(a[:], dtype=np.float32)

    def tuning(self):
        t = Tuning()
        self._check(lib.nd_get_tuning(self._h, C.byref(t)), "nd_get_tuning")
        return t

    def set_tuning(self, waves_per_cu=0, refill_min_idle=8, chunk_rays=0, track_latency=False, graphs=True,
                   round_cap=0, tile_width=0, refill_period=0, gbuffer_graphs=False, mc_pool_log2=0):
        """round_cap > 0 stops every wave of a persistent geometry pass after that many rounds;
        the call then raises NdError(ND_E_TIMEOUT). tile_width: 4, 8, 16 or 32 pixels per wave64
        tile row (0 = 8); refill_period: march steps between refills (0 = 1). graphs: the path
        trace replays a captured hipGraph of each chunk's passes; gbuffer_graphs: the persistent
        G-buffer passes are captured when a batch repeats the last launch record; mc_pool_log2: the
        path trace's pool of 2^k paths per chunk of samples (0 = sized by device memory)."""
        flags = ((abi.TUNE_TRACK_LATENCY if track_latency else 0) | (0 if graphs else abi.TUNE_NO_GRAPH)
                 | (abi.TUNE_GBUFFER_GRAPH if gbuffer_graphs else 0))
        t = Tuning(waves_per_cu, refill_min_idle, chunk_rays, flags, round_cap, tile_width, refill_period, mc_pool_log2)
        self._check(lib.nd_set_tuning(self._h, C.byref(t)), "nd_set_tuning")

    def set_cancel(self, flag):
        """Watches flag (a ctypes.c_int the caller keeps alive, or None to stop watching), like
        the export cancel flag (src/app/export_image.zig:46-47): while it is non-zero, calls
        return ND_E_CANCELLED; the path trace checks it between chunks of samples."""
        self._cancel = flag
        ptr = C.pointer(flag) if flag is not None else None
        self._check(lib.nd_set_cancel(self._h, ptr), "nd_set_cancel")

    def autotune(self, jobs, normal_mode=abi.NORMAL_ANALYTIC, shadows=True, shadow_max_steps=0):
        arr = (Job * len(jobs))()
        for i, jb in enumerate(jobs):
            arr[i].uniforms = C.pointer(jb[0])
            arr[i].rect = Rect(*(jb[2] if len(jb) > 2 else (0, 0, 0, 0)))
            arr[i].out = jb[1].out()
        opts = MarchOpts(normal_mode, shadow_max_steps, 1 if shadows else 0, abi.KERNEL_PERSISTENT)
        t = Tuning()
        self._check(lib.nd_autotune(self._h, arr, len(jobs), C.byref(opts), C.byref(t)), "nd_autotune")
        return t

    def sync(self):
        self._check(lib.nd_sync(self._h), "nd_sync")

    def adjoint(self, views, target=None, depth_out=None, dtdp_out=None):
        """Depth-loss gradient over views (list of Uniforms with the current parameters).

        target / depth_out / dtdp_out: per-view depth maps concatenated in view order
        (W*H floats each; dtdp has 6 floats per pixel), all numpy (host) or all torch
        tensors on the device. Returns (loss, grad[6], pixels_used, stats).
        Parameter order: bulb power, box scale, box min radius, camera x, y, z."""
        n = len(views)
        arr = (Uniforms * n)(*views)
        dev = any(x is not None and hasattr(x, "data_ptr") for x in (target, depth_out, dtdp_out))
        grad = (C.c_double * 6)()
        loss = C.c_double()
        used = C.c_uint64()
        io = AdjointIO(_ptr(target), _ptr(depth_out), _ptr(dtdp_out), C.addressof(grad), C.addressof(loss),
                       C.addressof(used), 1 if dev else 0, 0)
        st = BatchStats()
        self._check(lib.nd_adjoint(self._h, arr, n, C.byref(io), C.byref(st)), "nd_adjoint")
        return loss.value, np.array(grad[:]), int(used.value), st

    def radiance_adjoint(self, views, target=None, rgb_out=None, drgb_out=None):
        """Colour-loss gradient of the direct-lighting images of views (nd_radiance_adjoint).

        target / rgb_out / drgb_out: per-view maps concatenated in view order (W*H*3 floats;
        drgb has 18 per pixel, [c * 6 + k]), all numpy (host) or all torch tensors on the
        device; a pixel whose target red is negative is left out of the loss. Returns (loss,
        grad[6], pixels_used, stats) with the parameter order of adjoint()."""
        n = len(views)
        arr = (Uniforms * n)(*views)
        dev = any(x is not None and hasattr(x, "data_ptr") for x in (target, rgb_out, drgb_out))
        grad = (C.c_double * 6)()
        loss = C.c_double()
        used = C.c_uint64()
        io = abi.RadianceIO(_ptr(target), _ptr(rgb_out), _ptr(drgb_out), C.addressof(grad), C.addressof(loss),
                            C.addressof(used), 1 if dev else 0, 0)
        st = BatchStats()
        self._check(lib.nd_radiance_adjoint(self._h, arr, n, C.byref(io), C.byref(st)), "nd_radiance_adjoint")
        return loss.value, np.array(grad[:]), int(used.value), st

    def mesh_points(self, u, origin, cell, iso, span, local_offset, blocks, values=None):
        """cs_mesh_points: the DE minus iso on span^3 points of every block (blocks: (n, 4) u32
        or a device tensor). values: float32 (n * span^3,) numpy array or device tensor,
        allocated on the host when omitted. Returns (values, kernel ms)."""
        dev = hasattr(blocks, "data_ptr")
        if not dev:
            blocks = np.ascontiguousarray(blocks, np.uint32).reshape(-1, 4)
        n = int(blocks.shape[0])
        if values is None:
            values = np.zeros(n * span ** 3, np.float32)
        g = abi.MeshGrid((C.c_float * 3)(*origin), cell, iso, span, local_offset, 1 if dev else 0)
        ms = C.c_float(0.0)
        ub = Uniforms.from_buffer_copy(bytes(u))
        self._check(lib.nd_mesh_points(self._h, C.byref(ub), C.sizeof(Uniforms), C.byref(g), _ptr(blocks), n,
                                       _ptr(values), C.byref(ms)), "nd_mesh_points")
        return values, ms.value

    def mesh_probe(self, u, points):
        """cs_mesh_probe: (d, gradient) at points (n, 4) = (x, y, z, h); h < 0 without carves."""
        p = np.ascontiguousarray(points, np.float32).reshape(-1, 4)
        out = np.zeros((len(p), 4), np.float32)
        ub = Uniforms.from_buffer_copy(bytes(u))
        self._check(lib.nd_mesh_probe(self._h, C.byref(ub), C.sizeof(Uniforms), p.ctypes.data, len(p),
                                      out.ctypes.data, 0), "nd_mesh_probe")
        return out

    def accel_probe(self, u, pts):
        """The accelerator's skip distance at points (n, 3) with the device grid of u's scene
        and cascades (float32, host)."""
        p = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
        out = np.zeros(len(p), np.float32)
        ub = Uniforms.from_buffer_copy(bytes(u))
        self._check(lib.nd_accel_probe(self._h, C.byref(ub), C.sizeof(Uniforms), p.ctypes.data, len(p),
                                       out.ctypes.data), "nd_accel_probe")
        return out

    def path_trace(self, u, spp, rect=None, color=None, depth=None, normal=None):
        """Monte Carlo path trace of rect (default: the whole frame) with spp samples per pixel
        (sample i uses mc_sample = u.mc_sample + i). color (h, w, 3), depth (h, w) and normal
        (h, w, 3) are float32 numpy arrays or torch tensors on the device; color is allocated
        on the host when omitted. Returns (color, depth, normal, stats)."""
        W, H = int(u.resolution[0]), int(u.resolution[1])
        x, y, w, h = rect if rect is not None else (0, 0, W, H)
        if color is None:
            color = np.zeros((h, w, 3), np.float32)
        dev = hasattr(color, "data_ptr")
        out = ImageOut(_ptr(color), _ptr(depth), _ptr(normal), 1 if dev else 0, 0)
        st = TraceStats()
        ub = Uniforms.from_buffer_copy(bytes(u))
        self._check(lib.nd_path_trace(self._h, C.byref(ub), C.sizeof(Uniforms), spp, Rect(x, y, w, h), C.byref(out),
                                      C.byref(st)), "nd_path_trace")
        return color, depth, normal, st
