"""Benchmark of the distance-estimator ray-march pass on one MI355X.

Workload (one "step"): the 1080p Mandelbox batch - 16 seeded cameras on a sphere of radius 4,
1920x1080 each, amazingbox (-1.5, .5, 1, 1, 15 iterations), G-buffer (depth, normal, trap,
hit) plus one soft-shadow ray per hit toward the global light (.577, .577, -.577), k = 16,
at most 128 shadow steps. Outputs stay in HBM (torch tensors); the persistent batched
schedule renders the 16 frames in one launch per pass, the plain schedule launches one
kernel per frame. value = primary rays (with their shadow ray) per second of the persistent
schedule, whole job, wall clock around the timed steps.

Prints one JSON line. --gpus N > 1 exits: the path runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3         # MI355X FP32 vector (published)
METRIC = "Mrays/s persistent-batched vs plain @1080p Mandelbox, %FP32 peak; grad rays/s"
# HBM bytes per k_march launch on the 16-camera batch, from two rocprofv3 --pmc runs
# (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, factors from a 1 GiB copy; profiles/k_march_c2_traffic.md).
# Hardware counters cannot be read inside the timed run, so the measured figure is carried here.
K_MARCH_TRAFFIC_BYTES_16CAM = 2.477e9
# Hardware counters of the same k_march launch (profiles/k_march_c2_valu_counters.md, two
# rocprofv3 --pmc runs): active lanes per VALU instruction, the VALU lane-instruction rate as a
# fraction of the 128 FP32 lanes/clk/CU the FP32 peak implies, the FP32 FLOP rate the
# hardware counts (FMA 2, add/mul/transcendental 1; folds and clamps are not FLOPs to it), and
# the share of quad-cycles in which a SIMD issued two VALU instructions.
K_MARCH_COUNTERS_16CAM = {"lane_util": 0.902, "valu_lane_rate_frac": 0.631, "hw_fp32_tflops": 46.9,
                          "dual_issue_quad_cycles": 0.533, "source": "profiles/k_march_c2_valu_counters.md"}


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def flops_table():
    with open(os.path.join(ROOT, "tests", "golden", "flops_per_de.json")) as f:
        return json.load(f)


def make_jobs(nd, torch, n_cams):
    jobs = []
    for i in range(n_cams):
        u = nd.config_uniforms(2, i)
        W, H = int(u.resolution[0]), int(u.resolution[1])
        g = nd.GBuffer(W, H, device="cuda:0", fields=("depth", "normal", "trap", "hit", "shadow"))
        jobs.append((u, g))
    return jobs


def timed(r, jobs, kernel, steps, warmup, torch, abi, label):
    stats = None
    for i in range(warmup):
        stats = r.render_batch(jobs, shadows=True, shadow_max_steps=128, kernel=kernel)
        log("%s warmup %d: %.1f ms device" % (label, i, stats.ms_total))
    torch.cuda.synchronize()
    r.sync()
    agg = {"ms_march": 0.0, "ms_total": 0.0, "de_march": 0, "de_total": 0, "rays": 0, "hits": 0, "rounds": 0}
    t0 = time.perf_counter()
    for i in range(steps):
        stats = r.render_batch(jobs, shadows=True, shadow_max_steps=128, kernel=kernel)
        agg["ms_march"] += stats.ms_march
        agg["ms_total"] += stats.ms_total
        agg["de_march"] += stats.de_evals_march
        agg["de_total"] += stats.de_evals
        agg["rays"] += stats.rays
        agg["hits"] += stats.hits
        agg["rounds"] += stats.march_wave_rounds
        log("%s step %d: %.1f ms device, %d DE evals" % (label, i, stats.ms_total, stats.de_evals))
    r.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    agg["wall_s"] = wall
    return agg


def batch_jobs_leg(r, nd, torch, abi, reps=3):
    """256 uneven G-buffer jobs (Mandelbulb / Mandelbox, 64..1024 px, 8..20 iterations) in one
    persistent launch per pass, against one plain launch per job."""
    jobs = []
    for i in range(nd.config_job_count(4)):
        u = nd.config_uniforms(4, i)
        w, h = int(u.resolution[0]), int(u.resolution[1])
        jobs.append((u, nd.GBuffer(w, h, device="cuda:0", fields=("depth", "normal", "trap", "hit"))))
    rays = sum(j[1].w * j[1].h for j in jobs)
    t = r.tuning()
    r.set_tuning(t.waves_per_cu, t.refill_min_idle, t.chunk_rays, track_latency=True)
    res = {}
    for name, kernel in (("persistent", abi.KERNEL_PERSISTENT), ("plain", abi.KERNEL_PLAIN)):
        r.render_batch(jobs, shadows=False, kernel=kernel)  # warmup
        r.sync()
        t0 = time.perf_counter()
        ms = []
        for _ in range(reps):
            st = r.render_batch(jobs, shadows=False, kernel=kernel)
            ms.append(st.ms_total)
        wall = (time.perf_counter() - t0) / reps
        res[name] = {"ms_per_batch": round(wall * 1e3, 3), "device_ms": round(min(ms), 3),
                     "jobs_per_s": round(len(jobs) / wall, 1), "mrays_per_s": round(rays / wall / 1e6, 2)}
        if kernel == abi.KERNEL_PERSISTENT:
            lat = np.sort(r.job_latency_ms(len(jobs)))
            res[name]["latency_ms_p50"] = round(float(np.percentile(lat, 50)), 3)
            res[name]["latency_ms_p99"] = round(float(np.percentile(lat, 99)), 3)
            res[name]["de_evals"] = int(st.de_evals)
        log("c4 %s: %.1f ms per batch of %d jobs" % (name, wall * 1e3, len(jobs)))
    r.set_tuning(t.waves_per_cu, t.refill_min_idle, t.chunk_rays, track_latency=False)
    res["jobs"] = len(jobs)
    res["rays"] = rays
    res["persistent_vs_plain"] = round(res["persistent"]["mrays_per_s"] / res["plain"]["mrays_per_s"], 3)
    return res


def gradient_leg(r, nd, torch, reps=3):
    """Differentiable render: 8 views of 512x512, depth loss against a target, gradient of the
    loss with respect to (bulb power, box scale, box min radius, camera x/y/z)."""
    views = [nd.config_uniforms(5, i) for i in range(nd.config_job_count(5))]
    total = sum(int(v.resolution[0]) * int(v.resolution[1]) for v in views)
    target = torch.zeros(total, dtype=torch.float32, device="cuda:0")
    r.adjoint(views, depth_out=target)
    for v in views:  # perturbed parameters: the loss is not zero
        v.instances[0].params[0] = 8.2
    r.adjoint(views, target=target)
    r.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        loss, grad, used, st = r.adjoint(views, target=target)
    wall = (time.perf_counter() - t0) / reps
    res = {"grad_rays_per_s": round(total / wall, 1), "ms_per_call": round(wall * 1e3, 3),
           "device_ms": round(st.ms_total, 3), "adjoint_pass_ms": round(st.ms_surface, 3),
           "views": len(views), "pixels": total, "pixels_in_loss": used}
    # the colour loss of the same views lit by a point, a global and a ray light with soft
    # shadows (nd_radiance_adjoint: Monte Carlo off, reflection continuation cut)
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from tests.radiance_scene import radiance_scene
    cviews = [radiance_scene(view=i, size=(512, 512)) for i in range(len(views))]
    ctarget = torch.zeros(3 * total, dtype=torch.float32, device="cuda:0")
    r.radiance_adjoint(cviews, rgb_out=ctarget)
    for v in cviews:
        v.instances[0].params[0] = 8.2
    r.radiance_adjoint(cviews, target=ctarget)
    r.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        closs, cgrad, cused, cst = r.radiance_adjoint(cviews, target=ctarget)
    cwall = (time.perf_counter() - t0) / reps
    res["radiance"] = {"grad_rays_per_s": round(total / cwall, 1), "ms_per_call": round(cwall * 1e3, 3),
                       "device_ms": round(cst.ms_total, 3), "adjoint_pass_ms": round(cst.ms_surface, 3),
                       "pixels_in_loss": cused, "lights": 3}
    return res


def c1_leg(r, nd, torch, abi, flops, reps=5):
    """The Mandelbulb G-buffer workload: power 8, 8 iterations, 640x360, camera (0, 0, -2.5),
    depth + analytic normal + trap + hit, no lighting; persistent and plain schedules."""
    u = nd.config_uniforms(1, 0)
    W, H = int(u.resolution[0]), int(u.resolution[1])
    g = nd.GBuffer(W, H, device="cuda:0", fields=("depth", "normal", "trap", "hit"))
    f = flops["workloads"]["1"]["march"]["flops_per_de"]
    res = {"width": W, "height": H, "flops_per_de": round(f, 2)}
    for name, kernel in (("persistent", abi.KERNEL_PERSISTENT), ("plain", abi.KERNEL_PLAIN)):
        r.render_batch([(u, g)], shadows=False, kernel=kernel)
        r.sync()
        t0 = time.perf_counter()
        de = ms = ms_march = de_march = 0.0
        for _ in range(reps):
            st = r.render_batch([(u, g)], shadows=False, kernel=kernel)
            de += st.de_evals
            ms += st.ms_total
            ms_march += st.ms_march if kernel == abi.KERNEL_PERSISTENT else st.ms_total
            de_march += st.de_evals_march
        wall = (time.perf_counter() - t0) / reps
        tfl = f * de_march / (ms_march * 1e-3) / 1e12 if ms_march > 0 else 0.0
        res[name] = {"mrays_per_s": round(W * H / wall / 1e6, 2), "ms_per_frame": round(wall * 1e3, 3),
                     "device_ms": round(ms / reps, 3), "de_evals_per_s": round(de / (ms * 1e-3), 1),
                     "march_tflops": round(tfl, 3), "march_frac_fp32_peak": round(tfl / FP32_PEAK_TFLOPS, 4)}
        log("c1 %s: %.3f ms per frame" % (name, wall * 1e3))
    return res


def mesh_leg(r, nd, torch, flops, n=512, reps=3):
    """The mesher's distance pass (cs_mesh_points) over an n^3 grid of the Mandelbox scene in
    8^3 blocks, blocks and values in device memory: DE evaluations per second and FP32 rate."""
    u = nd.config_uniforms(2, 0)
    b = n // 8
    z, y, x = torch.meshgrid(torch.arange(b), torch.arange(b), torch.arange(b), indexing="ij")
    blocks = torch.stack([x.reshape(-1), y.reshape(-1), z.reshape(-1), torch.zeros(b ** 3, dtype=torch.int64)], 1)
    blocks = blocks.to(torch.int32).contiguous().to("cuda:0")
    vals = torch.empty(n ** 3, dtype=torch.float32, device="cuda:0")
    origin, cell = (-4.0, -4.0, -4.0), 8.0 / n
    r.mesh_points(u, origin, cell, 0.0, 8, 0, blocks, values=vals)
    ms = []
    for _ in range(reps):
        ms.append(r.mesh_points(u, origin, cell, 0.0, 8, 0, blocks, values=vals)[1])
    k = sum(ms) / len(ms)
    # the Mandelbox DE runs all 15 iterations wherever it is evaluated: the march's mean
    # FLOP per DE on this scene stands for the grid's
    f = flops["workloads"]["2"]["march"]["flops_per_de"]
    return {"grid": n, "points": n ** 3, "kernel_ms": round(k, 3), "de_evals_per_s": round(n ** 3 / (k * 1e-3), 1),
            "tflops": round(f * n ** 3 / (k * 1e-3) / 1e12, 3),
            "write_gb_per_s": round(4 * n ** 3 / (k * 1e-3) / 1e9, 2)}


def accel_leg(r, nd, torch, abi, jobs, steps=2):
    """The same 16-camera batch with the empty-space accelerator as the app enables it by
    default (64^3 cells x 6 levels, safety 0.92; one grid per camera, built on the device in
    the warmup and reused from the cache afterwards)."""
    acc_jobs = []
    for u, g in jobs:
        ua = nd.abi.Uniforms.from_buffer_copy(bytes(u))
        nd.enable_accel(ua, res=64, levels=6, safety=0.92)
        acc_jobs.append((ua, g))
    res = {}
    for name, kernel in (("persistent", abi.KERNEL_PERSISTENT), ("plain", abi.KERNEL_PLAIN)):
        r.render_batch(acc_jobs, shadows=True, shadow_max_steps=128, kernel=kernel)  # builds / caches the grids
        r.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            st = r.render_batch(acc_jobs, shadows=True, shadow_max_steps=128, kernel=kernel)
        r.sync()
        wall = (time.perf_counter() - t0) / steps
        res[name] = {"mrays_per_s": round(st.rays / wall / 1e6, 2), "ms_per_step": round(wall * 1e3, 3),
                     "de_evals": int(st.de_evals)}
        log("c2+accel %s: %.1f ms per step" % (name, wall * 1e3))
    res["persistent_vs_plain"] = round(res["persistent"]["mrays_per_s"] / res["plain"]["mrays_per_s"], 3)
    return res


def mc_leg(r, nd, spp=64):
    """Wavefront Monte Carlo path trace of the four-fractal smooth-union scene at 1920x1080 and
    64 samples per pixel (one frame)."""
    u = nd.config_uniforms(3, 0)
    W, H = int(u.resolution[0]), int(u.resolution[1])
    import torch
    col = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda:0")
    r.path_trace(u, 1, color=col)  # warmup (captures the per-chunk graph)
    r.sync()
    t0 = time.perf_counter()
    _, _, _, st = r.path_trace(u, spp, color=col)
    r.sync()
    wall = time.perf_counter() - t0
    log("c3 mc: %.1f ms for %d spp, %d segments" % (wall * 1e3, spp, st.segments))
    # progressive rendering: one call per sample, as an interactive host issues them; the
    # chunk's passes come from the captured hipGraph or are launched one by one
    per_sample = {}
    t = r.tuning()
    for name, graphs in (("graph", True), ("launches", False)):
        r.set_tuning(t.waves_per_cu, t.refill_min_idle, t.chunk_rays, graphs=graphs)
        r.path_trace(u, 1, color=col)
        r.sync()
        t0 = time.perf_counter()
        for i in range(min(8, spp)):
            u.mc_sample = float(i)
            r.path_trace(u, 1, color=col)
        r.sync()
        per_sample[name] = round((time.perf_counter() - t0) / min(8, spp) * 1e3, 3)
    r.set_tuning(t.waves_per_cu, t.refill_min_idle, t.chunk_rays)
    u.mc_sample = 0.0
    # the same frame with the primary hit cache on (primary_cache_key, as the reference's
    # multi-sample export sets it for a static camera): samples after the first reuse the
    # primary segment and the packed primary normal
    uc = nd.config_uniforms(3, 0)
    uc.primary_cache_key = 1.0
    r.path_trace(uc, 1, color=col)
    r.sync()
    t0 = time.perf_counter()
    _, _, _, stc = r.path_trace(uc, spp, color=col)
    r.sync()
    wall_c = time.perf_counter() - t0
    log("c3 mc with the primary hit cache: %.1f ms for %d spp" % (wall_c * 1e3, spp))
    cached = {"ms_per_frame": round(wall_c * 1e3, 2), "msamples_per_s": round(stc.paths / wall_c / 1e6, 3),
              "de_evals_per_path": round(stc.de_evals / stc.paths, 1)}
    # the same frame with adaptive sampling (threshold 0.02 from 16 samples, the app's
    # defaults): converged pixels start no more paths
    ua = nd.config_uniforms(3, 0)
    ua.adapt_enabled = 1.0
    ua.adapt_row_width = float(W)
    r.path_trace(ua, 1, color=col)
    r.sync()
    t0 = time.perf_counter()
    _, _, _, sta = r.path_trace(ua, spp, color=col)
    r.sync()
    wall_a = time.perf_counter() - t0
    log("c3 mc adaptive: %.1f ms for %d spp, %d segments (%d without)" % (wall_a * 1e3, spp, sta.segments, st.segments))
    adaptive = {"ms_per_frame": round(wall_a * 1e3, 2), "segments_per_path": round(sta.segments / sta.paths, 3),
                "threshold": 0.02, "min_samples": 16}
    f_de = flops_table()["workloads"]["3"]["march"]["flops_per_de"]
    tflops = f_de * st.de_evals / (st.ms_total / 1e3) / 1e12 if st.ms_total > 0 else 0.0
    return {"tflops": round(tflops, 3), "frac_fp32_peak": round(tflops / FP32_PEAK_TFLOPS, 4),
            "flops_per_de": round(f_de, 2),"spp": spp, "width": W, "height": H, "ms_per_frame": round(wall * 1e3, 2),
            "device_ms": round(st.ms_total, 2), "samples_per_s": round(st.paths / wall, 1),
            "msamples_per_s": round(st.paths / wall / 1e6, 3), "segments_per_s": round(st.segments / wall, 1),
            "segments_per_path": round(st.segments / st.paths, 3), "shadow_rays": int(st.shadow_rays),
            "de_evals_per_s": round(st.de_evals / wall, 1), "de_evals_per_path": round(st.de_evals / st.paths, 1),
            "ms_per_1spp_call": per_sample, "primary_cache": cached, "adaptive": adaptive}


def dream_leg(r, nd, spp=64, reps=2):
    """The reference's benchmark scenes (bench/wavefront_opaque.dream, bench/wavefront_glass.dream,
    kept under tests/golden/scenes) loaded by nd_scene_from_dream at their 960x540 export size and
    path traced at spp samples per pixel, best of reps, as the reference's -bench harness renders
    them (Monte Carlo on, the export's empty-space accelerator at res 64, 6 levels, safety 0.92).
    Not rendered here: the photon map an export traces first, and the glass scene's fog emitter
    (this library's path trace has neither). Also without the accelerator."""
    import torch
    here = os.path.dirname(os.path.abspath(__file__))
    out = {}
    for name in ("opaque", "glass"):
        u, info = nd.load_dream(os.path.join(here, "tests", "golden", "scenes", "wavefront_%s.dream" % name))
        r.bind_formulas(list(info.code_to_builtin[:info.n_codes]))
        u.mc_enabled = 1.0
        u.primary_cache_key = 1.0
        had_fog = u.fog_count > 0
        u.fog_count = 0.0
        W, H = int(u.resolution[0]), int(u.resolution[1])
        col = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda:0")
        res = {"width": W, "height": H, "spp": spp, "fog_removed": bool(had_fog),
               "photon_map": False}
        for label, accel in (("accel", True), ("no_accel", False)):
            v = nd.Uniforms.from_buffer_copy(u)
            if accel:
                nd.enable_accel(v, res=64, levels=6, safety=0.92)
            r.path_trace(v, 1, color=col)  # warmup
            r.sync()
            best, st = None, None
            for _ in range(reps):
                t0 = time.perf_counter()
                _, _, _, s_ = r.path_trace(v, spp, color=col)
                r.sync()
                wall = time.perf_counter() - t0
                if best is None or wall < best:
                    best, st = wall, s_
            log("scene %s (%s): %.1f ms for %d spp, %.2f ms per sample" % (name, label, best * 1e3, spp, best * 1e3 / spp))
            res[label] = {"ms_per_frame": round(best * 1e3, 2), "ms_per_sample": round(best * 1e3 / spp, 3),
                          "samples_per_s": round(st.paths / best, 1), "msamples_per_s": round(st.paths / best / 1e6, 3),
                          "segments_per_path": round(st.segments / st.paths, 3),
                          "de_evals_per_path": round(st.de_evals / st.paths, 1)}
        out[name] = res
    r.bind_formulas([0, 1, 2, 3])
    return out


def cpu_baseline(nd, target_s=40.0):
    """FP32 CPU restatement of the same per-ray math (the oracle instantiated for float),
    on host threads, over a band of rows of camera 0 at full 1920x1080."""
    import oracle
    # every core this process may run on, up to the host share of one GPU (16 threads)
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = max(1, min(16, usable))
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    u = nd.config_uniforms(2, 0)
    W, H = int(u.resolution[0]), int(u.resolution[1])
    rows = 4
    t0 = time.perf_counter()
    oracle.render_gbuffer(u, rect=(0, H // 2, W, rows), prec=1, normal_mode=0, shadows=True,
                          shadow_max_steps=128, threads=threads)
    dt = max(time.perf_counter() - t0, 1e-3)
    rows = int(max(4, min(H, rows * target_s / dt)))
    y0 = max(0, H // 2 - rows // 2)
    t0 = time.perf_counter()
    oracle.render_gbuffer(u, rect=(0, y0, W, rows), prec=1, normal_mode=0, shadows=True,
                          shadow_max_steps=128, threads=threads)
    dt = time.perf_counter() - t0
    return {"value": W * rows / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(), "usable_cpus": usable, "cpu_model": model,
            "sample": "camera 0 of the batch, rows %d..%d of 1920x1080 (%d rays), FP32 CPU restatement of "
                      "the same march + analytic normal + trap + 128-step shadow, %.1f s" % (y0, y0 + rows, W * rows, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cams", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--autotune", action="store_true")
    ap.add_argument("--no-legs", action="store_true", help="skip the path-trace, 256-job and gradient legs")
    ap.add_argument("--dream-spp", type=int, default=64, help="samples per pixel of the scene-file leg")
    ap.add_argument("--legs", default="c1,c3,c4,c5,accel,mesh,dream", help="comma list of extra legs: c1 Mandelbulb 640x360 "
                                                                    "G-buffer, c3 path trace, c4 256 "
                                                                 "jobs, c5 gradient, accel the batch with the "
                                                                 "empty-space accelerator, mesh the mesher's "
                                                                 "512^3 distance pass, dream the reference's two "
                                                                 "benchmark scene files")
    ap.add_argument("--c3-spp", type=int, default=64, help="samples per pixel of the path-trace leg's frames "
                                                          "(64 is the workload; fewer for profiling runs)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="before the timed steps, run one elementwise copy of MIB MiB (reads MIB, writes MIB): "
                         "the reference kernel for scaling hardware byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this path runs on one GPU; --gpus %d is not supported (use --gpus 1)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import torch
    import numericdream_amd as nd
    from numericdream_amd import abi

    r = nd.Renderer(0)
    if args.copy_calibration > 0:
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        torch.mul(a, 1.0, out=b)
        torch.cuda.synchronize()
        log("copy calibration: %d MiB read, %d MiB written" % (args.copy_calibration, args.copy_calibration))
        del a, b
    jobs = make_jobs(nd, torch, args.cams)
    log("jobs ready: %d x 1920x1080" % len(jobs))
    if args.autotune:
        t = r.autotune(jobs, shadows=True, shadow_max_steps=128)
        log("autotune: waves/CU %d, refill_min_idle %d, chunk %d" % (t.waves_per_cu, t.refill_min_idle,
                                                                     t.chunk_rays))
    pers = timed(r, jobs, abi.KERNEL_PERSISTENT, args.steps, args.warmup, torch, abi, "persistent")
    plain = None
    if not args.no_plain:
        plain = timed(r, jobs, abi.KERNEL_PLAIN, args.steps, args.warmup, torch, abi, "plain")

    fl = flops_table()["workloads"]["2"]
    f_march = fl["march"]["flops_per_de"]
    steps = args.steps
    rays = pers["rays"]
    value = rays / pers["wall_s"] / 1e6
    ms_step = pers["wall_s"] * 1e3 / steps
    # roofline of the dominant kernel (the persistent march pass), timed with HIP events on
    # the library's stream
    march_s = pers["ms_march"] / 1e3
    achieved = f_march * pers["de_march"] / march_s / 1e12 if march_s > 0 else 0.0
    res = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mrays/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded cameras, analytic fractal scene)",
        "config": {"workload": "c2: Mandelbox (-1.5,.5,1,1,15 it) G-buffer + 128-step soft shadow, %d cameras x "
                               "1920x1080, persistent batched" % len(jobs),
                   "cameras": len(jobs), "width": 1920, "height": 1080, "shadow_max_steps": 128},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                     "traffic": K_MARCH_TRAFFIC_BYTES_16CAM if len(jobs) == 16 else None,
                     "counters": K_MARCH_COUNTERS_16CAM if len(jobs) == 16 else None,
                     "kernel": "k_march<true, false, false> (persistent march pass: shared scene, no accelerator, tuned formulas)",
                     "flops_per_de": round(f_march, 2), "de_evals_per_launch": pers["de_march"] // steps,
                     "ms_per_launch": round(pers["ms_march"] / steps, 3),
                     "lane_occupancy": round(pers["de_march"] / (64.0 * pers["rounds"]), 4) if pers["rounds"] else None},
        "de_evals_per_s": round(pers["de_total"] / (pers["ms_total"] / 1e3), 1),
        "hits_per_step": pers["hits"] // steps,
    }
    if plain:
        pv = plain["rays"] / plain["wall_s"] / 1e6
        res["plain_mrays_per_s"] = round(pv, 3)
        res["persistent_vs_plain"] = round(value / pv, 3) if pv > 0 else None
        res["plain_ms_per_step"] = round(plain["wall_s"] * 1e3 / steps, 3)
    legs = set() if args.no_legs else set(x.strip() for x in args.legs.split(",") if x.strip())
    if "c1" in legs:
        res["c1_mandelbulb"] = c1_leg(r, nd, torch, abi, flops_table())
    if "c4" in legs:
        res["c4_batch"] = batch_jobs_leg(r, nd, torch, abi)
    if "c5" in legs:
        g = gradient_leg(r, nd, torch)
        res["grad_rays_per_s"] = g["grad_rays_per_s"]
        res["c5_gradient"] = g
    if "c3" in legs:
        res["c3_path_trace"] = mc_leg(r, nd, spp=args.c3_spp)
    if "accel" in legs:
        res["c2_accelerator"] = accel_leg(r, nd, torch, abi, jobs)
    if "dream" in legs:
        res["dream_scenes"] = dream_leg(r, nd, spp=args.dream_spp)
    if "mesh" in legs:
        res["mesh_points"] = mesh_leg(r, nd, torch, flops_table())
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        res["cpu_baseline"] = cpu_baseline(nd)
    print(json.dumps(res), flush=True)
    r.close()


if __name__ == "__main__":
    main()
