// nd_march.hip - geometry passes for gfx950: the primary march with on-hit bisection,
// the G-buffer surface pass (normal + orbit trap) and the light-0 soft shadow.
//
// Two schedules of the same per-ray math:
//   plain       one launch per frame, one lane per pixel, one wave64 per 8x8 tile; a wave
//               runs until its slowest ray is done (the reference's fragment dispatch,
//               fs_simple, template.wgsl:4521-4538).
//   persistent  three persistent passes over the whole batch of jobs: march (rays pulled
//               from a device queue, a finished lane is refilled with the next ray at once,
//               hits appended to a compacted hit list with one atomic per wave), surface
//               (normal + trap of every hit), shadow (refilled like the march).
// Reference semantics: march_to (template.wgsl:1859-1963), estimate_normal
// (:1781-1804), visibility_march / shadow_factor (:2282-2375), frag_out (:3762-3774).
#include <hip/hip_runtime.h>

#include "nd_accel.hpp"
#include "nd_de.hpp"
#include "nd_internal.h"
#include "nd_march_state.hpp"
#include "nd_rays.hpp"

namespace nd {

// ----------------------------------------------------------------------------- surface
NDD V3<float> vnorm(const V3<float>& g) {
    float il = __frsqrt_rn(dot3(g, g));
    return mk3(g.x * il, g.y * il, g.z * il);
}
template <bool EXT>
NDD V3<float> surface_normal(const DevScene& sc, int mode, const V3<float>& p, float eps, int* evals) {
    if (mode == 1) {  // tetrahedral, e = eps(t)
        const float K[4][3] = {{1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}, {1, 1, 1}};
        V3<float> g = mk3(0.f, 0.f, 0.f);
        for (int i = 0; i < 4; ++i) {
            V3<float> k = mk3(K[i][0], K[i][1], K[i][2]);
            float d = scene_de<false, EXT>(sc, mk3(p.x + k.x * eps, p.y + k.y * eps, p.z + k.z * eps), nullptr);
            g = mk3(g.x + k.x * d, g.y + k.y * d, g.z + k.z * d);
        }
        *evals += 4;
        return vnorm(g);
    }
    if (mode == 2) {  // central, h = eps / 2
        float h = 0.5f * eps;
        float gx = scene_de<false, EXT>(sc, mk3(p.x + h, p.y, p.z), nullptr) - scene_de<false, EXT>(sc, mk3(p.x - h, p.y, p.z), nullptr);
        float gy = scene_de<false, EXT>(sc, mk3(p.x, p.y + h, p.z), nullptr) - scene_de<false, EXT>(sc, mk3(p.x, p.y - h, p.z), nullptr);
        float gz = scene_de<false, EXT>(sc, mk3(p.x, p.y, p.z + h), nullptr) - scene_de<false, EXT>(sc, mk3(p.x, p.y, p.z - h), nullptr);
        *evals += 6;
        return vnorm(mk3(gx, gy, gz));
    }
    // analytic: forward-mode derivative of the DE
    using D3 = DualN<3>;
    V3<D3> q = mk3(dvar<3>(p.x, 0), dvar<3>(p.y, 1), dvar<3>(p.z, 2));
    D3 d = scene_de_d<D3, EXT, true>(sc, q);
    *evals += 1;
    return vnorm(mk3(d.d[0], d.d[1], d.d[2]));
}

// Light 0 direction and distance at a surface point (light_sample, template.wgsl:2178-2190):
// type 1 global (v = normalised direction, distance far), else point (v = position).
NDD void light_dir(int type, const V3<float>& v, float far, const V3<float>& p, V3<float>* dir, float* dist) {
    if (type == 1) {
        *dir = v;
        *dist = far;
    } else {
        V3<float> to = mk3(v.x - p.x, v.y - p.y, v.z - p.z);
        float dd = len3(to);
        float inv = frcp(fmaxf(dd, 1e-4f));
        *dir = mk3(to.x * inv, to.y * inv, to.z * inv);
        *dist = dd;
    }
}
NDD void light_at(const DevJob& J, const V3<float>& p, V3<float>* dir, float* dist) {
    light_dir(J.light_type, mk3(J.light_v[0], J.light_v[1], J.light_v[2]), J.light_far, p, dir, dist);
}

// With SHARED every job of the batch has the same scene: it is read through a
// kernel-argument pointer, so its fields are wave-uniform scalar loads.
template <bool SHARED>
NDD const DevScene& scene_of(const DevJob* __restrict__ jobs, const DevScene* __restrict__ shared, int job) {
    if (SHARED) return *shared;
    return jobs[job].scene;
}

// The whole per-pixel pipeline of the plain schedule; returns its DE evaluations.
template <bool EXT>
NDD int plain_pixel(const DevJob& J, const DevScene& sc, int x, int y) {
    const V3<float> dir = primary_dir(J, J.rect_x + x, J.rect_y + y);
    const V3<float> org = mk3(J.cam[0], J.cam[1], J.cam[2]);
    March m;
    march_init(m);
    float dist = 0.f;
    int st;
    const bool acc = J.accel != nullptr;
    const AccelView av = accel_view(J);
    AccelCache ac;
    accel_cache_reset(ac);
    do {
        // march_to's loop: at a probe the accelerator's skip, when it applies, takes the step
        // without a DE (template.wgsl:1890-1900); otherwise the DE decides
        float s = march_sample_t(m);
        const V3<float> ps = mk3(org.x + dir.x * s, org.y + dir.y * s, org.z + dir.z * s);
        st = (acc && m.probing) ? march_skip_step(m, J, accel_skip_cached(av, ac, ps), &dist) : -1;
        if (st < 0) st = march_update(m, J, s, scene_de<false, EXT>(sc, ps, nullptr), &dist);
    } while (st == MS_CONT);
    const size_t o = (size_t)y * J.rect_w + x;
    int evals = m.evals;
    if (st != MS_HIT) {
        if (J.out_depth) J.out_depth[o] = 0.f;
        if (J.out_normal) { J.out_normal[3 * o] = 0.f; J.out_normal[3 * o + 1] = 0.f; J.out_normal[3 * o + 2] = 0.f; }
        if (J.out_trap) J.out_trap[o] = 0.f;
        if (J.out_hit) J.out_hit[o] = 0;
        if (J.out_shadow) J.out_shadow[o] = 1.f;
        if (J.out_evals) J.out_evals[o] = (uint32_t)evals;
        if (J.out_pos) { J.out_pos[3 * o] = 0.f; J.out_pos[3 * o + 1] = 0.f; J.out_pos[3 * o + 2] = 0.f; }
        return evals;
    }
    const V3<float> hp = mk3(org.x + dir.x * dist, org.y + dir.y * dist, org.z + dir.z * dist);
    int sev = 0;
    const V3<float> n = surface_normal<EXT>(sc, J.normal_mode, hp, J.eps_c * dist + J.eps_f, &sev);
    float trap = 0.f;
    (void)scene_de<true, EXT>(sc, hp, &trap);
    float shadow = 1.f;
    if (J.shadow_on) {
        V3<float> ld; float ldist;
        light_at(J, hp, &ld, &ldist);
        if (dot3(n, ld) <= 0.f) {
            shadow = 0.f;
        } else {
            const V3<float> so = mk3(hp.x + n.x * 0.004f, hp.y + n.y * 0.004f, hp.z + n.z * 0.004f);
            float t = 0.004f, res = 1.f;
            accel_cache_reset(ac);
            for (int i = 0; i < J.shadow_steps; ++i) {
                if (t >= ldist) break;
                const V3<float> ps = mk3(so.x + ld.x * t, so.y + ld.y * t, so.z + ld.z * t);
                if (acc) {
                    const float skip = accel_skip_cached(av, ac, ps);
                    if (shadow_skip_applies(skip, J.light_k, res, t)) { t += skip; continue; }
                }
                const float d = scene_de<false, EXT>(sc, ps, nullptr);
                evals++;
                if (d < 0.0005f) { res = 0.f; break; }
                res = fminf(res, J.light_k * d * frcp(t));
                t += fmaxf(d, 0.001f);
            }
            shadow = fclamp(res, 0.f, 1.f);
        }
    }
    if (J.out_depth) J.out_depth[o] = dist;
    if (J.out_normal) { J.out_normal[3 * o] = n.x; J.out_normal[3 * o + 1] = n.y; J.out_normal[3 * o + 2] = n.z; }
    if (J.out_trap) J.out_trap[o] = trap;
    if (J.out_hit) J.out_hit[o] = 1;
    if (J.out_shadow) J.out_shadow[o] = shadow;
    if (J.out_evals) J.out_evals[o] = (uint32_t)evals;
    if (J.out_pos) { J.out_pos[3 * o] = hp.x; J.out_pos[3 * o + 1] = hp.y; J.out_pos[3 * o + 2] = hp.z; }
    return evals + sev + 1;
}

// ============================================================================ plain kernel
// One lane per pixel of one job; blockDim 64 = one tile (8x8 by default) per wave64. 6 waves per SIMD
// (80 VGPRs) is this kernel's fastest setting, about 7% ahead of the compiler's 4.
template <bool EXT>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(EXT ? 1 : 6, EXT ? 8 : 6))) k_plain(const DevJob* __restrict__ jobp, DevCounters* ctr) {
    const DevJob& J = *jobp;
    const DevScene& sc = J.scene;
    const int x = (blockIdx.x << J.tile_lg) + (int)(threadIdx.x & ((1u << J.tile_lg) - 1u));
    const int y = blockIdx.y * (64 >> J.tile_lg) + (int)(threadIdx.x >> J.tile_lg);
    const bool live = x < J.rect_w && y < J.rect_h;
    const int ev = live ? plain_pixel<EXT>(J, sc, x, y) : 0;
    unsigned long long v = (unsigned long long)ev;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if (threadIdx.x == 0 && ctr) atomicAdd(&ctr->de_march, v);
}

// ============================================================================ persistent passes
__global__ void k_zero_counters(DevCounters* c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        c->ray_next = 0; c->hit_count = 0; c->surf_next = 0; c->shadow_next = 0;
        c->de_march = 0; c->de_surface = 0; c->de_shadow = 0; c->march_rounds = 0; c->hits = 0;
        c->rounds_capped = 0;
        c->clock_start = wall_clock64();
    }
}

NDD uint64_t lanemask_lt() {
    const unsigned lane = __lane_id();
    return lane ? (~0ull >> (64 - lane)) : 0ull;
}
NDD bool lane_bit(uint64_t m) { return (m >> __lane_id()) & 1ull; }
NDD uint32_t uniform_u32(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }


struct PassArgs {
    const DevJob* jobs;
    const DevScene* shared;  // non-null when every job has the same scene
    const uint32_t* job_base;
    int n_jobs;
    uint32_t total_rays;
    DevCounters* ctr;
    DevHitBuffers hb;
    uint32_t chunk;          // ray slots per claim of the march pass
    uint32_t shadow_chunk;   // hits per claim of the shadow pass
    uint32_t min_idle;
    uint32_t round_cap;
    uint32_t refill_period;  // march steps between refills (a wave with every lane idle refills at once)
    uint32_t* ray_done;  // per ray slot: completion time in wall-clock ticks after clock_start, or null
};

NDD void write_miss(const DevJob& J, size_t o, int evals) {
    if (J.out_depth) J.out_depth[o] = 0.f;
    if (J.out_normal) { J.out_normal[3 * o] = 0.f; J.out_normal[3 * o + 1] = 0.f; J.out_normal[3 * o + 2] = 0.f; }
    if (J.out_pos) { J.out_pos[3 * o] = 0.f; J.out_pos[3 * o + 1] = 0.f; J.out_pos[3 * o + 2] = 0.f; }
    if (J.out_trap) J.out_trap[o] = 0.f;
    if (J.out_hit) J.out_hit[o] = 0;
    if (J.out_shadow) J.out_shadow[o] = 1.f;
    if (J.out_evals) J.out_evals[o] = (uint32_t)evals;
}

// The G-buffer writes of a finished primary ray: depth, hit and evaluations of a hit (the
// surface pass adds the rest), every field of a miss.
NDD void write_march_result(const DevJob& J, int px, int py, int st, float dist, int evals) {
    const size_t o = (size_t)py * J.rect_w + px;
    if (st == MS_HIT) {
        if (J.out_depth) J.out_depth[o] = dist;
        if (J.out_hit) J.out_hit[o] = 1;
        if (J.out_evals) J.out_evals[o] = (uint32_t)evals;
    } else {
        write_miss(J, o, evals);
    }
}

// Runs f(J) for the job of each active lane. The lanes of a wave nearly always hold rays of
// one job (a claim is 96 or more consecutive slots of a job), and then the job index is
// made wave-uniform: the record's fields and output pointers become scalar loads, instead
// of a chain of per-lane pointer loads each waited for before its store.
template <class F>
NDD void for_lane_job(const DevJob* __restrict__ jobs, int job, F&& f) {
    const int j0 = __builtin_amdgcn_readfirstlane(job);
    if (__builtin_amdgcn_ballot_w64(job != j0) == 0ull) f(jobs[j0]);
    else f(jobs[job]);
}

// --- pass 1: march. Every round each active lane takes one DE sample; a lane whose ray
// finished is refilled at once from the wave's current chunk of ray slots. The chunk and
// its job are wave-uniform, so a refill reads the job's camera and precision with scalar
// loads and the hot loop touches no per-lane memory.
// Left to itself the compiler gives this kernel 84 VGPRs (93 without a shared scene), 5 waves
// per SIMD. Forcing 7 waves with a shared scene (72 VGPRs, three spilled values) makes the
// pass about 7% faster, and 6 waves without one about 4%: more waves hide the DE's dependent
// latency better than the spills cost. The accelerated variant (about 100 VGPRs, 4 waves left
// to itself) is fastest at 6: 35.0 ms per 16-camera launch against 40.3 at 4 and 43.9 at 7.
template <bool SHARED, bool ACC, bool EXT>
__global__ void __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(EXT ? 1 : 6 + (SHARED && !ACC), EXT ? 8 : 6 + (SHARED && !ACC))))
k_march(PassArgs a, const DevJob* __restrict__ jobs,
                                                 const DevScene* __restrict__ shared) {
    uint32_t cnext = 0, cend = 0;          // wave-uniform chunk of ray slots [cnext, cend)
    int cjob = 0;                          // wave-uniform job of the chunk position
    uint32_t cjstart = 1, cjend = 0;       // ray slots of cjob
    uint32_t hnext = 0, hend = 0;          // wave-uniform chunk of hit-list slots
    bool work_left = true, first_claim = true;
    bool active = false;
    uint32_t ray = 0;
    int job = 0, px = 0, py = 0;
    uint32_t lat_ray = ~0u;                // retired ray slot whose completion time is written at the
                                           // lane's next refill (latency tracking)
    const unsigned long long c0 = a.ray_done ? a.ctr->clock_start : 0ull;
    V3<float> org = mk3(0.f, 0.f, 0.f), dir = mk3(0.f, 0.f, 1.f);
    March m;
    march_init(m);
    MarchParams mp = {};
    bool acc = false;  // the ray's job uses the accelerator
    AccelView av = {};
    AccelCache ac;
    accel_cache_reset(ac);
    unsigned long long de = 0, nhits = 0;
    uint32_t rounds_used = 0;  // rounds that evaluated at least one DE
    for (uint32_t round = 0;; ++round) {
        if (work_left) {
            uint64_t need = __builtin_amdgcn_ballot_w64(!active);
            if (((uint32_t)__popcll(need) >= a.min_idle && round % a.refill_period == 0u) || need == ~0ull) {
                // latency tracking reads the clock once per refill: a ray that retired since the
                // last refill finished no later than now
                if (a.ray_done && lat_ray != ~0u) {
                    a.ray_done[lat_ray] = (uint32_t)((unsigned long long)wall_clock64() - c0);
                    lat_ray = ~0u;
                }
                while (need != 0ull) {
                    if (cnext >= cend) {
                        const uint32_t base = claim_chunk(&a.ctr->ray_next, a.total_rays, a.chunk, first_claim);
                        first_claim = false;
                        if (base >= a.total_rays) { work_left = false; break; }
                        cnext = base;
                        cend = uniform_u32(min(base + a.chunk, a.total_rays));
                    }
                    if (cnext < cjstart || cnext >= cjend) {
                        cjob = (int)uniform_u32((uint32_t)find_job(a.job_base, a.n_jobs, cnext));
                        cjstart = jobs[cjob].ray_base;
                        cjend = cjstart + jobs[cjob].ray_count;
                    }
                    const DevJob& CJ = jobs[cjob];
                    const uint32_t avail = min(cend, cjend) - cnext;
                    const uint32_t rank = (uint32_t)__popcll(need & lanemask_lt());
                    const bool mine = lane_bit(need) && rank < avail;
                    if (mine) {
                        ray = cnext + rank;
                        if (slot_pixel(CJ, ray - cjstart, &px, &py)) {
                            job = cjob;
                            dir = primary_dir(CJ, CJ.rect_x + px, CJ.rect_y + py);
                            org = mk3(CJ.cam[0], CJ.cam[1], CJ.cam[2]);
                            mp = march_params(CJ);
                            acc = ACC && CJ.accel != nullptr;
                            if (acc) av = accel_view(CJ);
                            if (ACC) accel_cache_reset(ac);
                            march_init(m);
                            active = true;
                        }
                    }
                    cnext += min((uint32_t)__popcll(need), avail);
                    need = __builtin_amdgcn_ballot_w64(lane_bit(need) && !mine);
                }
            }
        }
        const uint64_t act = __builtin_amdgcn_ballot_w64(active);
        if (act == 0ull) {
            if (!work_left) break;
            continue;
        }
        ++rounds_used;
        bool hit_now = false;
        float dist = 0.f;
        int st = -1;
        bool used = false;  // the lane evaluated a DE this round (no accelerator skip applied)
        if (ACC || EXT || !SHARED) {
            if (active) {
                // march_to's loop: at a probe the accelerator's skip, when it applies, takes the
                // step without a DE (template.wgsl:1890-1900); a wave whose lanes all skip runs
                // no DE
                const DevScene& sc = scene_of<SHARED>(jobs, shared, job);
                const float s = march_sample_t(m);
                const V3<float> ps = mk3(org.x + dir.x * s, org.y + dir.y * s, org.z + dir.z * s);
                if (ACC && acc && m.probing) st = march_skip_step(m, mp, accel_skip_cached(av, ac, ps), &dist);
                used = st < 0;
                if (used) st = march_update(m, mp, s, scene_de<false, EXT>(sc, ps, nullptr), &dist);
            }
        } else {
            // Idle lanes step too, on the ray they last held (or the initial one): the wave
            // issues the same instructions either way, and without a branch around the step
            // the ray state is updated in place instead of copied in and out of the branch
            // every round. An idle lane's result is discarded, and its state is initialised
            // again when it takes a ray. The tuned DEs read no memory by position. Only with
            // one scene for every job: with several, an idle lane's stale job could hold
            // another formula than the active lanes', a divergent DE of its own.
            const DevScene& sc = scene_of<SHARED>(jobs, shared, job);
            const float s = march_sample_t(m);
            const V3<float> ps = mk3(org.x + dir.x * s, org.y + dir.y * s, org.z + dir.z * s);
            const int st_any = march_update(m, mp, s, scene_de<false, EXT>(sc, ps, nullptr), &dist);
            if (active) { st = st_any; used = true; }
        }
        de += (unsigned long long)__popcll(__builtin_amdgcn_ballot_w64(used));
        if (active) {
            if (st != MS_CONT) {
                hit_now = st == MS_HIT;
                for_lane_job(jobs, job, [&](const DevJob& J) { write_march_result(J, px, py, st, dist, m.evals); });
                active = false;
                if (a.ray_done) {
                    // timed at this lane's next refill; after the queue drained there is none,
                    // so the clock is read now (the tail of the pass)
                    if (work_left) lat_ray = ray;
                    else a.ray_done[ray] = (uint32_t)((unsigned long long)wall_clock64() - c0);
                }
            }
        }
        // append this round's hits to the compacted hit list through the wave's slot chunk
        const uint64_t hm = __builtin_amdgcn_ballot_w64(hit_now);
        if (hm) {
            const uint32_t nh = (uint32_t)__popcll(hm);
            nhits += nh;
            const uint32_t room = hend - hnext;
            uint32_t nbase = 0;
            if (nh > room) {
                if (__lane_id() == 0) nbase = atomicAdd(&a.ctr->hit_count, 64u);
                nbase = uniform_u32(nbase);
            }
            if (hit_now) {
                const uint32_t rank = (uint32_t)__popcll(hm & lanemask_lt());
                const uint32_t idx = rank < room ? hnext + rank : nbase + (rank - room);
                a.hb.ray[idx] = ray;
                a.hb.job[idx] = (uint16_t)job;
                a.hb.t[idx] = dist;
            }
            if (nh > room) { hnext = nbase + (nh - room); hend = nbase + 64u; }
            else hnext += nh;
        }
        if (round >= a.round_cap) {
            if (__lane_id() == 0) atomicAdd(&a.ctr->rounds_capped, 1u);
            break;
        }
    }
    // unused slots of the last hit chunk are marked empty for the later passes
    for (uint32_t i = hnext + __lane_id(); i < hend; i += 64) a.hb.ray[i] = ~0u;
    if (a.ray_done && lat_ray != ~0u) a.ray_done[lat_ray] = (uint32_t)((unsigned long long)wall_clock64() - c0);
    if (__lane_id() == 0 && rounds_used) {  // a wave that marched nothing adds nothing
        atomicAdd(&a.ctr->march_rounds, (unsigned long long)rounds_used);
        atomicAdd(&a.ctr->de_march, de);
        if (nhits) atomicAdd(&a.ctr->hits, nhits);
    }
}

// --- pass 2: surface. Normal + trap of every hit (fixed cost per hit), and the shadow
// ray of the hit, stored SoA for pass 3. 4 waves per SIMD: with the Mandelbox normal's
// parameters as plain floats, 4 waves are 2% faster than 5 on the 16-camera batch (5 were
// 3% ahead of 4 before that change).
template <bool SHARED, bool EXT>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(EXT ? 1 : 4, EXT ? 8 : 4)))
k_surface(PassArgs a, const DevJob* __restrict__ jobs,
                                                 const DevScene* __restrict__ shared) {
    const uint32_t n_hits = a.ctr->hit_count;
    unsigned long long de = 0;
    for (uint32_t i = blockIdx.x * 64 + threadIdx.x; i - threadIdx.x < n_hits; i += gridDim.x * 64) {
        const uint32_t r = i < n_hits ? a.hb.ray[i] : ~0u;
        if (i < n_hits && r == ~0u) a.hb.steps[i] = 0;  // empty slot of a partly used chunk
        if (r != ~0u) {
            const int job = a.hb.job[i];
            const DevScene& sc = scene_of<SHARED>(jobs, shared, job);
            // the hit's pixel, point, epsilon and light-0 terms from its job record
            const float t = a.hb.t[i];
            size_t o = 0;
            V3<float> hp;
            float eps = 0.f, lfar = 0.f;
            V3<float> lv;
            int nmode = 0, light_type = 0, shadow_on = 0, shadow_steps = 0;
            for_lane_job(jobs, job, [&](const DevJob& J) {
                int x, y;
                slot_pixel(J, r - J.ray_base, &x, &y);
                o = (size_t)y * J.rect_w + x;
                const V3<float> dir = primary_dir(J, J.rect_x + x, J.rect_y + y);
                hp = mk3(J.cam[0] + dir.x * t, J.cam[1] + dir.y * t, J.cam[2] + dir.z * t);
                eps = J.eps_c * t + J.eps_f;
                nmode = J.normal_mode;
                shadow_on = J.shadow_on;
                shadow_steps = J.shadow_steps;
                light_type = J.light_type;
                lv = mk3(J.light_v[0], J.light_v[1], J.light_v[2]);
                lfar = J.light_far;
            });
            int ev = 0;
            const V3<float> n = surface_normal<EXT>(sc, nmode, hp, eps, &ev);
            float trap = 0.f;
            (void)scene_de<true, EXT>(sc, hp, &trap);
            ev += 1;
            de += (unsigned long long)ev;
            // shadow ray of light 0 (shadow_factor, template.wgsl:2365-2375); steps = 0 marks
            // a hit that needs no march (shadows off, or the light is behind the surface)
            int steps = 0;
            float shadow_now = 1.f;
            V3<float> ld = mk3(0.f, 0.f, 0.f);
            float ldist = 0.f;
            if (shadow_on) {
                light_dir(light_type, lv, lfar, hp, &ld, &ldist);
                if (dot3(n, ld) <= 0.f) shadow_now = 0.f;
                else if (shadow_steps > 0 && 0.004f < ldist) steps = shadow_steps;
            }
            for_lane_job(jobs, job, [&](const DevJob& JO) {
                if (JO.out_normal) { JO.out_normal[3 * o] = n.x; JO.out_normal[3 * o + 1] = n.y; JO.out_normal[3 * o + 2] = n.z; }
                if (JO.out_pos) { JO.out_pos[3 * o] = hp.x; JO.out_pos[3 * o + 1] = hp.y; JO.out_pos[3 * o + 2] = hp.z; }
                if (JO.out_trap) JO.out_trap[o] = trap;
                if (steps == 0 && JO.out_shadow) JO.out_shadow[o] = shadow_now;
            });
            a.hb.opix[i] = (uint32_t)o;
            a.hb.steps[i] = (uint16_t)steps;
            a.hb.sox[i] = hp.x + n.x * 0.004f;
            a.hb.soy[i] = hp.y + n.y * 0.004f;
            a.hb.soz[i] = hp.z + n.z * 0.004f;
            a.hb.sdx[i] = ld.x;
            a.hb.sdy[i] = ld.y;
            a.hb.sdz[i] = ld.z;
            a.hb.sdist[i] = ldist;
        }
    }
    for (int off = 32; off > 0; off >>= 1) de += __shfl_down(de, off);
    if (threadIdx.x == 0) atomicAdd(&a.ctr->de_surface, de);
}

// --- pass 3: light-0 soft shadow of every hit (visibility_march on an opaque scene,
// template.wgsl:2282-2342), refilled like the march from the SoA shadow rays. Its rays are
// short (7 DEs per hit on the 16-camera batch), so refills and their loads are frequent:
// 7 waves per SIMD (the compiler alone gives 5) hide more of that latency, 7% faster.
template <bool SHARED, bool ACC, bool EXT>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(EXT || ACC ? 1 : 7, EXT || ACC ? 8 : 7)))
k_shadow(PassArgs a, const DevJob* __restrict__ jobs,
                                                 const DevScene* __restrict__ shared) {
    const uint32_t n_hits = a.ctr->hit_count;
    uint32_t cnext = 0, cend = 0;
    bool work_left = true, active = false, first_claim = true;
    uint32_t hit = 0;
    int job = 0, steps = 0, step = 0, sev = 0;  // step: march steps incl. accelerator skips; sev: DEs
    bool sacc = false;
    AccelView sav = {};
    AccelCache sac;
    accel_cache_reset(sac);
    V3<float> so = mk3(0.f, 0.f, 0.f), ld = mk3(0.f, 0.f, 1.f);
    float t = 0.f, res = 1.f, ldist = 0.f, k = 0.f;
    uint32_t sray = 0;           // ray slot of the current shadow ray
    uint32_t lat_ray = ~0u;      // retired ray slot timed at the next refill
    const unsigned long long c0 = a.ray_done ? a.ctr->clock_start : 0ull;
    unsigned long long de = 0;
    for (uint32_t round = 0;; ++round) {
        if (work_left) {
            uint64_t need = __builtin_amdgcn_ballot_w64(!active);
            if (((uint32_t)__popcll(need) >= a.min_idle && round % a.refill_period == 0u) || need == ~0ull) {
                if (a.ray_done && lat_ray != ~0u) {
                    a.ray_done[lat_ray] = (uint32_t)((unsigned long long)wall_clock64() - c0);
                    lat_ray = ~0u;
                }
                while (need != 0ull) {
                    if (cnext >= cend) {
                        const uint32_t base = claim_chunk(&a.ctr->shadow_next, n_hits, a.shadow_chunk, first_claim);
                        first_claim = false;
                        if (base >= n_hits) { work_left = false; break; }
                        cnext = base;
                        cend = uniform_u32(min(base + a.shadow_chunk, n_hits));
                    }
                    const uint32_t avail = cend - cnext;
                    const uint32_t rank = (uint32_t)__popcll(need & lanemask_lt());
                    const bool mine = lane_bit(need) && rank < avail;
                    if (mine) {
                        hit = cnext + rank;
                        steps = a.hb.steps[hit];
                        if (steps > 0) {
                            job = a.hb.job[hit];
                            if (a.ray_done) sray = a.hb.ray[hit];
                            so = mk3(a.hb.sox[hit], a.hb.soy[hit], a.hb.soz[hit]);
                            ld = mk3(a.hb.sdx[hit], a.hb.sdy[hit], a.hb.sdz[hit]);
                            ldist = a.hb.sdist[hit];
                            for_lane_job(jobs, job, [&](const DevJob& J) {
                                k = J.light_k;
                                sacc = ACC && J.accel != nullptr;
                                if (sacc) sav = accel_view(J);
                            });
                            if (ACC) accel_cache_reset(sac);
                            t = 0.004f; res = 1.f; step = 0; sev = 0;
                            active = true;
                        }
                    }
                    cnext += min((uint32_t)__popcll(need), avail);
                    need = __builtin_amdgcn_ballot_w64(lane_bit(need) && !mine);
                }
            }
        }
        const uint64_t act = __builtin_amdgcn_ballot_w64(active);
        if (act == 0ull) {
            if (!work_left) break;
            continue;
        }
        bool done = false, skipped = false;
        if (active) {
            const DevScene& sc = scene_of<SHARED>(jobs, shared, job);
            const V3<float> ps = mk3(so.x + ld.x * t, so.y + ld.y * t, so.z + ld.z * t);
            step++;
            float skip = 0.f;
            if (ACC && sacc) {
                skip = accel_skip_cached(sav, sac, ps);
                skipped = shadow_skip_applies(skip, k, res, t);
            }
            if (skipped) {
                t += skip;
                if (step >= steps || t >= ldist) done = true;
            } else {
                const float d = scene_de<false, EXT>(sc, ps, nullptr);
                sev++;
                if (d < 0.0005f) { res = 0.f; done = true; }
                else {
                    res = fminf(res, k * d * frcp(t));
                    t += fmaxf(d, 0.001f);
                    if (step >= steps || t >= ldist) done = true;
                }
            }
            if (done) {
                const uint32_t o = a.hb.opix[hit];
                for_lane_job(jobs, job, [&](const DevJob& J) {
                    if (J.out_shadow) J.out_shadow[o] = fclamp(res, 0.f, 1.f);
                    if (J.out_evals) J.out_evals[o] += (uint32_t)sev;
                });
                active = false;
                if (a.ray_done) {
                    if (work_left) lat_ray = sray;
                    else a.ray_done[sray] = (uint32_t)((unsigned long long)wall_clock64() - c0);
                }
            }
        }
        de += (unsigned long long)__popcll(act & ~__builtin_amdgcn_ballot_w64(skipped));
        if (round >= a.round_cap) {
            if (__lane_id() == 0) atomicAdd(&a.ctr->rounds_capped, 1u);
            break;
        }
    }
    if (a.ray_done && lat_ray != ~0u) a.ray_done[lat_ray] = (uint32_t)((unsigned long long)wall_clock64() - c0);
    if (__lane_id() == 0 && de) atomicAdd(&a.ctr->de_shadow, de);
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

// Per-job completion clock = clock_start + the latest completion tick over the job's ray
// slots (slots outside the rect keep 0). Grid (jobs, JOB_DONE_SLICES): each 256-thread
// workgroup reduces one slice of a job and merges it with one atomicMax into the zeroed
// job_done.
constexpr int JOB_DONE_SLICES = 16;
__global__ void __launch_bounds__(256) k_job_done(const DevJob* __restrict__ jobs, const uint32_t* __restrict__ ray_done,
                                                  const DevCounters* ctr, unsigned long long* job_done) {
    const DevJob& J = jobs[blockIdx.x];
    uint32_t m = 0;
    for (uint32_t i = blockIdx.y * 256 + threadIdx.x; i < J.ray_count; i += 256 * JOB_DONE_SLICES)
        m = max(m, ray_done[J.ray_base + i]);
    for (int off = 32; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_down((int)m, off));
    __shared__ uint32_t part[4];
    if (__lane_id() == 0) part[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = max(max(part[0], part[1]), max(part[2], part[3]));
        if (m) atomicMax(&job_done[blockIdx.x], ctr->clock_start + m);
    }
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0ull;
}

}  // namespace nd

// ----------------------------------------------------------------------------- launchers
hipError_t nd_launch_zero_counters(DevCounters* c, hipStream_t s) {
    hipLaunchKernelGGL(nd::k_zero_counters, dim3(1), dim3(64), 0, s, c);
    return hipGetLastError();
}

hipError_t nd_launch_plain(const DevJob* d_job, int tiles_x, int tiles_y, DevCounters* ctr, bool ext, hipStream_t s) {
    if (ext) hipLaunchKernelGGL(nd::k_plain<true>, dim3(tiles_x, tiles_y), dim3(64), 0, s, d_job, ctr);
    else hipLaunchKernelGGL(nd::k_plain<false>, dim3(tiles_x, tiles_y), dim3(64), 0, s, d_job, ctr);
    return hipGetLastError();
}

hipError_t nd_launch_persistent(const NdPassLaunch& L, hipStream_t s) {
    nd::PassArgs a;
    a.jobs = L.jobs;
    a.shared = L.shared;
    a.job_base = L.job_base;
    a.n_jobs = L.n_jobs;
    a.total_rays = L.total_rays;
    a.ctr = L.ctr;
    a.hb = L.hb;
    a.chunk = L.chunk;
    // Shadow rays are shorter (at most 40 steps) and start in bursts of hits: claims of four
    // times the march's keep the claim counter cooler (measured: 128-ray march claims with
    // 512-hit shadow claims beat 256 for both).
    a.shadow_chunk = 4 * L.chunk;
    a.min_idle = L.min_idle;
    a.round_cap = L.round_cap;
    a.refill_period = L.refill_period ? L.refill_period : 1u;
    a.ray_done = L.job_done_cycles ? L.ray_done : nullptr;
    hipLaunchKernelGGL(nd::k_zero_counters, dim3(1), dim3(64), 0, s, L.ctr);
    if (a.ray_done) {
        hipLaunchKernelGGL(nd::k_zero_u32, dim3(1024), dim3(256), 0, s, a.ray_done, L.total_rays);
        hipLaunchKernelGGL(nd::k_zero_u64, dim3((L.n_jobs + 255) / 256), dim3(256), 0, s, L.job_done_cycles, L.n_jobs);
    }
    if (L.ev[0]) hipEventRecord(L.ev[0], s);
    // ACC: some job uses the accelerator; EXT: some instance uses a formula beyond the four
    // tuned ones (those kernels read each job's own scene). The kernels without either keep
    // fewer registers.
    const dim3 b64(64);
    if (L.ext) {
        if (L.accel) hipLaunchKernelGGL((nd::k_march<false, true, true>), dim3(L.grid_march), b64, 0, s, a, L.jobs, L.shared);
        else hipLaunchKernelGGL((nd::k_march<false, false, true>), dim3(L.grid_march), b64, 0, s, a, L.jobs, L.shared);
    } else if (L.shared) {
        if (L.accel) hipLaunchKernelGGL((nd::k_march<true, true, false>), dim3(L.grid_march), b64, 0, s, a, L.jobs, L.shared);
        else hipLaunchKernelGGL((nd::k_march<true, false, false>), dim3(L.grid_march), b64, 0, s, a, L.jobs, L.shared);
    } else {
        if (L.accel) hipLaunchKernelGGL((nd::k_march<false, true, false>), dim3(L.grid_march), b64, 0, s, a, L.jobs, L.shared);
        else hipLaunchKernelGGL((nd::k_march<false, false, false>), dim3(L.grid_march), b64, 0, s, a, L.jobs, L.shared);
    }
    if (L.ev[1]) hipEventRecord(L.ev[1], s);
    if (L.surface) {
        if (L.ext) hipLaunchKernelGGL((nd::k_surface<false, true>), dim3(L.grid_surface), b64, 0, s, a, L.jobs, L.shared);
        else if (L.shared) hipLaunchKernelGGL((nd::k_surface<true, false>), dim3(L.grid_surface), b64, 0, s, a, L.jobs, L.shared);
        else hipLaunchKernelGGL((nd::k_surface<false, false>), dim3(L.grid_surface), b64, 0, s, a, L.jobs, L.shared);
    }
    if (L.ev[2]) hipEventRecord(L.ev[2], s);
    if (L.any_shadow) {
        if (L.ext) {
            if (L.accel) hipLaunchKernelGGL((nd::k_shadow<false, true, true>), dim3(L.grid_shadow), b64, 0, s, a, L.jobs, L.shared);
            else hipLaunchKernelGGL((nd::k_shadow<false, false, true>), dim3(L.grid_shadow), b64, 0, s, a, L.jobs, L.shared);
        } else if (L.shared) {
            if (L.accel) hipLaunchKernelGGL((nd::k_shadow<true, true, false>), dim3(L.grid_shadow), b64, 0, s, a, L.jobs, L.shared);
            else hipLaunchKernelGGL((nd::k_shadow<true, false, false>), dim3(L.grid_shadow), b64, 0, s, a, L.jobs, L.shared);
        } else {
            if (L.accel) hipLaunchKernelGGL((nd::k_shadow<false, true, false>), dim3(L.grid_shadow), b64, 0, s, a, L.jobs, L.shared);
            else hipLaunchKernelGGL((nd::k_shadow<false, false, false>), dim3(L.grid_shadow), b64, 0, s, a, L.jobs, L.shared);
        }
    }
    if (a.ray_done)
        hipLaunchKernelGGL(nd::k_job_done, dim3(L.n_jobs, nd::JOB_DONE_SLICES), dim3(256), 0, s, L.jobs, a.ray_done,
                           L.ctr, L.job_done_cycles);
    if (L.ev[3]) hipEventRecord(L.ev[3], s);
    return hipGetLastError();
}

template <typename K>
static uint32_t resident_of(K kernel) {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, 64, 0) != hipSuccess || blocks <= 0) return 8;
    return (uint32_t)blocks;
}

void nd_resident_march(uint32_t* r) {
    r[RES_MARCH] = resident_of(nd::k_march<true, false, false>);
    r[RES_MARCH_MIXED] = resident_of(nd::k_march<false, false, false>);
    r[RES_MARCH_ACC] = resident_of(nd::k_march<true, true, false>);
    r[RES_MARCH_ACC_MIXED] = resident_of(nd::k_march<false, true, false>);
    r[RES_MARCH_EXT] = resident_of(nd::k_march<false, false, true>);
    r[RES_MARCH_ACC_EXT] = resident_of(nd::k_march<false, true, true>);
    r[RES_SURFACE] = resident_of(nd::k_surface<true, false>);
    r[RES_SURFACE_MIXED] = resident_of(nd::k_surface<false, false>);
    r[RES_SURFACE_EXT] = resident_of(nd::k_surface<false, true>);
    r[RES_SHADOW] = resident_of(nd::k_shadow<true, false, false>);
    r[RES_SHADOW_MIXED] = resident_of(nd::k_shadow<false, false, false>);
    r[RES_SHADOW_ACC] = resident_of(nd::k_shadow<true, true, false>);
    r[RES_SHADOW_ACC_MIXED] = resident_of(nd::k_shadow<false, true, false>);
    r[RES_SHADOW_EXT] = resident_of(nd::k_shadow<false, false, true>);
    r[RES_SHADOW_ACC_EXT] = resident_of(nd::k_shadow<false, true, true>);
}
