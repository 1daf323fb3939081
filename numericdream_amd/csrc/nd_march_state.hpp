// nd_march_state.hpp - the per-ray state machine of march_to (template.wgsl:1859-1963):
// sphere-tracing probe, on-hit refinement and bisection, shared by every march pass.
#pragma once

#include "nd_de.hpp"

namespace nd {

enum { MS_CONT = 0, MS_HIT = 1, MS_MISS = 2 };
struct March {
    float t, lo, hi;
    int i, refines, n, evals;
    int probing, bisecting, resolved;
};
NDD void march_init(March& m) {
    m.t = 0.f; m.lo = 0.f; m.hi = 0.f;
    m.i = 0; m.refines = 0; m.n = 0; m.evals = 0;
    m.probing = 1; m.bisecting = 0; m.resolved = 0;
}
NDD float march_sample_t(const March& m) { return m.probing ? m.t : 0.5f * (m.lo + m.hi); }
// March precision of one ray, held in registers by the persistent passes.
struct MarchParams {
    float eps_c, eps_f, max_dist, footprint;
    int max_steps, bisect_budget, refine_budget, hq;
};
NDD MarchParams march_params(const DevJob& J) {
    MarchParams p;
    p.eps_c = J.eps_c; p.eps_f = J.eps_f; p.max_dist = J.max_dist; p.footprint = J.footprint;
    p.max_steps = J.max_steps; p.bisect_budget = J.bisect_budget; p.refine_budget = J.refine_budget; p.hq = J.hq;
    return p;
}
// march_update without probe refinement (refine_budget == 0: the export settings
// high_quality = 1, refine_hq = 20 give plain sphere tracing, then bisection of
// [t_prev, t] on the first d < eps). Same results as the general update below for that
// case, with less state: a probe's hi is t + d and becomes the next t; the sample budget
// max_steps + bisect_budget cannot be reached before max_steps probes or the bisection
// budget end the ray, so n is not tracked.
template <class MP>
NDD int march_update_norefine(March& m, const MP& J, float s, float d, float* dist) {
    m.evals++;
    const float eps = J.eps_c * s + J.eps_f;
    const bool below = d < eps;
    if (m.bisecting) {
        if (below) m.hi = s; else m.lo = s;
        m.refines++;
        if (m.hi - m.lo <= eps || m.refines >= J.bisect_budget) { *dist = m.hi; return MS_HIT; }
        return MS_CONT;
    }
    if (below) {
        if (J.bisect_budget > 0 && s - m.lo > eps) {
            m.bisecting = 1; m.probing = 0; m.hi = s; m.refines = 0;
            return MS_CONT;
        }
        *dist = s;
        return MS_HIT;
    }
    m.lo = m.t;
    m.t = m.t + d;
    m.i++;
    if (m.t > J.max_dist || m.i >= J.max_steps) { *dist = m.t; return MS_MISS; }
    return MS_CONT;
}

// One accelerator step at a probe (template.wgsl:1890-1900), given the grid's safe distance
// at the probe point: when it exceeds eps(t), t advances by it without using a DE sample and
// the step counts as a march step. Returns MS_CONT after such a skip, MS_MISS (*dist = t)
// when the skip ends the ray at the distance or step limit, and -1 when no skip applies and
// the DE sample at t decides (march_update).
template <class MP>
NDD int march_skip_step(March& m, const MP& J, float skip, float* dist) {
    const float s = m.t;
    if (!(skip > J.eps_c * s + J.eps_f)) return -1;
    m.lo = s;
    m.t = s + skip;
    m.i++;
    m.n++;
    if (m.t > J.max_dist || m.i >= J.max_steps) { *dist = m.t; return MS_MISS; }
    return MS_CONT;
}

// One step of march_to's loop after the DE d at sample_t s. *dist receives the result.
// J is a DevJob (plain kernel, wave-uniform) or MarchParams (persistent, per lane).
template <class MP>
NDD int march_update(March& m, const MP& J, float s, float d, float* dist) {
    if (J.refine_budget == 0) return march_update_norefine(m, J, s, d, dist);
    const int max_samples = J.max_steps * (1 + J.refine_budget) + J.bisect_budget;
    m.n++;
    m.evals++;
    const float eps = J.eps_c * s + J.eps_f;
    if (m.bisecting) {
        if (d < eps) m.hi = s; else m.lo = s;
        m.refines++;
        if (m.hi - m.lo <= eps || m.refines >= J.bisect_budget) { *dist = m.hi; return MS_HIT; }
        if (m.n >= max_samples) { *dist = m.t; return MS_MISS; }
        return MS_CONT;
    }
    if (d < eps) {
        if (J.refine_budget == 0 && J.bisect_budget > 0 && s - m.lo > eps) {
            m.bisecting = 1; m.probing = 0; m.hi = s; m.refines = 0;
            if (m.n >= max_samples) { *dist = m.t; return MS_MISS; }
            return MS_CONT;
        }
        *dist = s;
        return MS_HIT;
    }
    if (m.probing) {
        m.lo = m.t; m.hi = m.t + d; m.refines = 0; m.resolved = 0; m.probing = 0;
    } else {
        if (d >= m.hi - s) m.resolved = 1; else m.hi = s;
        m.refines++;
    }
    if (m.resolved || m.refines >= J.refine_budget) {
        if (J.hq && !m.resolved && J.refine_budget > 0)
            m.hi = fminf(m.hi, m.lo + fmaxf(m.t, 1.f) * J.footprint);
        m.t = m.hi;
        m.i++;
        if (m.t > J.max_dist || m.i >= J.max_steps) { *dist = m.t; return MS_MISS; }
        m.probing = 1;
    }
    if (m.n >= max_samples) { *dist = m.t; return MS_MISS; }
    return MS_CONT;
}

// The wave's next chunk of a work queue of n items whose shared head is *head: the first
// item of the chunk, n or more once the queue is drained. Device-scope atomics on the one
// address every wave of the launch shares are serialised, about 24 ns each: 7,168 waves that
// each take one on an empty queue hold the launch for 0.17 ms. With peek set (a wave's first
// claim), lane 0 reads the head before it takes the atomic, so the waves of a launch whose
// queue is empty leave with a load. Later claims take the atomic alone: a read before each of
// them doubles the traffic on the shared address, which busy passes feel.
NDD uint32_t claim_chunk(uint32_t* head, uint32_t n, uint32_t chunk, bool peek) {
    uint32_t base = n;
    if (__lane_id() == 0) {
        if (peek) base = __hip_atomic_load(head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!peek || base < n) base = atomicAdd(head, chunk);
    }
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)base);
}

}  // namespace nd
