// nd_de.hpp - device distance estimators for gfx950.
//
// The four tuned formula plug-ins (mandelbrot.wgsl, formulas/amazingbox_box.wgsl,
// formulas/menger3_ifs.wgsl, formulas/kleinian_kleinian.wgsl; the others are in
// nd_formulas.hpp), the formula wrapper with
// orbit traps (fr.zig:490-513, template.wgsl:1207-1289), the instance transform and the
// seven combine modes (template.wgsl:1600-1779), written as HIP device templates over the
// scalar type:
//   float      the march, shadow and trap evaluations
//   DualN<N>   forward-mode derivatives: N = 3 gives the analytic normal, N = 6 the
//              adjoint (3 spatial + 3 parameter tangents)
// All arithmetic is FP32. Reciprocal, square root, log2 and exp2 use the hardware
// transcendental unit (v_rcp_f32, v_sqrt_f32, v_log_f32, v_exp_f32).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "nd_dev.h"

#define NDD __device__ __forceinline__

namespace nd {

// ----------------------------------------------------------------------------- float math
NDD float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
NDD float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
NDD float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }  // v_sqrt_f32, 1 ulp
NDD float flog2(float x) { return __builtin_amdgcn_logf(x); }
NDD float fexp2(float x) { return __builtin_amdgcn_exp2f(x); }
NDD float fln(float x) { return __builtin_amdgcn_logf(x) * 0.69314718055994531f; }
NDD float fmin_(float a, float b) { return fminf(a, b); }
NDD float fmax_(float a, float b) { return fmaxf(a, b); }
NDD float fclamp(float x, float lo, float hi) { return __builtin_amdgcn_fmed3f(x, lo, hi); }  // v_med3_f32
NDD float fabs_(float x) { return fabsf(x); }
// One rounding per operation, for arithmetic the reference's FP32 results depend on bit for
// bit (hashes of float bits, lattice hashes, grid centres). HIP compiles with
// -ffp-contract=fast and this toolchain's __fmul_rn / __fadd_rn are plain operators, so
// a*b + c would otherwise become one fused multiply-add; contraction is off inside these.
NDD float mul_rn(float a, float b) {
#pragma clang fp contract(off)
    return a * b;
}
NDD float add_rn(float a, float b) {
#pragma clang fp contract(off)
    return a + b;
}
NDD float sub_rn(float a, float b) {
#pragma clang fp contract(off)
    return a - b;
}

// ----------------------------------------------------------------------------- dual numbers
template <int N>
struct DualN {
    float v;
    float d[N];
};
template <int N> NDD DualN<N> dconst(float x) {
    DualN<N> r; r.v = x;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = 0.f;
    return r;
}
template <int N> NDD DualN<N> dvar(float x, int slot) {
    DualN<N> r = dconst<N>(x);
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = (k == slot) ? 1.f : 0.f;
    return r;
}
template <int N> NDD DualN<N> dchain(const DualN<N>& a, float fv, float dfa) {
    DualN<N> r; r.v = fv;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = dfa * a.d[k];
    return r;
}

// Scalar traits: S is float or DualN<N>.
template <class S> struct Sc;
template <> struct Sc<float> {
    static NDD float c(float x) { return x; }
    static NDD float v(float x) { return x; }
};
template <int N> struct Sc<DualN<N>> {
    static NDD DualN<N> c(float x) { return dconst<N>(x); }
    static NDD float v(const DualN<N>& x) { return x.v; }
};
template <class S> NDD float vv(const S& x) { return Sc<S>::v(x); }

template <int N> NDD DualN<N> operator+(const DualN<N>& a, const DualN<N>& b) {
    DualN<N> r; r.v = a.v + b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
    return r; }
template <int N> NDD DualN<N> operator-(const DualN<N>& a, const DualN<N>& b) {
    DualN<N> r; r.v = a.v - b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
    return r; }
template <int N> NDD DualN<N> operator-(const DualN<N>& a) {
    DualN<N> r; r.v = -a.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
    return r; }
template <int N> NDD DualN<N> operator*(const DualN<N>& a, const DualN<N>& b) {
    DualN<N> r; r.v = a.v * b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r; }
template <int N> NDD DualN<N> operator*(const DualN<N>& a, float s) {
    DualN<N> r; r.v = a.v * s;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * s;
    return r; }
template <int N> NDD DualN<N> operator*(float s, const DualN<N>& a) { return a * s; }
template <int N> NDD DualN<N> operator+(const DualN<N>& a, float s) { DualN<N> r = a; r.v += s; return r; }
template <int N> NDD DualN<N> operator-(const DualN<N>& a, float s) { DualN<N> r = a; r.v -= s; return r; }
template <int N> NDD DualN<N> operator/(const DualN<N>& a, const DualN<N>& b) {
    float ib = frcp(b.v);
    DualN<N> r; r.v = a.v * ib;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r; }
template <int N> NDD bool operator<(const DualN<N>& a, const DualN<N>& b) { return a.v < b.v; }
template <int N> NDD bool operator>(const DualN<N>& a, const DualN<N>& b) { return a.v > b.v; }
template <int N> NDD bool operator<=(const DualN<N>& a, const DualN<N>& b) { return a.v <= b.v; }
template <int N> NDD bool operator>=(const DualN<N>& a, const DualN<N>& b) { return a.v >= b.v; }

// Generic scalar functions (float overloads + dual overloads).
NDD float sdiv(float a, float b) { return fdiv(a, b); }
template <int N> NDD DualN<N> sdiv(const DualN<N>& a, const DualN<N>& b) { return a / b; }
NDD float ssqrt(float a) { return fsqrt(a); }
template <int N> NDD DualN<N> ssqrt(const DualN<N>& a) {
    float s = fsqrt(a.v);
    return dchain(a, s, s > 0.f ? 0.5f * frcp(s) : 0.f); }
NDD float sabs(float a) { return fabsf(a); }
template <int N> NDD DualN<N> sabs(const DualN<N>& a) { return a.v < 0.f ? -a : a; }
NDD float smin_(float a, float b) { return fminf(a, b); }
template <int N> NDD DualN<N> smin_(const DualN<N>& a, const DualN<N>& b) { return b.v < a.v ? b : a; }
NDD float smax_(float a, float b) { return fmaxf(a, b); }
template <int N> NDD DualN<N> smax_(const DualN<N>& a, const DualN<N>& b) { return b.v > a.v ? b : a; }
template <class S> NDD S sclamp(const S& x, const S& lo, const S& hi) { return smin_(smax_(x, lo), hi); }
NDD float sln(float a) { return fln(a); }
template <int N> NDD DualN<N> sln(const DualN<N>& a) { return dchain(a, fln(a.v), frcp(a.v)); }
NDD float srint(float a) { return __builtin_rintf(a); }
template <int N> NDD DualN<N> srint(const DualN<N>& a) { return dconst<N>(__builtin_rintf(a.v)); }
// r^e for r >= 0 through exp2(e log2 r)
NDD float spow(float r, float e) { return fexp2(e * flog2(r)); }
template <int N> NDD DualN<N> spow(const DualN<N>& r, const DualN<N>& e) {
    float lr = flog2(r.v);
    float p = fexp2(e.v * lr);
    DualN<N> o; o.v = p;
    float dr = r.v != 0.f ? e.v * fexp2((e.v - 1.f) * lr) : 0.f;
    float de = r.v > 0.f ? p * lr * 0.69314718055994531f : 0.f;
#pragma unroll
    for (int k = 0; k < N; ++k) o.d[k] = dr * r.d[k] + de * e.d[k];
    return o; }
NDD float sacos(float a) { return acosf(a); }
template <int N> NDD DualN<N> sacos(const DualN<N>& a) {
    float s = fsqrt(fmaxf(1.f - a.v * a.v, 0.f));
    return dchain(a, acosf(a.v), s > 0.f ? -frcp(s) : 0.f); }
NDD float satan2(float y, float x) { return atan2f(y, x); }
template <int N> NDD DualN<N> satan2(const DualN<N>& y, const DualN<N>& x) {
    float den = x.v * x.v + y.v * y.v;
    float id = den > 0.f ? frcp(den) : 0.f;
    DualN<N> r; r.v = atan2f(y.v, x.v);
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = (x.v * y.d[k] - y.v * x.d[k]) * id;
    return r; }
NDD void ssincos(float a, float* s, float* c) { sincosf(a, s, c); }
template <int N> NDD void ssincos(const DualN<N>& a, DualN<N>* s, DualN<N>* c) {
    float sv, cv; sincosf(a.v, &sv, &cv);
    *s = dchain(a, sv, cv); *c = dchain(a, cv, -sv); }

// ----------------------------------------------------------------------------- vec3
template <class S> struct V3 { S x, y, z; };
template <class S> NDD V3<S> mk3(S x, S y, S z) { V3<S> r; r.x = x; r.y = y; r.z = z; return r; }
template <class S> NDD V3<S> cst3(float x, float y, float z) { return mk3(Sc<S>::c(x), Sc<S>::c(y), Sc<S>::c(z)); }
template <class S> NDD V3<S> operator+(const V3<S>& a, const V3<S>& b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class S> NDD V3<S> operator-(const V3<S>& a, const V3<S>& b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class S> NDD V3<S> operator*(const V3<S>& a, const S& s) { return mk3(a.x * s, a.y * s, a.z * s); }
template <int N> NDD V3<DualN<N>> operator*(const V3<DualN<N>>& a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
template <class S> NDD S dot3(const V3<S>& a, const V3<S>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class S> NDD S len3(const V3<S>& a) { return ssqrt(dot3(a, a)); }
template <class S> NDD V3<float> val3(const V3<S>& a) { return mk3(vv(a.x), vv(a.y), vv(a.z)); }

// ----------------------------------------------------------------------------- parameters
// Per-formula parameters as scalars of type S. For float they come from the host-hoisted
// constants h[]; for dual numbers they are rebuilt from p[] with the adjoint tangent of
// each @param attached (DevInst::dparam).
// I is DevInst or DevMixin: both carry p, h, dparam, rot_on and rot_cs.
template <class S, class I> NDD S prm(const I& in, int k) {
    if constexpr (std::is_same<S, float>::value) {
        return in.p[k];
    } else {
        S r = Sc<S>::c(in.p[k]);
        int slot = in.dparam[k];
#pragma unroll
        for (int j = 0; j < (int)(sizeof(r.d) / sizeof(float)); ++j) r.d[j] = (j == slot) ? 1.f : 0.f;
        return r;
    }
}

template <class S> struct Carry { V3<S> z; S dr; };

// formula-internal Euler rotation: z' = Rz (Ry (Rx z)) (template.wgsl:1087-1093, 1393-1408)
template <class S> NDD V3<S> rotate_cs(const V3<S>& z, const float* cs) {
    const float cx = cs[0], sx = cs[1], cy = cs[2], sy = cs[3], cz = cs[4], sz = cs[5];
    V3<S> a = mk3(z.x, z.y * cx - z.z * sx, z.y * sx + z.z * cx);
    V3<S> b = mk3(a.x * cy + a.z * sy, a.y, a.z * cy - a.x * sy);
    return mk3(b.x * cz - b.y * sz, b.x * sz + b.y * cz, b.z);
}

// ----------------------------------------------------------------------------- Mandelbulb
// bulb_power_step (template.wgsl:1019-1031), general power through acos/atan2/pow/sincos.
template <class S> NDD Carry<S> bulb_step_trig(const Carry<S>& c, const V3<S>& add, const S& power, bool& escaped) {
    S r = len3(c.z);
    if (vv(r) > 2.0f) { escaped = true; return c; }
    S theta = sacos(sclamp(sdiv(c.z.z, r), Sc<S>::c(-1.f), Sc<S>::c(1.f))) * power;
    S phi = satan2(c.z.y, c.z.x) * power;
    S rp = spow(r, power - Sc<S>::c(1.f));
    S zr = rp * r;
    S dr = rp * power * c.dr + Sc<S>::c(1.f);
    S st, ct, sp, cp;
    ssincos(theta, &st, &ct);
    ssincos(phi, &sp, &cp);
    Carry<S> o;
    o.z = mk3(zr * st * cp + add.x, zr * st * sp + add.y, zr * ct + add.z);
    o.dr = dr;
    return o;
}

// Power 8 without transcendentals: with rho^2 = x^2 + y^2, cos(theta) = z / r,
// sin(theta) = rho / r (theta in [0, pi]) and cos(phi) + i sin(phi) = (x + i y) / rho,
// z^8 = (Im[(z + i rho)^8] Re[(x + i y)^8] / rho^8, Im[..] Im[(x + i y)^8] / rho^8,
//        Re[(z + i rho)^8]).  dr' = 8 r^7 dr + 1.
NDD Carry<float> bulb_step_pow8(const Carry<float>& c, const V3<float>& add, bool& escaped) {
    float x = c.z.x, y = c.z.y, z = c.z.z;
    float rho2 = x * x + y * y;
    float r2 = rho2 + z * z;
    if (r2 > 4.0f) { escaped = true; return c; }
    float rho = fsqrt(rho2);
    // (z + i rho)^8 by three squarings
    float a = z * z - rho2, b = 2.f * z * rho;
    float a2 = a * a - b * b, b2 = 2.f * a * b;
    float ct8 = a2 * a2 - b2 * b2, st8 = 2.f * a2 * b2;  // r^8 cos(8 theta), r^8 sin(8 theta)
    // ((x + i y) / rho)^8 = cos(8 phi) + i sin(8 phi); phi = atan2(0, 0) = 0 on the z axis
    float irho = rho2 > 0.f ? frcp(rho) : 0.f;
    float ux = rho2 > 0.f ? x * irho : 1.f, uy = y * irho;
    float u = ux * ux - uy * uy, v = 2.f * ux * uy;
    float u2 = u * u - v * v, v2 = 2.f * u * v;
    float cphi = u2 * u2 - v2 * v2, sphi = 2.f * u2 * v2;
    float r4 = r2 * r2;
    float r = fsqrt(r2);
    Carry<float> o;
    o.z = mk3(st8 * cphi + add.x, st8 * sphi + add.y, ct8 + add.z);
    o.dr = 8.f * (r4 * r2) * r * c.dr + 1.f;
    return o;
}

// bulb_log_escape (template.wgsl:1006-1010)
template <class S> NDD S bulb_finalize(const Carry<S>& c) {
    S r = len3(c.z);
    S l = Sc<S>::c(0.5f) * sln(smax_(r, Sc<S>::c(1e-6f))) * sdiv(r, smax_(c.dr, Sc<S>::c(1e-6f)));
    return smin_(l, r);
}

// ----------------------------------------------------------------------------- Mandelbox
// amazingbox de_step (formulas/amazingbox_box.wgsl:18-33) with box_fold and sphere_fold
// (template.wgsl:947-969). posc = pos * CScale + julia, hoisted out of the iteration.
struct BoxF {
    float S, absS, m2, F2, kF, L, dr_add;
};
template <class S> struct BoxT { S Sc_, m2, F2, L; };
// Branch-free: the fold factor is selected, the reciprocal always computed (a zero r2 is
// covered by the r2 < m2 arm), the rotation is a compile-time choice and the fold factor
// is folded into the scale: z' = z (t S) + posc, dr' = dr (t |S|) + max(|cs|, 1).
// F1: the fixed radius is exactly 1, so F2 / r2 is the reciprocal itself (1 x is exact: the
// same results with one multiply less per iteration).
template <bool ROT, bool CLAMP, bool F1 = false>
NDD Carry<float> box_step(const Carry<float>& c, const V3<float>& posc, const BoxF& b, const float* rot_cs) {
    V3<float> z;
    z.x = fclamp(c.z.x, -b.L, b.L) * 2.f - c.z.x;
    z.y = fclamp(c.z.y, -b.L, b.L) * 2.f - c.z.y;
    z.z = fclamp(c.z.z, -b.L, b.L) * 2.f - c.z.z;
    const float r2 = dot3(z, z);
    // sphere_fold's three arms (template.wgsl:959-969) as one clamp: with m2 <= F2,
    // F2 / r2 > F2 / m2 exactly when r2 < m2 and F2 / r2 <= 1 exactly when r2 >= F2, so
    // t = clamp(F2 / r2, 1, F2 / m2). The host keeps m2 > F2 off this path (kF < 1).
    float t = F1 ? frcp(r2) : b.F2 * frcp(r2);
    if (CLAMP) {
        t = __builtin_amdgcn_fmed3f(t, 1.f, b.kF);
    } else {
        t = r2 < b.F2 ? t : 1.f;
        t = r2 < b.m2 ? b.kF : t;
    }
    Carry<float> o;
    if (ROT) {
        z = rotate_cs(mk3(z.x * t, z.y * t, z.z * t), rot_cs);
        o.z = mk3(z.x * b.S + posc.x, z.y * b.S + posc.y, z.z * b.S + posc.z);
        o.dr = c.dr * (t * b.absS) + b.dr_add;
    } else {
        const float ts = t * b.S;
        o.z = mk3(z.x * ts + posc.x, z.y * ts + posc.y, z.z * ts + posc.z);
        // t > 0, so |t S| = t |S| bit for bit: the FMA takes it through its abs modifier
        o.dr = c.dr * fabsf(ts) + b.dr_add;
    }
    return o;
}
template <class S, class I> NDD Carry<S> box_step_t(const Carry<S>& c, const V3<S>& posc, const S& scale, const S& m2,
                                                    const S& F2, const S& L, float dr_add, const I& in) {
    S nL = -L;
    V3<S> z = mk3(sclamp(c.z.x, nL, L) * 2.f - c.z.x, sclamp(c.z.y, nL, L) * 2.f - c.z.y,
                  sclamp(c.z.z, nL, L) * 2.f - c.z.z);
    S r2 = dot3(z, z);
    S dr = c.dr;
    if (vv(r2) < vv(m2)) {
        S t = sdiv(F2, m2);
        z = z * t; dr = dr * t;
    } else if (vv(r2) < vv(F2)) {
        S t = sdiv(F2, r2);
        z = z * t; dr = dr * t;
    }
    if (in.rot_on) z = rotate_cs(z, in.rot_cs);
    Carry<S> o;
    o.z = mk3(z.x * scale + posc.x, z.y * scale + posc.y, z.z * scale + posc.z);
    o.dr = dr * sabs(scale) + Sc<S>::c(dr_add);
    return o;
}

// ----------------------------------------------------------------------------- Menger3
// formulas/menger3_ifs.wgsl:12-24 with sort_desc_abs (template.wgsl:1036-1054)
template <class S> NDD V3<S> sort_desc_abs(const V3<S>& z) {
    S a = sabs(z.x), b = sabs(z.y), c = sabs(z.z), t;
    if (vv(a) < vv(b)) { t = a; a = b; b = t; }
    if (vv(a) < vv(c)) { t = a; a = c; c = t; }
    if (vv(b) < vv(c)) { t = b; b = c; c = t; }
    return mk3(a, b, c);
}
template <class S, class I> NDD Carry<S> menger_step(const Carry<S>& c, const I& in) {
    V3<S> z = sort_desc_abs(c.z);
    if (in.rot_on) z = rotate_cs(z, in.rot_cs);
    S sc = prm<S>(in, 4);
    S sm1 = sc - Sc<S>::c(1.f);
    V3<S> o = mk3(prm<S>(in, 1) * sm1, prm<S>(in, 2) * sm1, prm<S>(in, 3) * sm1);
    z = mk3(z.x * sc - o.x, z.y * sc - o.y, z.z * sc - o.z);
    if (vv(z.z) < -0.5f * vv(o.z)) z.z = z.z + o.z;
    Carry<S> r;
    r.z = z;
    r.dr = c.dr * sc;
    return r;
}
template <class S> NDD S menger_finalize(const Carry<S>& c) {
    V3<S> q = mk3(smax_(sabs(c.z.x) - Sc<S>::c(1.f), Sc<S>::c(0.f)), smax_(sabs(c.z.y) - Sc<S>::c(1.f), Sc<S>::c(0.f)),
                  smax_(sabs(c.z.z) - Sc<S>::c(1.f), Sc<S>::c(0.f)));
    return sdiv(len3(q), c.dr);
}

// ----------------------------------------------------------------------------- Kleinian
// formulas/kleinian_kleinian.wgsl:16-33 with box_fold, lattice_wrap (round half to even)
// and sphere_invert_offset (template.wgsl:975-989)
template <class S, class I> NDD Carry<S> kleinian_step(const Carry<S>& c, const V3<S>& posc, const I& in) {
    S L = prm<S>(in, 0);
    S nL = -L;
    V3<S> z = mk3(sclamp(c.z.x, nL, L) * 2.f - c.z.x, sclamp(c.z.y, nL, L) * 2.f - c.z.y,
                  sclamp(c.z.z, nL, L) * 2.f - c.z.z);
    S cell = prm<S>(in, 5);
    const float inv_cell = in.h[6];
    z = mk3(z.x - cell * srint(z.x * inv_cell), z.y - cell * srint(z.y * inv_cell), z.z - cell * srint(z.z * inv_cell));
    V3<S> ctr = mk3(prm<S>(in, 1), prm<S>(in, 2), prm<S>(in, 3));
    S R = prm<S>(in, 4);
    S R2 = R * R;
    V3<S> d = z - ctr;
    S r2 = smax_(dot3(d, d), Sc<S>::c(1e-6f));
    S dr = c.dr;
    if (vv(r2) < vv(R2)) {
        S t = sdiv(R2, r2);
        z = ctr + d * t;
        dr = dr * t;
    }
    Carry<S> o;
    o.z = z + posc;
    o.dr = dr + Sc<S>::c(in.h[11]);
    return o;
}

// ----------------------------------------------------------------------------- orbit trap
// orbit_trap_distance / begin / update / finish (template.wgsl:1227-1289)
NDD float trap_distance(const DevInst& in, const V3<float>& z) {
    V3<float> q = mk3(z.x - in.trap_center[0], z.y - in.trap_center[1], z.z - in.trap_center[2]);
    switch (in.trap_shape) {
        case 1: return fabsf(len3(q) - in.trap_radius);
        case 2: {
            float ex = fabsf(q.x) - in.trap_box[0], ey = fabsf(q.y) - in.trap_box[1], ez = fabsf(q.z) - in.trap_box[2];
            V3<float> e = mk3(fmaxf(ex, 0.f), fmaxf(ey, 0.f), fmaxf(ez, 0.f));
            return fabsf(len3(e) + fminf(fmaxf(ex, fmaxf(ey, ez)), 0.f));
        }
        case 3: return fminf(fabsf(q.x), fminf(fabsf(q.y), fabsf(q.z)));
        case 4: {
            float a = fsqrt(q.x * q.x + q.z * q.z) - in.trap_radius;
            return fabsf(fsqrt(a * a + q.y * q.y) - in.trap_tube);
        }
        default: return len3(q);
    }
}
struct TrapAcc { float x, y, z, w; };
NDD TrapAcc trap_begin(const DevInst& in) {
    bool high = in.trap_mode != 1 && in.trap_mode != 2;
    TrapAcc a; a.x = high ? 1e6f : 0.f; a.y = 0.f; a.z = 0.f; a.w = 1e6f;
    return a;
}
NDD void trap_update(TrapAcc& a, const DevInst& in, const V3<float>& z, const V3<float>& zp, int it) {
    float d = trap_distance(in, z);
    a.w = d;
    if (in.trap_mode == 1) a.x = fmaxf(a.x, d);
    else if (in.trap_mode == 2) {
        if (z.x != zp.x || z.y != zp.y || z.z != zp.z) { a.x += d; a.y += 1.f; }
    } else if (in.trap_mode == 4) {
        if (d < a.x) { a.x = d; a.z = (float)it; }
    } else a.x = fminf(a.x, d);
}
NDD float trap_finish(const TrapAcc& a, const DevInst& in, int iters) {
    if (in.trap_mode == 2) return a.y > 0.5f ? a.x / fmaxf(a.y, 1.f) : a.w;
    if (in.trap_mode == 3) return a.w;
    if (in.trap_mode == 4) return (a.z + 1.f) / (float)max(iters, 1);
    return a.x;
}

// The Mandelbox iteration with the uniform choices (rotation, clamp form) taken outside
// the loop.
template <bool ROT, bool CLAMP, bool TRAP, bool F1 = false>
NDD Carry<float> box_loop(Carry<float> c, const V3<float>& posc, const BoxF& b, const DevInst& in, int iters,
                          TrapAcc& acc) {
    for (int it = 0; it < iters; ++it) {
        V3<float> zp = c.z;
        c = box_step<ROT, CLAMP, F1>(c, posc, b, in.rot_cs);
        if (TRAP) trap_update(acc, in, c.z, zp, it);
    }
    return c;
}

}  // namespace nd

#include "nd_formulas.hpp"

namespace nd {

// ----------------------------------------------------------------------------- formula wrapper
// formula_N (fr.zig:490-513): iterate de_step de_iterations times from (pos, 1), update the
// trap after every step, finalize. The float path exits early once every lane of the wave
// has escaped a Mandelbulb (bulb_power_step is the identity after escape, and so is the
// trap update), which changes no result.
// EXT: the instance may use any formula of nd_formulas.hpp; without it only the four above
// are compiled in (others give NO_SURFACE_DE) and the kernel keeps their register budget.
template <bool TRAP, bool EXT = false>
NDD float formula_de_f(const DevInst& in, const V3<float>& pos, float* trap) {
    if constexpr (EXT) {
        if (in.hybrid) return ext_de_f<TRAP>(in, pos, trap);
    }
    Carry<float> c; c.z = pos; c.dr = 1.f;
    const int iters = in.iters;
    TrapAcc acc;
    if (TRAP) acc = trap_begin(in);
    switch (in.formula) {
        case NDF_BULB: {
            bool esc = false;
            if (in.bulb_pow8) {
                for (int it = 0; it < iters; ++it) {
                    V3<float> zp = c.z;
                    c = bulb_step_pow8(c, pos, esc);
                    if (TRAP) trap_update(acc, in, c.z, zp, it);
                    if (__builtin_amdgcn_ballot_w64(!esc) == 0) break;
                }
            } else {
                const float power = in.h[0];
                for (int it = 0; it < iters; ++it) {
                    V3<float> zp = c.z;
                    c = bulb_step_trig<float>(c, pos, power, esc);
                    if (TRAP) trap_update(acc, in, c.z, zp, it);
                    if (__builtin_amdgcn_ballot_w64(!esc) == 0) break;
                }
            }
            if (TRAP) *trap = trap_finish(acc, in, iters);
            return bulb_finalize<float>(c);
        }
        case NDF_BOX: {
            BoxF b;
            b.S = in.h[0]; b.absS = in.h[1]; b.m2 = in.h[2]; b.F2 = in.h[3]; b.kF = in.h[4]; b.L = in.h[5];
            b.dr_add = in.h[10];
            V3<float> posc = mk3(pos.x * in.h[6] + in.h[7], pos.y * in.h[6] + in.h[8], pos.z * in.h[6] + in.h[9]);
            const bool clamp = b.kF >= 1.f;
            if (in.rot_on) {
                if (clamp) c = box_loop<true, true, TRAP>(c, posc, b, in, iters, acc);
                else c = box_loop<true, false, TRAP>(c, posc, b, in, iters, acc);
            } else if (clamp) {
                if (b.F2 == 1.f) c = box_loop<false, true, TRAP, true>(c, posc, b, in, iters, acc);
                else c = box_loop<false, true, TRAP>(c, posc, b, in, iters, acc);
            } else {
                c = box_loop<false, false, TRAP>(c, posc, b, in, iters, acc);
            }
            if (TRAP) *trap = trap_finish(acc, in, iters);
            return fdiv(len3(c.z), fabsf(c.dr));
        }
        case NDF_MENGER: {
            for (int it = 0; it < iters; ++it) {
                V3<float> zp = c.z;
                c = menger_step<float>(c, in);
                if (TRAP) trap_update(acc, in, c.z, zp, it);
            }
            if (TRAP) *trap = trap_finish(acc, in, iters);
            return menger_finalize<float>(c);
        }
        case NDF_KLEINIAN: {
            V3<float> posc = mk3(pos.x * in.h[7] + in.h[8], pos.y * in.h[7] + in.h[9], pos.z * in.h[7] + in.h[10]);
            for (int it = 0; it < iters; ++it) {
                V3<float> zp = c.z;
                c = kleinian_step<float>(c, posc, in);
                if (TRAP) trap_update(acc, in, c.z, zp, it);
            }
            if (TRAP) *trap = trap_finish(acc, in, iters);
            return fdiv(len3(c.z), fabsf(c.dr));
        }
        default:
            if constexpr (EXT) return ext_de_f<TRAP>(in, pos, trap);
            if (TRAP) *trap = 0.f;
            return 1e6f;  // NO_SURFACE_DE
    }
}

// The Mandelbox step of the dual path with its parameters as plain floats (CONSTP below):
// the same value arithmetic as box_step_t, without the zero parameter tangents that a dual
// constant carries through every product.
template <class S, class I> NDD Carry<S> box_step_fc(const Carry<S>& c, const V3<S>& posc, float scale, float t_in,
                                                     float m2, float F2, float L, float dr_add, const I& in) {
    const S nL = Sc<S>::c(-L), pL = Sc<S>::c(L);
    V3<S> z = mk3(sclamp(c.z.x, nL, pL) * 2.f - c.z.x, sclamp(c.z.y, nL, pL) * 2.f - c.z.y,
                  sclamp(c.z.z, nL, pL) * 2.f - c.z.z);
    S r2 = dot3(z, z);
    S dr = c.dr;
    if (vv(r2) < m2) {
        z = mk3(z.x * t_in, z.y * t_in, z.z * t_in); dr = dr * t_in;
    } else if (vv(r2) < F2) {
        S t = sdiv(Sc<S>::c(F2), r2);
        z = z * t; dr = dr * t;
    }
    if (in.rot_on) z = rotate_cs(z, in.rot_cs);
    Carry<S> o;
    o.z = mk3(z.x * scale + posc.x, z.y * scale + posc.y, z.z * scale + posc.z);
    o.dr = dr * fabsf(scale) + dr_add;
    return o;
}

// Dual-number version (normals, adjoint). The trap is not needed on these paths. CONSTP:
// the tangents are of the position only (analytic normals), so the formula's parameters
// are plain constants; without it they carry their adjoint tangent slots (dparam).
template <class S, bool EXT = false, bool CONSTP = false>
NDD S formula_de_d(const DevInst& in, const V3<S>& pos) {
    if constexpr (EXT) {
        if (in.hybrid) return ext_de_d<S>(in, pos);
    }
    Carry<S> c; c.z = pos; c.dr = Sc<S>::c(1.f);
    const int iters = in.iters;
    switch (in.formula) {
        case NDF_BULB: {
            S power = prm<S>(in, 0);
            bool esc = false;
            for (int it = 0; it < iters && !esc; ++it) c = bulb_step_trig<S>(c, pos, power, esc);
            return bulb_finalize<S>(c);
        }
        case NDF_BOX: {
            if constexpr (CONSTP) {
                const float scale = in.p[0], m = in.p[1], F = in.p[2], L = in.p[3], cs = in.p[8];
                const float m2 = m * m, F2 = F * F;
                const float t_in = vv(sdiv(Sc<S>::c(F2), Sc<S>::c(m2)));
                V3<S> posc = mk3(pos.x * cs + in.p[9], pos.y * cs + in.p[10], pos.z * cs + in.p[11]);
                for (int it = 0; it < iters; ++it) c = box_step_fc<S>(c, posc, scale, t_in, m2, F2, L, in.h[10], in);
                return sdiv(len3(c.z), sabs(c.dr));
            }
            S scale = prm<S>(in, 0);
            S m = prm<S>(in, 1), F = prm<S>(in, 2);
            S m2 = m * m, F2 = F * F;
            S L = prm<S>(in, 3);
            S cs = prm<S>(in, 8);
            V3<S> posc = mk3(pos.x * cs + prm<S>(in, 9), pos.y * cs + prm<S>(in, 10), pos.z * cs + prm<S>(in, 11));
            for (int it = 0; it < iters; ++it) c = box_step_t<S>(c, posc, scale, m2, F2, L, in.h[10], in);
            return sdiv(len3(c.z), sabs(c.dr));
        }
        case NDF_MENGER: {
            for (int it = 0; it < iters; ++it) c = menger_step<S>(c, in);
            return menger_finalize<S>(c);
        }
        case NDF_KLEINIAN: {
            S cs = prm<S>(in, 7);
            V3<S> posc = mk3(pos.x * cs + prm<S>(in, 8), pos.y * cs + prm<S>(in, 9), pos.z * cs + prm<S>(in, 10));
            for (int it = 0; it < iters; ++it) c = kleinian_step<S>(c, posc, in);
            return sdiv(len3(c.z), sabs(c.dr));
        }
        default:
            if constexpr (EXT) return ext_de_d<S>(in, pos);
            return Sc<S>::c(1e6f);
    }
}

// ----------------------------------------------------------------------------- instance & scene
// instance_local / instance_de_factor (template.wgsl:1600-1630)
template <class S> NDD V3<S> instance_local(const DevInst& in, const V3<S>& p) {
    if (in.identity) return p;
    V3<S> q = mk3(p.x - Sc<S>::c(in.offset[0]), p.y - Sc<S>::c(in.offset[1]), p.z - Sc<S>::c(in.offset[2]));
    const float* f = in.frame;
    V3<S> l = mk3(q.x * f[0] + q.y * f[1] + q.z * f[2], q.x * f[3] + q.y * f[4] + q.z * f[5],
                  q.x * f[6] + q.y * f[7] + q.z * f[8]);
    return mk3(l.x * in.inv_s[0], l.y * in.inv_s[1], l.z * in.inv_s[2]);
}

// smooth combine family and combine_de (template.wgsl:1651-1730); returns (d, h).
NDD void combine_de(int mode, float d, float di, float k, float* od, float* oh) {
    const bool hard = k <= 0.0001f;
    const float closer = di < d ? 0.f : 1.f;
    if (mode == 0 || hard) {
        // every mode degenerates to its hard operation (min / max / max(-cut, base))
        if (mode == 2) { *od = fmaxf(d, di); *oh = di > d ? 0.f : 1.f; return; }
        if (mode == 3) { *od = fmaxf(-di, d); *oh = -di > d ? 0.f : 1.f; return; }
        *od = fminf(d, di); *oh = closer; return;
    }
    const float ik = frcp(k);
    if (mode == 1 || mode == 6) {
        float h = fclamp(0.5f + 0.5f * (di - d) * ik, 0.f, 1.f);
        float m = di * (1.f - h) + d * h;  // mix(di, d, h)
        *od = mode == 6 ? m : m - k * h * (1.f - h);
        *oh = h;
        return;
    }
    if (mode == 2) {
        float h = fclamp(0.5f - 0.5f * (di - d) * ik, 0.f, 1.f);
        *od = di * (1.f - h) + d * h + k * h * (1.f - h);
        *oh = h;
        return;
    }
    if (mode == 3) {  // ssub(cut = di, base = d)
        float h = fclamp(0.5f - 0.5f * (d + di) * ik, 0.f, 1.f);
        *od = d * (1.f - h) + (-di) * h + k * h * (1.f - h);
        *oh = 1.f - h;
        return;
    }
    if (mode == 4) {
        *od = fminf(di - fmaxf(k - d, 0.f), d - fmaxf(k - di, 0.f));
        *oh = closer;
        return;
    }
    float ra = fmaxf(k - d, 0.f), rb = fmaxf(k - di, 0.f);
    *od = fminf(di - ra * (k - rb) * ik, d - rb * (k - ra) * ik);
    *oh = closer;
}

// scene_de (template.wgsl:1756-1779) over a DevScene; TRAP also returns the trap blended
// like hit_material blends materials (template.wgsl:2128-2155).
template <bool TRAP, bool EXT = false>
NDD float scene_de(const DevScene& sc, const V3<float>& p, float* trap) {
    float d = sc.empty_de, tr = 0.f;
    bool first = true;
    const int count = sc.count;
    for (int i = 0; i < count; ++i) {
        const DevInst& in = sc.inst[i];
        if (in.hidden) continue;
        float ti = 0.f;
        float di = formula_de_f<TRAP, EXT>(in, instance_local<float>(in, p), &ti) * in.de_factor;
        if (first) {
            d = di; tr = ti; first = false;
        } else {
            float nd, h;
            combine_de(in.combine, d, di, in.blend_k, &nd, &h);
            if (TRAP) tr = ti * (1.f - h) + tr * h;
            d = nd;
        }
    }
    if (first) { if (TRAP) *trap = 0.f; return sc.empty_de; }
    for (int i = 0; i < sc.carve_count; ++i) {
        const float* c = sc.carves[i];
        float cut = len3(mk3(p.x - c[0], p.y - c[1], p.z - c[2])) - c[3];
        float k = 0.2f * c[3];
        if (k <= 0.0001f) d = fmaxf(-cut, d);
        else {
            float h = fclamp(0.5f - 0.5f * (d + cut) / k, 0.f, 1.f);
            d = d * (1.f - h) + (-cut) * h + k * h * (1.f - h);
        }
    }
    if (TRAP) *trap = tr;
    return d;
}

// Dual-number scene DE (union/smooth combines differentiated through their active branch).
template <class S, bool EXT = false, bool CONSTP = false>
NDD S scene_de_d(const DevScene& sc, const V3<S>& p) {
    S d = Sc<S>::c(sc.empty_de);
    bool first = true;
    for (int i = 0; i < sc.count; ++i) {
        const DevInst& in = sc.inst[i];
        if (in.hidden) continue;
        S di = formula_de_d<S, EXT, CONSTP>(in, instance_local<S>(in, p)) * in.de_factor;
        if (first) { d = di; first = false; continue; }
        const float k = in.blend_k;
        const int mode = in.combine;
        if (mode == 0 || k <= 0.0001f) {
            if (mode == 2) d = smax_(d, di);
            else if (mode == 3) d = smax_(-di, d);
            else d = smin_(d, di);
        } else if (mode == 1 || mode == 6) {
            float hv = fclamp(0.5f + 0.5f * (vv(di) - vv(d)) / k, 0.f, 1.f);
            S h = (hv > 0.f && hv < 1.f) ? Sc<S>::c(0.5f) + (di - d) * (0.5f / k) : Sc<S>::c(hv);
            S m = di * (Sc<S>::c(1.f) - h) + d * h;
            d = mode == 6 ? m : m - h * (Sc<S>::c(1.f) - h) * k;
        } else if (mode == 2) {
            float hv = fclamp(0.5f - 0.5f * (vv(di) - vv(d)) / k, 0.f, 1.f);
            S h = (hv > 0.f && hv < 1.f) ? Sc<S>::c(0.5f) - (di - d) * (0.5f / k) : Sc<S>::c(hv);
            d = di * (Sc<S>::c(1.f) - h) + d * h + h * (Sc<S>::c(1.f) - h) * k;
        } else if (mode == 3) {
            float hv = fclamp(0.5f - 0.5f * (vv(d) + vv(di)) / k, 0.f, 1.f);
            S h = (hv > 0.f && hv < 1.f) ? Sc<S>::c(0.5f) - (d + di) * (0.5f / k) : Sc<S>::c(hv);
            d = d * (Sc<S>::c(1.f) - h) - di * h + h * (Sc<S>::c(1.f) - h) * k;
        } else if (mode == 4) {
            d = smin_(di - smax_(Sc<S>::c(k) - d, Sc<S>::c(0.f)), d - smax_(Sc<S>::c(k) - di, Sc<S>::c(0.f)));
        } else {
            S ra = smax_(Sc<S>::c(k) - d, Sc<S>::c(0.f)), rb = smax_(Sc<S>::c(k) - di, Sc<S>::c(0.f));
            d = smin_(di - ra * (Sc<S>::c(k) - rb) * (1.f / k), d - rb * (Sc<S>::c(k) - ra) * (1.f / k));
        }
    }
    if (first) return Sc<S>::c(sc.empty_de);
    for (int i = 0; i < sc.carve_count; ++i) {
        const float* c = sc.carves[i];
        S cut = len3(mk3(p.x - Sc<S>::c(c[0]), p.y - Sc<S>::c(c[1]), p.z - Sc<S>::c(c[2]))) - Sc<S>::c(c[3]);
        float k = 0.2f * c[3];
        if (k <= 0.0001f) d = smax_(-cut, d);
        else {
            float hv = fclamp(0.5f - 0.5f * (vv(d) + vv(cut)) / k, 0.f, 1.f);
            S h = (hv > 0.f && hv < 1.f) ? Sc<S>::c(0.5f) - (d + cut) * (0.5f / k) : Sc<S>::c(hv);
            d = d * (Sc<S>::c(1.f) - h) - cut * h + h * (Sc<S>::c(1.f) - h) * k;
        }
    }
    return d;
}

}  // namespace nd
