// nd_formulas.hpp - the formula plug-ins beyond the four tuned ones.
//
// One struct per formulas/*.wgsl file (nd_builtin_formula 4 and up): de_step as step<S>(),
// de_finalize as fin<S>(), over the scalar type S of nd_de.hpp (float for the march,
// shadow and trap; DualN<N> for analytic normals). de_iterations is resolved on the host
// (DevInst::iters), and so are the formula rotation (rotate_xyz_deg of the rotation
// @params: DevInst::rot_on / rot_cs) and the cos / sin of a twist @param (h[14], h[15]).
// Parameters are read with prm<S>(in, k) = FractalInstance.params[k], @param order.
//
// Included by nd_de.hpp after the shared helpers, before the formula wrapper.
#pragma once

#include "nd_hip.h"

namespace nd {

// template.wgsl:59-68
constexpr float K_PI = 3.14159265f;
constexpr float K_TAU = 6.28318530718f;
constexpr float K_EPS = 1e-4f;
constexpr float K_EPS_FINE = 1e-6f;
constexpr float K_EPS_TINY = 1e-12f;
constexpr float K_DEG = 3.14159265358979323846f / 180.0f;  // WGSL radians()
constexpr float K_INFINITE = 1e30f;

// ----------------------------------------------------------------------------- scalar helpers
template <class S> NDD S C(float x) { return Sc<S>::c(x); }
template <class S, class I> NDD S P(const I& in, int k) { return prm<S>(in, k); }

NDD float sfloor(float a) { return floorf(a); }
template <int N> NDD DualN<N> sfloor(const DualN<N>& a) { return dconst<N>(floorf(a.v)); }
NDD float sexp(float a) { return expf(a); }
template <int N> NDD DualN<N> sexp(const DualN<N>& a) { float e = expf(a.v); return dchain(a, e, e); }
NDD float ssign(float a) { return a > 0.f ? 1.f : (a < 0.f ? -1.f : 0.f); }  // WGSL sign
template <int N> NDD DualN<N> ssign(const DualN<N>& a) { return dconst<N>(ssign(a.v)); }
template <class S> NDD S ssin(const S& a) { S s, c; ssincos(a, &s, &c); return s; }
template <class S> NDD S scos(const S& a) { S s, c; ssincos(a, &s, &c); return c; }
NDD float satan(float a) { return atanf(a); }
template <int N> NDD DualN<N> satan(const DualN<N>& a) { return dchain(a, atanf(a.v), 1.f / (1.f + a.v * a.v)); }
NDD float stan(float a) { return tanf(a); }
template <int N> NDD DualN<N> stan(const DualN<N>& a) {
    float t = tanf(a.v);
    return dchain(a, t, 1.f + t * t); }
// pow(x, y) of WGSL for x > 0
template <class S> NDD S spowc(const S& x, float y) { return spow(x, C<S>(y)); }
template <class S> NDD S smix(const S& a, const S& b, const S& t) { return a * (C<S>(1.f) - t) + b * t; }

// The float bits of a value whose low bits carry a level counter (the packed-carry formulas).
template <class S> NDD uint32_t sbits(const S& x) { return __float_as_uint(vv(x)); }
NDD float with_value(float, float v) { return v; }
template <int N> NDD DualN<N> with_value(DualN<N> x, float v) { x.v = v; return x; }
// select(m, min_mag, |m| < min_mag) with its low bits replaced by code (radioactive_pack
// and its siblings): the derivative passes through unchanged.
template <class S> NDD S pack_low(const S& m, uint32_t mask, uint32_t code) {
    const float v = vv(m);
    const float safe = fabsf(v) < 1e-30f ? 1e-30f : v;
    return with_value(m, __uint_as_float((__float_as_uint(safe) & ~mask) | code));
}

// ----------------------------------------------------------------------------- vec3 helpers
template <class S> NDD V3<S> vneg(const V3<S>& a) { return mk3(-a.x, -a.y, -a.z); }
template <class S> NDD V3<S> vabs3(const V3<S>& a) { return mk3(sabs(a.x), sabs(a.y), sabs(a.z)); }
template <class S> NDD V3<S> vmax0(const V3<S>& a) {
    return mk3(smax_(a.x, C<S>(0.f)), smax_(a.y, C<S>(0.f)), smax_(a.z, C<S>(0.f))); }
template <class S> NDD V3<S> vsub1(const V3<S>& a, float s) { return mk3(a.x - s, a.y - s, a.z - s); }
template <class S> NDD S len2(const S& x, const S& y) { return ssqrt(x * x + y * y); }
template <class S, class I> NDD V3<S> vrot(const V3<S>& z, const I& in) { return in.rot_on ? rotate_cs(z, in.rot_cs) : z; }
// rotation about z by the hoisted twist angle (x ca - y sa, x sa + y ca, z)
template <class S, class I> NDD V3<S> vtwist(const V3<S>& z, const I& in) {
    const float ca = in.h[14], sa = in.h[15];
    return mk3(z.x * ca - z.y * sa, z.x * sa + z.y * ca, z.z);
}

// box_fold (template.wgsl:949-951)
template <class S> NDD V3<S> box_fold3(const V3<S>& z, const S& L) {
    const S nL = -L;
    return mk3(sclamp(z.x, nL, L) * 2.f - z.x, sclamp(z.y, nL, L) * 2.f - z.y, sclamp(z.z, nL, L) * 2.f - z.z);
}
// sphere_fold (template.wgsl:959-969), radii squared
template <class S> NDD void sphere_fold3(V3<S>& z, S& dr, const S& m2, const S& F2) {
    const S r2 = dot3(z, z);
    if (vv(r2) < vv(m2)) {
        const S t = sdiv(F2, m2);
        z = z * t; dr = dr * t;
    } else if (vv(r2) < vv(F2)) {
        const S t = sdiv(F2, r2);
        z = z * t; dr = dr * t;
    }
}
// sphere_invert_offset (template.wgsl:975-983)
template <class S> NDD void sphere_invert3(V3<S>& z, S& dr, const V3<S>& ctr, const S& R2) {
    const V3<S> d = z - ctr;
    const S r2 = smax_(dot3(d, d), C<S>(K_EPS_FINE));
    if (vv(r2) < vv(R2)) {
        const S t = sdiv(R2, r2);
        z = ctr + d * t; dr = dr * t;
    }
}
// lattice_wrap (template.wgsl:987-989); round() is half to even
template <class S> NDD V3<S> lattice_wrap3(const V3<S>& p, const S& cell) {
    return mk3(p.x - cell * srint(sdiv(p.x, cell)), p.y - cell * srint(sdiv(p.y, cell)),
               p.z - cell * srint(sdiv(p.z, cell)));
}
// tetra_fold (template.wgsl:1058-1076)
template <class S> NDD V3<S> tetra_fold3(V3<S> v) {
    if (vv(v.x) + vv(v.y) < 0.f) { S t = -v.y; v.y = -v.x; v.x = t; }
    if (vv(v.x) + vv(v.z) < 0.f) { S t = -v.z; v.z = -v.x; v.x = t; }
    if (vv(v.y) + vv(v.z) < 0.f) { S t = -v.z; v.z = -v.y; v.y = t; }
    return v;
}
// the x < y swap that opens the diagonal folds
template <class S> NDD V3<S> swap_xy_desc(V3<S> z) {
    if (vv(z.x) < vv(z.y)) { S t = z.x; z.x = z.y; z.y = t; }
    return z;
}
// Icosahedral mirrors n1 = normalize(-phi, phi - 1, 1), n2 = normalize(1, -phi, phi + 1),
// n3 = (0, 0, -1), each applied when dot(z, n) > 0 (formulas/icosahedralifs_ifs.wgsl:18-34).
constexpr float ICO_N1X = -0.80901699437494742f, ICO_N1Y = 0.30901699437494742f, ICO_N1Z = 0.5f;
constexpr float ICO_N2X = 0.30901699437494742f, ICO_N2Y = -0.5f, ICO_N2Z = 0.80901699437494742f;
template <class S> NDD V3<S> mirror_pos(V3<S> z, float nx, float ny, float nz) {
    const S t = z.x * nx + z.y * ny + z.z * nz;
    if (vv(t) > 0.f) {
        const S t2 = t * 2.f;
        z = mk3(z.x - t2 * nx, z.y - t2 * ny, z.z - t2 * nz);
    }
    return z;
}
template <class S> NDD V3<S> icosa_mirrors(V3<S> z) {
    z = mirror_pos(z, ICO_N1X, ICO_N1Y, ICO_N1Z);
    z = mirror_pos(z, ICO_N2X, ICO_N2Y, ICO_N2Z);
    return mirror_pos(z, 0.f, 0.f, -1.f);
}
// wrap_angle (template.wgsl:1080-1082)
template <class S> NDD S wrap_angle(const S& a, float period) { return a - sfloor(a * (1.f / period)) * period; }
// Polar wedge fold of the xy plane into [0, period / 2] (starkifs / kaleidojulia / prismkifs).
template <class S> NDD V3<S> wedge_fold(const V3<S>& z, float degree) {
    const float period = K_TAU / degree;
    S a = wrap_angle(satan2(z.y, z.x), period);
    if (vv(a) > period * 0.5f) a = C<S>(period) - a;
    const S r = len2(z.x, z.y);
    S sa, ca;
    ssincos(a, &sa, &ca);
    return mk3(r * ca, r * sa, z.z);
}
// z * s - o * (s - 1), the IFS scaling about the offset point
template <class S> NDD V3<S> ifs_scale(const V3<S>& z, const S& s, const V3<S>& o) {
    const S sm1 = s - 1.f;
    return mk3(z.x * s - o.x * sm1, z.y * s - o.y * sm1, z.z * s - o.z * sm1);
}
template <class S, class I> NDD V3<S> p3(const I& in, int k0, int k1, int k2) { return mk3(P<S>(in, k0), P<S>(in, k1), P<S>(in, k2)); }

// Spherical power map of the bulb family: dir(theta * power + tp, phi * power + pp) for the
// unit direction of z (bulb_power_step, template.wgsl:1019-1031, with the phase @params).
template <class S> NDD V3<S> bulb_dir(const V3<S>& z, const S& r, const S& power, float tp, float pp) {
    const S theta = sacos(sclamp(sdiv(z.z, r), C<S>(-1.f), C<S>(1.f))) * power + tp;
    const S phi = satan2(z.y, z.x) * power + pp;
    S st, ct, sp, cp;
    ssincos(theta, &st, &ct);
    ssincos(phi, &sp, &cp);
    return mk3(st * cp, st * sp, ct);
}

// ----------------------------------------------------------------------------- finalizers
template <class S> NDD S fin_scaled(const Carry<S>& c) { return sdiv(len3(c.z), sabs(c.dr)); }  // template.wgsl:993-995
template <class S> NDD S fin_plain(const Carry<S>& c) { return sdiv(len3(c.z), c.dr); }         // template.wgsl:999-1001
template <class S> NDD S fin_unit_box(const Carry<S>& c) { return sdiv(len3(vmax0(vsub1(vabs3(c.z), 1.f))), c.dr); }
template <class S> NDD S prim_div(const S& d, const Carry<S>& c) { return sdiv(d, smax_(sabs(c.dr), C<S>(K_EPS_FINE))); }

// primitive_de_step (template.wgsl:1105-1109)
template <class S> NDD Carry<S> prim_step(const Carry<S>& c, const S& size) {
    const S unit = sdiv(C<S>(1.f), smax_(size, C<S>(K_EPS)));
    const S f = sdiv(unit, smax_(sabs(c.dr), C<S>(K_EPS_FINE)));
    Carry<S> o; o.z = c.z * f; o.dr = unit;
    return o;
}

#define ND_STEP template <class S, class I> static NDD Carry<S> step(const Carry<S>& c, const V3<S>& pos, const I& in)
#define ND_FIN template <class S> static NDD S fin(const Carry<S>& c)
#define ND_RET(zz, dd) do { Carry<S> o_; o_.z = (zz); o_.dr = (dd); return o_; } while (0)

// ============================================================================= primitives
// formulas/sphere_primitive.wgsl:6-12
struct FSphere {
    ND_STEP { (void)pos; return prim_step(c, P<S>(in, 0)); }
    ND_FIN { return prim_div(len3(c.z) - 1.f, c); }
};
// formulas/cube_primitive.wgsl:6-15
struct FCube {
    ND_STEP { (void)pos; return prim_step(c, P<S>(in, 0)); }
    ND_FIN {
        const V3<S> q = vsub1(vabs3(c.z), 1.f);
        const S inside = smin_(smax_(q.x, smax_(q.y, q.z)), C<S>(0.f));
        return prim_div(len3(vmax0(q)) + inside, c);
    }
};
// formulas/cylinder_primitive.wgsl:6-15
struct FCylinder {
    ND_STEP { (void)pos; return prim_step(c, P<S>(in, 0)); }
    ND_FIN {
        const S ex = len2(c.z.x, c.z.z) - 1.f, ey = sabs(c.z.y) - 1.f;
        const S ox = smax_(ex, C<S>(0.f)), oy = smax_(ey, C<S>(0.f));
        return prim_div(len2(ox, oy) + smin_(smax_(ex, ey), C<S>(0.f)), c);
    }
};
// formulas/torus_primitive.wgsl:7-16 (the step reads pos, not the carry)
struct FTorus {
    ND_STEP {
        (void)c;
        const S tube = smax_(P<S>(in, 1), C<S>(K_EPS));
        const S it = sdiv(C<S>(1.f), tube);
        const V3<S> sc = pos * it;
        ND_RET(mk3(len2(sc.x, sc.z) - P<S>(in, 0) * it, sc.y, C<S>(0.f)), it);
    }
    ND_FIN { return prim_div(len2(c.z.x, c.z.y) - 1.f, c); }
};
// formulas/pyramid_primitive.wgsl:2-21; side plane normal (2, 1) / sqrt(5)
struct FPyramid {
    ND_STEP { (void)pos; return prim_step(c, P<S>(in, 0)); }
    ND_FIN {
        const S m = smax_(sabs(c.z.x), sabs(c.z.z));
        const S side = (m * 2.f + c.z.y - 1.f) * 0.44721359549995794f;
        const S base = -(c.z.y + 1.f);
        return prim_div(smax_(side, base), c);
    }
};
// formulas/prism_primitive.wgsl:8-27
struct FPrism {
    ND_STEP {
        (void)pos;
        const Carry<S> q0 = prim_step(c, P<S>(in, 0));
        const S hh = sdiv(P<S>(in, 1), smax_(P<S>(in, 0), C<S>(K_EPS)));
        const float wedge = K_PI / fmaxf(rintf(in.p[2]), 3.f);
        float cw, sw;
        sincosf(wedge, &sw, &cw);
        const S a = wrap_angle(satan2(q0.z.x, q0.z.z), 2.f * wedge) - wedge;
        S sa, ca;
        ssincos(a, &sa, &ca);
        const S l = len2(q0.z.x, q0.z.z);
        S qx = l * ca - cw, qy = l * sabs(sa) - sw;
        qy = qy + sclamp(-qy, C<S>(0.f), C<S>(sw));
        const S edge = vv(qx) > 0.f ? len2(qx, qy) : -len2(qx, qy);
        ND_RET(mk3(edge, sabs(q0.z.y) - hh, C<S>(0.f)), q0.dr);
    }
    ND_FIN {
        const S ex = c.z.x, ey = c.z.y;
        return prim_div(len2(smax_(ex, C<S>(0.f)), smax_(ey, C<S>(0.f))) + smin_(smax_(ex, ey), C<S>(0.f)), c);
    }
};

// ============================================================================= tetrahedral / cubic IFS
// formulas/sierpinski3_ifs.wgsl:13-25
struct FSierpinski3 {
    ND_STEP {
        (void)pos;
        const S s = P<S>(in, 0);
        ND_RET(ifs_scale(vrot(tetra_fold3(c.z), in), s, p3<S>(in, 2, 3, 4)), c.dr * s);
    }
    ND_FIN { return fin_plain(c); }
};
// formulas/cubejulia_ifs_julia.wgsl:13-25
struct FCubeJulia {
    ND_STEP {
        (void)pos;
        const S s = P<S>(in, 0);
        ND_RET(ifs_scale(vrot(sort_desc_abs(c.z), in), s, p3<S>(in, 1, 2, 3)), c.dr * s);
    }
    ND_FIN { return fin_plain(c); }
};
// formulas/juliakifs_ifs_julia.wgsl:13-25
struct FJuliaKifs {
    ND_STEP {
        (void)pos;
        const S s = P<S>(in, 0);
        ND_RET(ifs_scale(vrot(tetra_fold3(c.z), in), s, p3<S>(in, 1, 2, 3)), c.dr * s);
    }
    ND_FIN { return fin_plain(c); }
};
// formulas/diagonaltetra_ifs_fold.wgsl:13-33
struct FDiagonalTetra {
    ND_STEP {
        (void)pos;
        const S s = P<S>(in, 0);
        ND_RET(ifs_scale(vrot(tetra_fold3(swap_xy_desc(c.z)), in), s, p3<S>(in, 2, 3, 4)), c.dr * s);
    }
    ND_FIN { return fin_plain(c); }
};
// formulas/starkifs_ifs_fold.wgsl:10-40
struct FStarKifs {
    ND_STEP {
        (void)pos;
        const S s = P<S>(in, 1);
        const V3<S> z = wedge_fold(c.z, fmaxf(floorf(in.p[0]), 2.f));
        ND_RET(ifs_scale(vrot(tetra_fold3(z), in), s, p3<S>(in, 3, 4, 5)), c.dr * s);
    }
    ND_FIN { return fin_plain(c); }
};
// formulas/kaleidojulia_ifs_fold_julia.wgsl:10-40
struct FKaleidoJulia {
    ND_STEP {
        (void)pos;
        const S s = P<S>(in, 1);
        const V3<S> z = wedge_fold(c.z, fmaxf(floorf(in.p[0]), 2.f));
        ND_RET(ifs_scale(vrot(sort_desc_abs(z), in), s, p3<S>(in, 2, 3, 4)), c.dr * s);
    }
    ND_FIN { return fin_plain(c); }
};
// formulas/bulbtetra_hybrid_bulb_ifs.wgsl:13-38: a bulb step without the trailing +1 on dr
// (skipped once r > 2), then the tetrahedral IFS step
struct FBulbTetra {
    ND_STEP {
        const S power = P<S>(in, 0), s = P<S>(in, 1);
        const S r0 = len3(c.z);
        V3<S> z = c.z;
        S dr = c.dr;
        if (vv(r0) <= 2.f) {
            const V3<S> d = bulb_dir(z, r0, power, in.p[6] * K_DEG, in.p[7] * K_DEG);
            const S rp = spow(r0, power - 1.f);
            const S zr = rp * r0;
            dr = rp * power * dr;
            z = mk3(d.x * zr + pos.x, d.y * zr + pos.y, d.z * zr + pos.z);
        }
        ND_RET(ifs_scale(tetra_fold3(z), s, p3<S>(in, 3, 4, 5)), dr * s);
    }
    ND_FIN { return fin_plain(c); }
};
// formulas/hybridkifs_hybrid_ifs.wgsl:13-35
struct FHybridKifs {
    ND_STEP {
        (void)pos;
        const S s = P<S>(in, 0);
        const V3<S> o = p3<S>(in, 2, 3, 4);
        V3<S> z = ifs_scale(vrot(sort_desc_abs(tetra_fold3(c.z)), in), s, o);
        const S oz = o.z * (s - 1.f);
        if (vv(z.z) < -0.5f * vv(oz)) z.z = z.z + oz;
        ND_RET(z, c.dr * s);
    }
    ND_FIN { return fin_unit_box(c); }
};
// formulas/kaleidoscopicifs_ifs_twist.wgsl:13-37
struct FKaleidoscopicIfs {
    ND_STEP {
        (void)pos;
        const S s = P<S>(in, 0);
        const V3<S> o = p3<S>(in, 3, 4, 5);
        V3<S> z = ifs_scale(sort_desc_abs(c.z), s, o);
        const S oz = o.z * (s - 1.f);
        if (vv(z.z) < -0.5f * vv(oz)) z.z = z.z + oz;
        ND_RET(vrot(vtwist(z, in), in), c.dr * s);
    }
    ND_FIN { return fin_unit_box(c); }
};

// ============================================================================= icosahedral IFS
// formulas/icosahedralifs_ifs.wgsl:13-44
struct FIcosahedralIfs {
    ND_STEP {
        (void)pos;
        const S s = P<S>(in, 0);
        const V3<S> o = p3<S>(in, 1, 2, 4);
        const V3<S> z = vrot(icosa_mirrors(vabs3(c.z)), in);
        ND_RET((z - o) * s + o, c.dr * s);
    }
    ND_FIN { return fin_plain(c); }
};
// formulas/icosahedraljulia_ifs_julia.wgsl:16-48
struct FIcosahedralJulia {
    ND_STEP {
        (void)pos;
        const S s = P<S>(in, 0);
        const V3<S> o = p3<S>(in, 1, 2, 7);
        const V3<S> z = vrot(icosa_mirrors(vabs3(c.z)), in);
        ND_RET((z - o) * s + o + p3<S>(in, 3, 4, 5), c.dr * s);
    }
    ND_FIN { return fin_plain(c); }
};
// formulas/icosahedraltwist_ifs_twist.wgsl:13-49
struct FIcosahedralTwist {
    ND_STEP {
        (void)pos;
        const S s = P<S>(in, 0);
        const V3<S> o = p3<S>(in, 1, 2, 5);
        const V3<S> z = (icosa_mirrors(vabs3(c.z)) - o) * s + o;
        ND_RET(vrot(vtwist(z, in), in), c.dr * s);
    }
    ND_FIN { return fin_plain(c); }
};
// formulas/diagonalicosa_ifs_fold.wgsl:13-51
struct FDiagonalIcosa {
    ND_STEP {
        (void)pos;
        const S s = P<S>(in, 0);
        const V3<S> o = p3<S>(in, 1, 2, 4);
        const V3<S> z = vrot(icosa_mirrors(swap_xy_desc(vabs3(c.z))), in);
        ND_RET((z - o) * s + o, c.dr * s);
    }
    ND_FIN { return fin_plain(c); }
};

// ============================================================================= box-fold family
// z * scale + add, dr * |scale| + dr_add: the tail shared by the Mandelbox relatives
template <class S> NDD Carry<S> box_tail(const V3<S>& z, const S& dr, const S& scale, const V3<S>& add, const S& dr_add) {
    Carry<S> o;
    o.z = mk3(z.x * scale + add.x, z.y * scale + add.y, z.z * scale + add.z);
    o.dr = dr * sabs(scale) + dr_add;
    return o;
}
// pos * cs + julia, max(|cs|, 1)
template <class S, class I> NDD V3<S> cj_add(const V3<S>& pos, const I& in, int cs) {
    const S k = P<S>(in, cs);
    return mk3(pos.x * k + P<S>(in, cs + 1), pos.y * k + P<S>(in, cs + 2), pos.z * k + P<S>(in, cs + 3));
}
template <class S, class I> NDD S cj_dr(const I& in, int cs) { return C<S>(fmaxf(fabsf(in.p[cs]), 1.f)); }
template <class S> NDD S sq(const S& x) { return x * x; }

// formulas/juliabox_box_julia.wgsl:16-37
struct FJuliaBox {
    ND_STEP {
        (void)pos;
        V3<S> z = box_fold3(c.z, P<S>(in, 3));
        S dr = c.dr;
        sphere_fold3(z, dr, sq(P<S>(in, 1)), sq(P<S>(in, 2)));
        return box_tail(vrot(z, in), dr, P<S>(in, 0), p3<S>(in, 4, 5, 6), C<S>(1.f));
    }
    ND_FIN { return fin_scaled(c); }
};
// formulas/amazingsurf_box_fold.wgsl:17-38 (TAIL 1: |z.z|), formulas/mandelboxfoam_box_fold.wgsl
// (TAIL 2: |z.x| and |z.z|), formulas/mandelboxsort_box_fold.wgsl (PRE 1: sort_desc_abs
// first), formulas/mandelboxdiag_box_fold.wgsl (PRE 2: the x < y swap first); slots
// [scale, minR, fixR, limit, iters, rx, ry, rz, cs, jx, jy, jz]
template <int PRE, int TAIL>
struct FBoxVariant {
    ND_STEP {
        V3<S> z = c.z;
        if (PRE == 1) z = sort_desc_abs(z);
        if (PRE == 2) z = swap_xy_desc(z);
        z = box_fold3(z, P<S>(in, 3));
        S dr = c.dr;
        sphere_fold3(z, dr, sq(P<S>(in, 1)), sq(P<S>(in, 2)));
        Carry<S> o = box_tail(vrot(z, in), dr, P<S>(in, 0), cj_add(pos, in, 8), cj_dr<S>(in, 8));
        if (TAIL == 2) o.z.x = sabs(o.z.x);
        if (TAIL >= 1) o.z.z = sabs(o.z.z);
        return o;
    }
    ND_FIN { return fin_scaled(c); }
};
// formulas/mandelboxinv_box_fold.wgsl:18-46
struct FMandelboxInv {
    ND_STEP {
        const S tin = sdiv(sq(P<S>(in, 0)), smax_(dot3(c.z, c.z), C<S>(K_EPS_FINE)));
        V3<S> z = box_fold3(c.z * tin, P<S>(in, 4));
        S dr = c.dr * tin;
        sphere_fold3(z, dr, sq(P<S>(in, 2)), sq(P<S>(in, 3)));
        return box_tail(vrot(z, in), dr, P<S>(in, 1), cj_add(pos, in, 9), cj_dr<S>(in, 9));
    }
    ND_FIN { return fin_scaled(c); }
};
// formulas/mandelboxtwist_box_twist.wgsl:17-42
struct FMandelboxTwist {
    ND_STEP {
        V3<S> z = box_fold3(c.z, P<S>(in, 3));
        S dr = c.dr;
        sphere_fold3(z, dr, sq(P<S>(in, 1)), sq(P<S>(in, 2)));
        return box_tail(vrot(vtwist(z, in), in), dr, P<S>(in, 0), cj_add(pos, in, 8), cj_dr<S>(in, 8));
    }
    ND_FIN { return fin_scaled(c); }
};
// formulas/prismkifs_box_fold.wgsl:14-52
struct FPrismKifs {
    ND_STEP {
        V3<S> z = box_fold3(wedge_fold(c.z, fmaxf(floorf(in.p[0]), 2.f)), P<S>(in, 4));
        S dr = c.dr;
        sphere_fold3(z, dr, sq(P<S>(in, 2)), sq(P<S>(in, 3)));
        return box_tail(vrot(z, in), dr, P<S>(in, 1), cj_add(pos, in, 9), cj_dr<S>(in, 9));
    }
    ND_FIN { return fin_scaled(c); }
};
// formulas/buffalo_box.wgsl:16-36 (ABS) and formulas/apollonian_kleinian.wgsl:16-36 (no
// abs); slots [scale, minR, fixR, iters, rx, ry, rz, cs, jx, jy, jz]
template <bool ABS>
struct FSphereFoldMandel {
    ND_STEP {
        V3<S> z = ABS ? vabs3(c.z) : c.z;
        S dr = c.dr;
        sphere_fold3(z, dr, sq(P<S>(in, 1)), sq(P<S>(in, 2)));
        return box_tail(vrot(z, in), dr, P<S>(in, 0), cj_add(pos, in, 7), cj_dr<S>(in, 7));
    }
    ND_FIN { return fin_scaled(c); }
};
// formulas/buffalojulia_box_julia.wgsl:15-36 (ABS) and
// formulas/apollonianjulia_kleinian_julia.wgsl:15-36; slots [scale, minR, fixR, jx, jy, jz,
// iters, rx, ry, rz]
template <bool ABS>
struct FSphereFoldJulia {
    ND_STEP {
        (void)pos;
        V3<S> z = ABS ? vabs3(c.z) : c.z;
        S dr = c.dr;
        sphere_fold3(z, dr, sq(P<S>(in, 1)), sq(P<S>(in, 2)));
        return box_tail(vrot(z, in), dr, P<S>(in, 0), p3<S>(in, 3, 4, 5), C<S>(1.f));
    }
    ND_FIN { return fin_scaled(c); }
};
// formulas/buffalotwist_box_twist.wgsl:16-41
struct FBuffaloTwist {
    ND_STEP {
        V3<S> z = vabs3(c.z);
        S dr = c.dr;
        sphere_fold3(z, dr, sq(P<S>(in, 1)), sq(P<S>(in, 2)));
        return box_tail(vrot(vtwist(z, in), in), dr, P<S>(in, 0), cj_add(pos, in, 7), cj_dr<S>(in, 7));
    }
    ND_FIN { return fin_scaled(c); }
};

// ============================================================================= Kleinian relatives
// formulas/kleinianjulia_kleinian_julia.wgsl:14-34
struct FKleinianJulia {
    ND_STEP {
        (void)pos;
        V3<S> z = lattice_wrap3(c.z, P<S>(in, 8));
        S dr = c.dr;
        sphere_invert3(z, dr, p3<S>(in, 0, 1, 2), sq(P<S>(in, 3)));
        ND_RET(z + p3<S>(in, 4, 5, 6), dr + 1.f);
    }
    ND_FIN { return fin_scaled(c); }
};
// formulas/pseudokleinian_kleinian.wgsl:15-34
struct FPseudoKleinian {
    ND_STEP {
        V3<S> z = lattice_wrap3(c.z, P<S>(in, 4));
        S dr = c.dr;
        sphere_invert3(z, dr, p3<S>(in, 0, 1, 2), sq(P<S>(in, 3)));
        ND_RET(z + cj_add(pos, in, 6), dr + cj_dr<S>(in, 6));
    }
    ND_FIN { return fin_scaled(c); }
};

// ============================================================================= bulb family
// formulas/juliabulb_bulb_julia.wgsl:12-30
struct FJuliaBulb {
    ND_STEP {
        (void)pos;
        const S r = len3(c.z);
        if (vv(r) > 2.f) return c;
        const S power = P<S>(in, 0);
        const V3<S> d = bulb_dir(c.z, r, power, in.p[5] * K_DEG, in.p[6] * K_DEG);
        const S rp = spow(r, power - 1.f);
        const S zr = rp * r;
        ND_RET(d * zr + p3<S>(in, 2, 3, 4), rp * power * c.dr + 1.f);
    }
    ND_FIN { return bulb_finalize(c); }
};
// formulas/mandelbulbmirror_bulb_fold.wgsl:13-32
struct FMandelbulbMirror {
    ND_STEP {
        const V3<S> zin = mk3(c.z.x, -c.z.y, c.z.z);
        const S r = len3(zin);
        if (vv(r) > 2.f) return c;
        const S power = P<S>(in, 0);
        const V3<S> d = bulb_dir(zin, r, power, in.p[2] * K_DEG, in.p[3] * K_DEG);
        const S zr = spow(r, power);
        const S dr = spow(r, power - 1.f) * power * c.dr + cj_dr<S>(in, 4);
        ND_RET(d * zr + cj_add(pos, in, 4), dr);
    }
    ND_FIN { return bulb_finalize(c); }
};
// formulas/mandelbulbtwist_bulb_twist.wgsl:14-37
struct FMandelbulbTwist {
    ND_STEP {
        const S r = len3(c.z);
        if (vv(r) > 2.f) return c;
        const S power = P<S>(in, 0);
        const V3<S> d = bulb_dir(c.z, r, power, in.p[3] * K_DEG, in.p[4] * K_DEG);
        const S zr = spow(r, power);
        const S dr = spow(r, power - 1.f) * power * c.dr + cj_dr<S>(in, 5);
        ND_RET(vtwist(d * zr, in) + cj_add(pos, in, 5), dr);
    }
    ND_FIN { return bulb_finalize(c); }
};
// formulas/mandelbulbdual_hybrid_bulb.wgsl:14-45
struct FMandelbulbDual {
    ND_STEP {
        const S r1 = len3(c.z);
        if (vv(r1) > 2.f) return c;
        const float tp = in.p[3] * K_DEG, pp = in.p[4] * K_DEG;
        const S pw1 = P<S>(in, 0), pw2 = P<S>(in, 2);
        V3<S> z = bulb_dir(c.z, r1, pw1, tp, pp) * spow(r1, pw1);
        S dr = spow(r1, pw1 - 1.f) * pw1 * c.dr;
        const S r2 = len3(z);
        const V3<S> d2 = bulb_dir(z, r2, pw2, tp, pp);
        dr = spow(r2, pw2 - 1.f) * pw2 * dr;
        ND_RET(d2 * spow(r2, pw2) + cj_add(pos, in, 5), dr + cj_dr<S>(in, 5));
    }
    ND_FIN { return bulb_finalize(c); }
};
// formulas/quatjulia_bulb_julia.wgsl:7-38: the quaternion Julia set z^2 + c in the plane
// spanned by 1 and the imaginary axis of c; the sign of dr marks the first iteration.
struct FQuatJulia {
    ND_STEP {
        (void)pos;
        const float cx = in.p[1], cy = in.p[2], cz = in.p[3];
        const float cl = sqrtf(cx * cx + cy * cy + cz * cz);
        V3<S> s = c.z;
        S dr = -c.dr;
        if (vv(c.dr) > 0.f) {
            float ax = 1.f, ay = 0.f, az = 0.f;
            if (cl > K_EPS_FINE) { ax = cx / cl; ay = cy / cl; az = cz / cl; }
            const V3<S> v = mk3(c.z.y, c.z.z, P<S>(in, 4));
            const S along = v.x * ax + v.y * ay + v.z * az;
            const V3<S> w = mk3(v.x - along * ax, v.y - along * ay, v.z - along * az);
            s = mk3(c.z.x, along, len3(w));
            dr = c.dr;
        }
        const S r2 = dot3(s, s);
        if (vv(r2) > 256.f) ND_RET(s, -dr);
        const V3<S> q = mk3(s.x * s.x - s.y * s.y - s.z * s.z + P<S>(in, 0), s.x * s.y * 2.f + cl, s.x * s.z * 2.f);
        ND_RET(q, ssqrt(r2) * dr * -2.f);
    }
    ND_FIN {
        const S r = len3(c.z);
        return sdiv(sln(smax_(r, C<S>(K_EPS_FINE))) * r * 0.5f, smax_(sabs(c.dr), C<S>(K_EPS_FINE)));
    }
};
// formulas/bulbbox_hybrid_bulb_box.wgsl:20-52
struct FBulbBox {
    ND_STEP {
        const S power = P<S>(in, 0);
        const S r0 = len3(c.z);
        V3<S> z = c.z;
        S dr = c.dr;
        if (vv(r0) <= 2.f) {
            const V3<S> d = bulb_dir(z, r0, power, in.p[6] * K_DEG, in.p[7] * K_DEG);
            dr = spow(r0, power - 1.f) * power * dr;
            z = d * spow(r0, power);
        }
        z = box_fold3(z, P<S>(in, 4));
        sphere_fold3(z, dr, sq(P<S>(in, 2)), sq(P<S>(in, 3)));
        return box_tail(vrot(z, in), dr, P<S>(in, 1), cj_add(pos, in, 11), cj_dr<S>(in, 11));
    }
    ND_FIN { return bulb_finalize(c); }
};

// ============================================================================= other folds
// formulas/mandalay_box_fold.wgsl:19-74
template <class S> NDD S mandalay_fold(V3<S> q, const S& fo, const S& g) {
    if (vv(q.z) > vv(q.y)) { S t = q.y; q.y = q.z; q.z = t; }
    const S vx = q.x - fo * 2.f;
    const S vy = q.y - fo * 4.f;
    const S v = smax_(sabs(vx + fo) - fo, vy);
    const S v1 = smax_(vx - g, q.y);
    return smin_(smin_(v, v1), q.x);
}
struct FMandalay {
    ND_STEP {
        const V3<S> cc = pos * 4.f + p3<S>(in, 12, 13, 14);
        V3<S> z = c.z;
        S dr = c.dr;
        if (vv(dr) == 1.f) { z = cc; dr = C<S>(4.f); }
        if (vv(dot3(z, z)) > 1e20f) ND_RET(z, dr);
        const S L = P<S>(in, 8);
        z = mk3(z.x - sclamp(z.x, -L, L) * 2.f, z.y - sclamp(z.y, -L, L) * 2.f, z.z - sclamp(z.z, -L, L) * 2.f);
        const V3<S> sg = mk3(ssign(z.x), ssign(z.y), ssign(z.z));
        const V3<S> a = vabs3(z);
        const S gap = P<S>(in, 5);
        z = mk3(mandalay_fold(a, P<S>(in, 2), P<S>(in, 9) * gap) * sg.x,
                mandalay_fold(mk3(a.y, a.z, a.x), P<S>(in, 3), P<S>(in, 10) * gap) * sg.y,
                mandalay_fold(mk3(a.z, a.x, a.y), P<S>(in, 4), P<S>(in, 11) * gap) * sg.z);
        const float mr = in.p[1];
        const S t = sclamp(sdiv(C<S>(1.f), smax_(dot3(z, z), C<S>(K_EPS_TINY))), C<S>(1.f), C<S>(1.f / (mr * mr)));
        z = z * t;
        dr = dr * t;
        const S s = P<S>(in, 0);
        z = mk3(z.x * s + cc.x, z.y * s + cc.y, (z.z * s + cc.z) * P<S>(in, 6));
        ND_RET(z, dr * sabs(s) + 4.f);
    }
    ND_FIN { return sdiv(len3(c.z) - 30.f, sabs(c.dr)); }
};
// formulas/wobblekifs_ifs_twist_julia.wgsl:19-65
template <class S> NDD void wob_rot(S& x, S& y, const S& a) {
    S sa, ca;
    ssincos(a, &sa, &ca);
    const S nx = ca * x + sa * y, ny = ca * y - sa * x;
    x = nx; y = ny;
}
struct FWobbleKifs {
    ND_STEP {
        const S s = P<S>(in, 0);
        const float t = in.p[7] * 4.f;
        const S len = len3(pos);
        V3<S> z = c.z;
        if (vv(c.dr) == 1.f) {
            z = mk3(z.y, z.x, z.z);
            wob_rot(z.x, z.y, C<S>(in.p[15] + sinf(t - 0.25f) * in.p[14]));
        }
        const S phx = ssin(len * in.p[12] + t) * in.p[8] + in.p[5];
        const S phy = scos(len * in.p[13] + (t - 0.2f)) * in.p[9] + in.p[6];
        z = vabs3(z) * s + p3<S>(in, 2, 3, 4);
        wob_rot(z.x, z.z, phx * 3.14f + scos(len + t) * in.p[10]);
        wob_rot(z.y, z.z, phy * 3.14f + ssin(len + t) * in.p[11]);
        ND_RET(z, c.dr * s);
    }
    ND_FIN {
        const V3<S> a = vabs3(c.z);
        const S m = smax_(smax_(a.x, a.y), smax_(a.z, C<S>(1e-20f)));
        const V3<S> q = mk3(sdiv(a.x, m), sdiv(a.y, m), sdiv(a.z, m));
        const V3<S> q3 = mk3(q.x * q.x * q.x, q.y * q.y * q.y, q.z * q.z * q.z);
        const S n = spowc(q3.x * q3.x + q3.y * q3.y + q3.z * q3.z, 1.f / 6.f);
        return sdiv(m * n, c.dr) - 0.25f;
    }
};

// ============================================================================= packed-carry formulas
// These keep a running minimum in dr, with an iteration level in its low mantissa bits.
// formulas/radioactive_ifs_carve_twist.wgsl:9-42
struct FRadioactive {
    ND_STEP {
        (void)pos;
        const uint32_t level = sbits(c.dr) & 31u;
        V3<S> v = c.z;
        S sv = c.dr;
        if (level == 0u) sv = v.y - P<S>(in, 5);
        const S i = P<S>(in, 1) * spow(P<S>(in, 4), C<S>((float)level));
        float sn, cs;
        sincosf(in.p[2], &sn, &cs);
        v = mk3(v.x * cs + v.z * sn, v.y, v.z * cs - v.x * sn);
        const S w = i * 2.f;
        const S ic = i * P<S>(in, 3);
        v = mk3(ic - sabs(v.x - w * sfloor(sdiv(v.x, w)) - i), ic - sabs(v.y - w * sfloor(sdiv(v.y, w)) - i),
                ic - sabs(v.z - w * sfloor(sdiv(v.z, w)) - i));
        sv = smax_(sv, smin_(smin_(v.x, v.y), v.z));
        ND_RET(v, pack_low(sv, 31u, min(level + 1u, 31u)));
    }
    ND_FIN { return c.dr; }
};
// formulas/recursivesphere_ifs_carve_twist.wgsl:12-65
template <class S> NDD S smin_poly(const S& a, const S& b, const S& k) {  // smin, template.wgsl:1653-1659
    if (vv(k) <= 0.0001f) return smin_(a, b);
    const S h = sclamp(sdiv(b - a, k) * 0.5f + 0.5f, C<S>(0.f), C<S>(1.f));
    return smix(b, a, h) - k * h * (C<S>(1.f) - h);
}
struct FRecursiveSphere {
    ND_STEP {
        (void)pos;
        const uint32_t level = sbits(c.dr) & 31u;
        const S world = sdiv(P<S>(in, 1), C<S>(5.2f));
        V3<S> z = c.z;
        S best = c.dr;
        if (level == 0u) {
            z = mk3(sdiv(z.x, world), sdiv(z.y, world), sdiv(z.z, world));
            best = C<S>(K_INFINITE);
        }
        const S shrink = spow(P<S>(in, 5), C<S>((float)level));
        const S radius = sdiv(C<S>(5.2f), P<S>(in, 8)) * shrink;
        const S fin = P<S>(in, 6) * shrink;
        S d = len3(z) - radius * P<S>(in, 8);
        V3<S> corners = vabs3(z) + mk3(radius * -0.57735027f, radius * -0.57735027f, radius * -0.57735027f);
        const S hole = len3(corners) - radius * P<S>(in, 2);
        const V3<S> ap = vabs3(z) * 0.7071f;
        S cut = smax_(hole, fin - sabs(ap.x - ap.y));
        cut = smax_(cut, fin - sabs(ap.y - ap.z));
        cut = smax_(cut, fin - sabs(ap.z - ap.x));
        cut = smin_(cut, hole + P<S>(in, 3) * shrink);
        d = -smin_poly(-d, cut, P<S>(in, 7) * shrink);
        best = smin_(best, d * world);
        const S a = sdiv(P<S>(in, 4), shrink);
        S sa, ca;
        ssincos(a, &sa, &ca);
        corners = mk3(ca * corners.x - sa * corners.z, corners.y, sa * corners.x + ca * corners.z);
        z = mk3(corners.x, corners.z, -corners.y);
        if (level + 1u >= (uint32_t)fmaxf(in.p[0], 1.f))
            best = smin_(best, (len3(z) - radius * P<S>(in, 5) * P<S>(in, 8)) * world);
        ND_RET(z, pack_low(best, 31u, min(level + 1u, 31u)));
    }
    ND_FIN { return c.dr; }
};
// formulas/tree_ifs.wgsl:9-75
template <class S> NDD S round_cone(const V3<S>& p, const S& r1, const S& r2, const S& h) {
    const S qx = len2(p.x, p.z), qy = p.y;
    const S b = sdiv(r1 - r2, h);
    const S a = ssqrt(C<S>(1.f) - b * b);
    const S k = qy * a - qx * b;
    if (vv(k) < 0.f) return len2(qx, qy) - r1;
    if (vv(k) > vv(a * h)) return len2(qx, qy - h) - r2;
    return qx * a + qy * b - r1;
}
struct FTree {
    ND_STEP {
        (void)pos;
        const uint32_t level = sbits(c.dr) & 31u;
        const S ratio = P<S>(in, 3), thick = P<S>(in, 5);
        V3<S> z = c.z;
        S m = c.dr;
        S lh = C<S>(1.f);
        if (level == 0u) {
            z = mk3(sdiv(z.x, C<S>(0.6f)), sdiv(z.y + 1.f, C<S>(0.6f)), sdiv(z.z, C<S>(0.6f)));
            lh = P<S>(in, 6);
            m = round_cone(z, thick, thick * ratio, lh) * 0.6f;
        }
        z.y = z.y - lh;
        float stw, ctw;
        sincosf(in.p[4] * K_DEG, &stw, &ctw);
        z = mk3(z.x * ctw - z.z * stw, z.y, z.x * stw + z.z * ctw);
        const float sector = K_TAU / fmaxf(rintf(in.p[1]), 1.f);
        S ang = satan2(z.z, z.x);
        ang = sabs(ang - srint(ang * (1.f / sector)) * sector);
        const S rxz = len2(z.x, z.z);
        S sa, ca;
        ssincos(ang, &sa, &ca);
        z = mk3(rxz * ca, z.y, rxz * sa);
        float sth, cth;
        sincosf(in.p[2] * K_DEG, &sth, &cth);
        z = mk3(sdiv(z.x * cth - z.y * sth, ratio), sdiv(z.x * sth + z.y * cth, ratio), sdiv(z.z, ratio));
        const uint32_t next = min(level + 1u, 31u);
        const S wpl = spow(ratio, C<S>((float)next)) * 0.6f;
        m = smin_(m, round_cone(z, thick, thick * ratio, C<S>(1.f)) * wpl);
        if ((int)next >= (int)in.p[0] && in.p[7] > 0.f)
            m = smin_(m, (len3(mk3(z.x, z.y - 1.f, z.z)) - P<S>(in, 7)) * wpl);
        ND_RET(z, pack_low(m, 31u, next));
    }
    ND_FIN { return c.dr; }
};
// formulas/kleinianlimit_kleinian.wgsl:9-76 (7-bit step counter, 127 = done)
struct FKleinianLimit {
    ND_STEP {
        const uint32_t stp = sbits(c.dr) & 127u;
        if (stp == 127u) return c;
        const S a = P<S>(in, 0), b = P<S>(in, 1);
        const S bx = P<S>(in, 2), bz = P<S>(in, 3);
        V3<S> z = c.z;
        S df = c.dr;
        if (stp == 0u) { z.y = z.y + 1.f; df = C<S>(1.f); }
        const S ba = sdiv(b, a);
        z.x = z.x + ba * z.y;
        z.x = z.x - bx * 2.f * sfloor(sdiv(z.x + bx, bx * 2.f));
        z.z = z.z - bz * 2.f * sfloor(sdiv(z.z + bz, bz * 2.f));
        z.x = z.x - ba * z.y;
        const S xs = z.x + b * 0.5f;
        const S sep = a * 0.5f + (a * 2.f - 1.95f) * (ssign(vv(b)) * 0.25f) * ssign(vv(xs)) *
                      (C<S>(1.f) - sexp(-(C<S>(7.2f) - (C<S>(1.95f) - a) * 15.f) * sabs(xs)));
        if (vv(z.y) >= vv(sep)) z = mk3(-b - z.x, a - z.y, -z.z);
        const S ir = sdiv(C<S>(1.f), smax_(dot3(z, z), C<S>(1e-20f)));
        z = z * -ir;
        z.x = -b - z.x;
        z.y = a + z.y;
        df = df * ir;
        if ((int)stp + 1 >= (int)in.p[5]) {
            const S y = smin_(z.y, a - z.y);
            const S de = sdiv(smin_(y, C<S>(0.3f)), smax_(df, C<S>(2.f)));
            const S cut = pos.y + 1.f - P<S>(in, 4) * a;
            ND_RET(z, pack_low(smax_(de, cut), 127u, 127u));
        }
        ND_RET(z, pack_low(df, 127u, stp + 1u));
    }
    ND_FIN { return c.dr; }
};
// formulas/fbmsdf_primitive_carve.wgsl:11-107
// The hash's product reaches ~2.5e5 before its fract, so it is computed exactly as written
// in FP32: every product and sum rounded on its own (mul_rn / add_rn, no fused multiply-add).
NDD float fbm_fract(float x) { return sub_rn(x, floorf(x)); }
NDD float fbm_hash(float x, float y, float z) {
    const float qx = mul_rn(17.f, fbm_fract(add_rn(mul_rn(x, 0.3183099f), 0.11f)));
    const float qy = mul_rn(17.f, fbm_fract(add_rn(mul_rn(y, 0.3183099f), 0.17f)));
    const float qz = mul_rn(17.f, fbm_fract(add_rn(mul_rn(z, 0.3183099f), 0.13f)));
    return fbm_fract(mul_rn(mul_rn(mul_rn(qx, qy), qz), add_rn(add_rn(qx, qy), qz)));
}
template <class S> NDD S fbm_lattice(const V3<S>& p, float radius) {
    const float ix = floorf(vv(p.x)), iy = floorf(vv(p.y)), iz = floorf(vv(p.z));
    const V3<S> f = mk3(p.x - ix, p.y - iy, p.z - iz);
    S d = C<S>(1e10f);
    for (int k = 0; k < 8; ++k) {
        const float cx = (float)(k & 1), cy = (float)((k >> 1) & 1), cz = (float)((k >> 2) & 1);
        const float r = fbm_hash(ix + cx, iy + cy, iz + cz);
        d = smin_(d, len3(mk3(f.x - cx, f.y - cy, f.z - cz)) - r * r * radius);
    }
    return d;
}
template <class S> NDD S fbm_simplex(const V3<S>& p, float radius) {
    const float k1 = 0.333333333f, k2 = 0.166666667f;
    const float s = (vv(p.x) + vv(p.y) + vv(p.z)) * k1;
    const float ix = floorf(vv(p.x) + s), iy = floorf(vv(p.y) + s), iz = floorf(vv(p.z) + s);
    const float u = (ix + iy + iz) * k2;
    const V3<S> d0 = mk3(p.x - (ix - u), p.y - (iy - u), p.z - (iz - u));
    // e = step(d0.yzx, d0)
    const float ex = vv(d0.x) >= vv(d0.y) ? 1.f : 0.f, ey = vv(d0.y) >= vv(d0.z) ? 1.f : 0.f,
                ez = vv(d0.z) >= vv(d0.x) ? 1.f : 0.f;
    const float i1x = ex * (1.f - ez), i1y = ey * (1.f - ex), i1z = ez * (1.f - ey);
    const float i2x = 1.f - ez * (1.f - ex), i2y = 1.f - ex * (1.f - ey), i2z = 1.f - ey * (1.f - ez);
    const V3<S> d1 = mk3(d0.x - (i1x - k2), d0.y - (i1y - k2), d0.z - (i1z - k2));
    const V3<S> d2 = mk3(d0.x - (i2x - 2.f * k2), d0.y - (i2y - 2.f * k2), d0.z - (i2z - 2.f * k2));
    const float k3 = 1.f - 3.f * k2;
    const V3<S> d3 = mk3(d0.x - k3, d0.y - k3, d0.z - k3);
    const float r0 = fbm_hash(ix, iy, iz);
    const float r1 = fbm_hash(ix + i1x, iy + i1y, iz + i1z);
    const float r2 = fbm_hash(ix + i2x, iy + i2y, iz + i2z);
    const float r3 = fbm_hash(ix + 1.f, iy + 1.f, iz + 1.f);
    const float rs = radius * (0.55f / 0.7f);
    return smin_(smin_(len3(d0) - r0 * r0 * rs, len3(d1) - r1 * r1 * rs),
                 smin_(len3(d2) - r2 * r2 * rs, len3(d3) - r3 * r3 * rs));
}
struct FFbmSdf {
    ND_STEP {
        (void)pos;
        const uint32_t level = sbits(c.dr) & 31u;
        V3<S> z = c.z;
        S d = c.dr;
        if (level == 0u) {
            const V3<S> q = vsub1(vabs3(z), in.p[1]);
            d = smin_(smax_(q.x, smax_(q.y, q.z)), C<S>(0.f)) + len3(vmax0(q));
            const float o = 0.5f + in.p[7];
            z = mk3(z.x + o, z.y + o, z.z + o);
        }
        const S sc = spow(P<S>(in, 4), C<S>((float)level));
        const S base = in.p[6] > 0.5f ? fbm_simplex(z, in.p[2]) : fbm_lattice(z, in.p[2]);
        const S n = sc * base;
        const S k = smax_(P<S>(in, 3) * sc, C<S>(K_EPS_FINE));
        const S h = smax_(k - sabs(d + n), C<S>(0.f));
        d = smax_(d, -n) + sdiv(h * h * 0.25f, k);
        const S l = P<S>(in, 5);
        z = mk3((z.y * -0.80f - z.z * 0.60f) * l, (z.x * 0.80f + z.y * 0.36f - z.z * 0.48f) * l,
                (z.x * 0.60f - z.y * 0.48f + z.z * 0.64f) * l);
        ND_RET(z, pack_low(d, 31u, min(level + 1u, 31u)));
    }
    ND_FIN { return c.dr; }
};

// ============================================================================= procedural
// formulas/seascape_primitive.wgsl:14-108: a height field of summed choppy wave octaves,
// bounded by a slab; the integer hash (LCG_MULTIPLIER / LCG_INCREMENT, template.wgsl:76-81)
// of the lattice cell.
NDD float sea_hash(int cx, int cy) {
    uint32_t vx = (uint32_t)cx * 1664525u + 1013904223u, vy = (uint32_t)cy * 1664525u + 1013904223u;
    vx += vy * 1664525u;
    vy += vx * 1664525u;
    vx ^= vx >> 16;
    vy ^= vy >> 16;
    vx += vy * 1664525u;
    vy += vx * 1664525u;
    vx ^= vx >> 16;
    vy ^= vy >> 16;
    return (float)(vx ^ vy) * 2.3283064365386963e-10f;
}
template <class S> NDD S sea_noise(const S& px, const S& py) {
    const float ix = floorf(vv(px)), iy = floorf(vv(py));
    const S fx = px - ix, fy = py - iy;
    const int cx = (int)ix, cy = (int)iy;
    const S ux = fx * fx * (C<S>(3.f) - fx * 2.f), uy = fy * fy * (C<S>(3.f) - fy * 2.f);
    const S a = smix(C<S>(sea_hash(cx, cy)), C<S>(sea_hash(cx + 1, cy)), ux);
    const S b = smix(C<S>(sea_hash(cx, cy + 1)), C<S>(sea_hash(cx + 1, cy + 1)), ux);
    return smix(a, b, uy) * 2.f - 1.f;
}
template <class S> NDD S sea_octave(S ux, S uy, float choppy) {
    const S n = sea_noise(ux, uy);
    ux = ux + n;
    uy = uy + n;
    S sx, cx, sy, cy;
    ssincos(ux, &sx, &cx);
    ssincos(uy, &sy, &cy);
    S wx = C<S>(1.f) - sabs(sx), wy = C<S>(1.f) - sabs(sy);
    wx = smix(wx, sabs(cx), wx);
    wy = smix(wy, sabs(cy), wy);
    const S crest = spowc(smax_(wx * wy, C<S>(0.f)), 0.65f);
    return spowc(smax_(C<S>(1.f) - crest, C<S>(0.f)), choppy);
}
struct FSeascape {
    ND_STEP {
        (void)pos;
        const int octaves = min(max((int)in.p[0], 1), 8);
        const float size = fmaxf(in.p[1], K_EPS);
        const S iq = sdiv(C<S>(1.f), smax_(sabs(c.dr), C<S>(K_EPS_FINE)) * size);
        const V3<S> q = c.z * iq;
        const float sea = fmaxf(in.p[2], 0.f), time = in.p[5], extent = fmaxf(in.p[6], K_EPS);
        const float depth = fmaxf(in.p[7], 0.f), fgain = fmaxf(in.p[8], 0.f);
        const float again = fminf(fmaxf(in.p[9], 0.f), 0.95f), relax = fminf(fmaxf(in.p[10], 0.f), 1.f);
        const float hmax = 2.f * sea * (1.f - powf(again, (float)octaves)) / (1.f - again);
        float reach = hmax;
        S above = q.y - reach, below = -q.y;
        S ux = q.x * 0.75f, uy = q.z;
        float freq = in.p[4], amp = sea, choppy = fmaxf(in.p[3], 1.f), stretch = 1.f, slope = 0.f;
        S h = C<S>(0.f);
        for (int i = 0; i < octaves; ++i) {
            const S d = sea_octave((ux + time) * freq, (uy + time) * freq, choppy) +
                        sea_octave((ux - time) * freq, (uy - time) * freq, choppy);
            h = h + d * amp;
            reach -= 2.f * amp;
            slope += 2.f * amp * freq * stretch * (3.5f + 1.05f * choppy);
            const float is = 1.f / sqrtf(1.f + slope * slope);
            above = smax_(above, (q.y - h - fmaxf(reach, 0.f)) * is);
            below = smax_(below, (h - q.y) * is);
            const S nx = ux * 1.6f + uy * 1.2f, ny = uy * 1.6f - ux * 1.2f;  // uv * mat2x2f(1.6, 1.2, -1.2, 1.6)
            ux = nx;
            uy = ny;
            freq *= fgain;
            amp *= again;
            stretch *= 2.f;
            choppy = choppy * (1.f - relax) + relax;
        }
        const S wave = vv(q.y) >= vv(h) ? above : -below;
        const S ex = sabs(q.x) - extent, ey = smax_(-q.y - depth, q.y - hmax), ez = sabs(q.z) - extent;
        const S bounds = len3(vmax0(mk3(ex, ey, ez))) + smin_(smax_(ex, smax_(ey, ez)), C<S>(0.f));
        const S d = smax_(wave, bounds);
        const S elev = sclamp(h * (1.f / fmaxf(hmax, K_EPS)), C<S>(0.f), C<S>(1.f));
        ND_RET(mk3(C<S>(0.f), elev, C<S>(0.f)), d * size);
    }
    ND_FIN { return c.dr; }
};

// formulas/dragon_ifs_twist.wgsl:7-123: the nearest point of a twisted dragon (or Levy)
// curve by a depth-first descent over its two affine maps, pruned by bounding spheres.
struct M3 { float m[3][3]; };  // m[col][row], WGSL mat3x3f columns
NDD M3 m3_rz(float a) { float c = cosf(a), s = sinf(a); return M3{{{c, s, 0.f}, {-s, c, 0.f}, {0.f, 0.f, 1.f}}}; }
NDD M3 m3_rx(float a) { float c = cosf(a), s = sinf(a); return M3{{{1.f, 0.f, 0.f}, {0.f, c, s}, {0.f, -s, c}}}; }
NDD M3 m3_mul(const M3& a, const M3& b) {
    M3 r;
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) r.m[j][i] = a.m[0][i] * b.m[j][0] + a.m[1][i] * b.m[j][1] + a.m[2][i] * b.m[j][2];
    return r;
}
NDD void m3_apply(const M3& a, const float* v, float* o) {  // a * v
    for (int i = 0; i < 3; ++i) o[i] = a.m[0][i] * v[0] + a.m[1][i] * v[1] + a.m[2][i] * v[2];
}
NDD M3 m3_inverse(const M3& m) {  // dragon_inverse: transpose of the cofactor columns / det
    const float* c0 = m.m[0]; const float* c1 = m.m[1]; const float* c2 = m.m[2];
    float a[3] = {c1[1] * c2[2] - c1[2] * c2[1], c1[2] * c2[0] - c1[0] * c2[2], c1[0] * c2[1] - c1[1] * c2[0]};
    float b[3] = {c2[1] * c0[2] - c2[2] * c0[1], c2[2] * c0[0] - c2[0] * c0[2], c2[0] * c0[1] - c2[1] * c0[0]};
    float d[3] = {c0[1] * c1[2] - c0[2] * c1[1], c0[2] * c1[0] - c0[0] * c1[2], c0[0] * c1[1] - c0[1] * c1[0]};
    const float id = 1.f / (c0[0] * a[0] + c0[1] * a[1] + c0[2] * a[2]);
    M3 r;
    for (int i = 0; i < 3; ++i) { r.m[i][0] = a[i] * id; r.m[i][1] = b[i] * id; r.m[i][2] = d[i] * id; }
    return r;
}
NDD float dragon_radius(const float* c, const float* ob, const M3& qa, const M3& qb, float s) {
    float ac[3], bc[3];
    m3_apply(qa, c, ac);
    m3_apply(qb, c, bc);
    const float x0 = s * ac[0] - c[0], x1 = s * ac[1] - c[1], x2 = s * ac[2] - c[2];
    const float y0 = ob[0] + s * bc[0] - c[0], y1 = ob[1] + s * bc[1] - c[1], y2 = ob[2] + s * bc[2] - c[2];
    return fmaxf(sqrtf(x0 * x0 + x1 * x1 + x2 * x2), sqrtf(y0 * y0 + y1 * y1 + y2 * y2)) / (1.f - s);
}
NDD float dragon_curve_t(uint32_t path, int n, bool levy, float x) {
    float t0 = 0.f, span = 1.f;
    for (int k = 0; k < n; ++k) {
        if ((path >> k) & 1u) { t0 += levy ? 0.5f * span : span; span *= levy ? 0.5f : -0.5f; }
        else span *= 0.5f;
    }
    return t0 + span * x;
}
template <class S> NDD V3<S> vmul_mt(const V3<S>& z, const M3& a) {  // z * a = transpose(a) z
    return mk3(z.x * a.m[0][0] + z.y * a.m[0][1] + z.z * a.m[0][2], z.x * a.m[1][0] + z.y * a.m[1][1] + z.z * a.m[1][2],
               z.x * a.m[2][0] + z.y * a.m[2][1] + z.z * a.m[2][2]);
}
struct FDragon {
    ND_STEP {
        (void)pos;
        const int n = min(max((int)in.p[0], 1), 20);
        const float th = in.p[1] * K_DEG;
        const float s = 0.5f / cosf(th);
        const bool levy = in.p[4] > 0.5f;
        const float thick = in.p[5];
        const M3 qa = m3_mul(m3_rz(th), m3_rx(in.p[2] * K_DEG));
        const M3 qb = m3_mul(m3_rz(levy ? -th : K_PI - th), m3_rx(in.p[3] * K_DEG));
        const float ob[3] = {levy ? 0.5f : 1.f, levy ? 0.5f * tanf(th) : 0.f, 0.f};
        M3 ma, mb;
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) {
                const float e = i == j ? 1.f : 0.f;
                ma.m[j][i] = e - (qa.m[j][i] + qb.m[j][i]) * (0.5f * s);
                mb.m[j][i] = e - qb.m[j][i] * s;
            }
        const float hob[3] = {0.5f * ob[0], 0.5f * ob[1], 0.5f * ob[2]};
        float cen[3], mid[3];
        m3_apply(m3_inverse(ma), hob, cen);
        m3_apply(m3_inverse(mb), ob, mid);
        for (int i = 0; i < 3; ++i) mid[i] *= 0.5f;
        const float rc = dragon_radius(cen, ob, qa, qb, s), rm = dragon_radius(mid, ob, qa, qb, s);
        const float* cc = rm < rc ? mid : cen;
        const float radius = fminf(rc, rm);
        const V3<S> cv = mk3(C<S>(cc[0]), C<S>(cc[1]), C<S>(cc[2]));
        V3<S> zs[21];
        zs[0] = c.z * 0.5f + cv;
        int depth = 0;
        bool fresh = true;
        uint32_t path = 0u, second = 0u;
        S best = C<S>(K_INFINITE);
        float best_t = 0.f;
        const float is = 1.f / s;
        for (int visit = 0; visit < 2048 && depth >= 0; ++visit) {
            const V3<S> z = zs[depth];
            const uint32_t bit = 1u << depth;
            if (fresh) {
                const float here = powf(s, (float)depth);
                const S bound = (len3(z - cv) - radius - thick * powf(s, (float)(n - depth))) * here;
                if (vv(bound) >= vv(best)) { --depth; fresh = false; continue; }
                if (depth == n) {
                    const S x = sclamp(z.x, C<S>(0.f), C<S>(1.f));
                    const S d = (len3(mk3(z.x - x, z.y, z.z)) - thick) * here;
                    if (vv(d) < vv(best)) { best = d; best_t = dragon_curve_t(path, n, levy, vv(x)); }
                    --depth;
                    fresh = false;
                    continue;
                }
            } else if (second & bit) {
                --depth;
                continue;
            }
            const V3<S> za = vmul_mt(z, qa) * is;
            const V3<S> zb = vmul_mt(mk3(z.x - ob[0], z.y - ob[1], z.z - ob[2]), qb) * is;
            const bool a_first = vv(dot3(za - cv, za - cv)) <= vv(dot3(zb - cv, zb - cv));
            const bool pick_b = a_first != fresh;
            second = fresh ? (second & ~bit) : (second | bit);
            path = pick_b ? (path | bit) : (path & ~bit);
            zs[depth + 1] = pick_b ? zb : za;
            ++depth;
            fresh = true;
        }
        for (int k = 0; k <= depth; ++k)
            best = smin_(best, (len3(zs[k] - cv) - radius - thick * powf(s, (float)(n - k))) * powf(s, (float)k));
        ND_RET(mk3(C<S>(best_t), C<S>(0.f), C<S>(0.f)), sdiv(best * 2.f, smax_(sabs(c.dr), C<S>(K_EPS_FINE))));
    }
    ND_FIN { return c.dr; }
};

// formulas/romanesco_ifs.wgsl:8-146: a cone of spiral-packed buds, each bud a smaller cone;
// dr carries the running minimum with a 4-bit level and a 10-bit world scale code.
template <class S> NDD S rom_cone(const V3<S>& z, float s, float tip) {
    const float r2 = fminf(tip, 0.9f * s);
    return round_cone(z, C<S>(s), C<S>(r2), C<S>(1.f - r2 / s));
}
template <class S> NDD S rom_pack(const S& m, float wpl, uint32_t level) {
    const float q = fminf(fmaxf(rintf((2.f - log2f(wpl)) * 32.f), 1.f), 1023.f);
    return pack_low(m, 16383u, ((uint32_t)q << 4) | level);
}
struct RomFrame { float ex, ez, ax, ay, az, cx, cy, cz, bx, by, bz, scale; };
NDD RomFrame rom_frame(float c0, float c1, float u0, float sc, float cc, float lean, float gap, float sink) {
    const float sn = expf(u0 + c0);
    const float th = c1 / sc;
    const float erx = cosf(th), erz = sinf(th);
    const float etx = -erz, etz = erx;  // et = (-er.z, 0, er.x)
    const float nx = erx * cc, ny = sc, nz = erz * cc;
    const float ux = -erx * sc, uy = cc, uz = -erz * sc;
    float sl, cl;
    sincosf(lean, &sl, &cl);
    const float ax = nx * cl + ux * sl, ay = ny * cl + uy * sl, az = nz * cl + uz * sl;
    const float scale = sn * gap;
    RomFrame f;
    f.ex = etx; f.ez = etz;
    f.ax = ax; f.ay = ay; f.az = az;
    // cross(et, axis) with et.y = 0
    f.cx = 0.f * az - etz * ay; f.cy = etz * ax - etx * az; f.cz = etx * ay - 0.f * ax;
    f.bx = (erx * sc) * sn - ax * (sink * scale);
    f.by = 1.f + (-cc) * sn - ay * (sink * scale);
    f.bz = (erz * sc) * sn - az * (sink * scale);
    f.scale = scale;
    return f;
}
template <class S> NDD V3<S> rom_local(const V3<S>& z, const RomFrame& f) {
    const V3<S> v = mk3(z.x - f.bx, z.y - f.by, z.z - f.bz);
    const float is = 1.f / f.scale;
    return mk3((v.x * f.ex + v.z * f.ez) * is, (v.x * f.ax + v.y * f.ay + v.z * f.az) * is,
               (v.x * f.cx + v.y * f.cy + v.z * f.cz) * is);
}
struct FRomanesco {
    ND_STEP {
        (void)pos;
        const uint32_t bits = sbits(c.dr) & 16383u;
        if (bits != 0u && vv(dot3(c.z, c.z)) > 100.f) return c;
        const uint32_t level = bits & 15u;
        float wpl = exp2f(2.f - (float)(bits >> 4) / 32.f);
        V3<S> z = c.z;
        S m = c.dr;
        const float alpha = in.p[1] * K_DEG;
        const float sa = sinf(alpha);
        const float lean = in.p[4] * K_DEG, sink = in.p[5], tip = in.p[6];
        float fi = 3.f, fj = 5.f;
        for (int k = 1; k < (int)in.p[2]; ++k) { const float nx = fi + fj; fi = fj; fj = nx; }
        const float g = 0.381966f;
        const float ei = fi * g - rintf(fi * g), ej = fj * g - rintf(fj * g);
        const float lam = sqrtf((ei * ei - ej * ej) / (fj * fj - fi * fi));
        const float kappa = sqrtf(fi * fi * lam * lam + ei * ei);
        const float t = tanf(alpha) / cosf(lean);
        const float ac = atanf(t * sa / (K_PI * kappa * in.p[3] + t * cosf(alpha)));
        const float sc = sinf(ac), cc = cosf(ac);
        if (bits == 0u) {
            wpl = 1.5f;
            z = mk3(z.x * (1.f / 1.5f), (z.y + 0.5f) * (1.f / 1.5f), z.z * (1.f / 1.5f));
            m = smax_(rom_cone(z, sc, tip), -z.y) * wpl;
        }
        const bool last = (int)level + 1 >= (int)in.p[0];
        const float per = K_TAU * sc;
        const float vax = -fi * lam * per, vay = ei * per, vbx = -fj * lam * per, vby = ej * per;
        const float spacing = per * kappa;
        const float slant = fmaxf(sqrtf(vv(z.x) * vv(z.x) + vv(z.z) * vv(z.z)) * sc - (vv(z.y) - 1.f) * cc, 1e-5f);
        const float u0 = logf(1.f / cc) - 0.5f * spacing;
        const float dx = fminf(logf(slant), u0 + spacing) - u0, dy = atan2f(vv(z.z), vv(z.x)) * sc;
        const float det = vax * vby - vay * vbx;
        const float x0 = rintf((dx * vby - dy * vbx) / det), y0 = rintf((vax * dy - vay * dx) / det);
        const float gap = sinf(alpha - ac) / ((1.f - sink) * cosf(lean));
        RomFrame bf = rom_frame(0.f, 0.f, u0, sc, cc, lean, gap, sink);
        V3<S> best = rom_local(z, bf);
        S best_outer = rom_cone(best, sa, tip) * bf.scale;
        S near = last ? best_outer : rom_cone(best, sc, tip) * bf.scale;
        float best_scale = bf.scale;
        for (int i = -1; i <= 1; ++i)
            for (int j = -1; j <= 1; ++j) {
                const float cx = x0 + (float)i, cy = y0 + (float)j;
                if (cx * fi + cy * fj < -0.5f) continue;
                const RomFrame f = rom_frame(cx * vax + cy * vbx, cx * vay + cy * vby, u0, sc, cc, lean, gap, sink);
                const V3<S> l = rom_local(z, f);
                const S outer = rom_cone(l, sa, tip) * f.scale;
                near = smin_(near, last ? outer : rom_cone(l, sc, tip) * f.scale);
                if (vv(outer) < vv(best_outer)) { best = l; best_outer = outer; best_scale = f.scale; }
            }
        m = smin_(m, near * wpl);
        wpl = wpl * best_scale;
        ND_RET(best, rom_pack(m, wpl, min(level + 1u, 15u)));
    }
    ND_FIN { return with_value(c.dr, __uint_as_float(sbits(c.dr) & ~16383u)); }
};

// ============================================================================= image
// formulas/heightmap_primitive.wgsl:10-64 with image_info / image_texel / image_sample
// (template.wgsl:1110-1193). The host packer puts the slot's image into h[]: h[0] texel
// offset (word index), h[1] width (0: no image bound), h[2] height, h[3..6] the per-channel
// slope bound, h[7..8] the image buffer's device address (nd_image_hoist).
struct ImgInfo {
    const uint32_t* buf;
    uint32_t base, w, h;
    float slope[4];
};
template <class I> NDD ImgInfo img_info(const I& in) {
    ImgInfo f;
    f.base = __float_as_uint(in.h[0]);
    f.w = __float_as_uint(in.h[1]);
    f.h = __float_as_uint(in.h[2]);
    for (int c = 0; c < 4; ++c) f.slope[c] = in.h[3 + c];
    const uint64_t lo = __float_as_uint(in.h[7]), hi = __float_as_uint(in.h[8]);
    f.buf = (const uint32_t*)(uintptr_t)(lo | (hi << 32));
    if (!f.buf) f.w = 0;
    return f;
}
// one texel's channels and its two dilated slope bounds (8-bit codes, 16 steps per octave
// below 2^1; code 0 is 0)
struct ImgTexel { float v[4], nr[4], fr[4]; };
NDD float img_slope(uint32_t code) {
    return code == 0u ? 0.f : fexp2(((float)code - 255.f) * (1.f / 16.f) + 1.f);
}
NDD ImgTexel img_texel(const ImgInfo& f, uint32_t x, uint32_t y) {
    const uint32_t* t = f.buf + f.base + (y * f.w + x) * 4u;
    const uint32_t rg = t[0], ba = t[1], nw = t[2], fw = t[3];
    ImgTexel o;
    const uint32_t q[4] = {rg & 0xffffu, rg >> 16, ba & 0xffffu, ba >> 16};
    for (int c = 0; c < 4; ++c) {
        o.v[c] = (float)q[c] * (1.f / 65535.f);
        o.nr[c] = img_slope((nw >> (8 * c)) & 0xffu);
        o.fr[c] = img_slope((fw >> (8 * c)) & 0xffu);
    }
    return o;
}
template <class S> struct ImgSample { S v[4], nr[4], fr[4]; };
// bilinear at uv in [0, 1]^2, clamped at the edges: the weights carry uv's tangents
template <class S> NDD ImgSample<S> img_sample(const ImgInfo& f, const S& u, const S& v) {
    ImgSample<S> o;
    for (int c = 0; c < 4; ++c) { o.v[c] = C<S>(0.f); o.nr[c] = C<S>(0.f); o.fr[c] = C<S>(0.f); }
    if (f.w == 0u || f.h == 0u) return o;
    const uint32_t lx = f.w - 1u, ly = f.h - 1u;
    const S tx = sclamp(u * (float)f.w - 0.5f, C<S>(0.f), C<S>((float)lx));
    const S ty = sclamp(v * (float)f.h - 0.5f, C<S>(0.f), C<S>((float)ly));
    const uint32_t x0 = min((uint32_t)vv(tx), lx), y0 = min((uint32_t)vv(ty), ly);
    const uint32_t x1 = min(x0 + 1u, lx), y1 = min(y0 + 1u, ly);
    const S fx = tx - (float)x0, fy = ty - (float)y0;
    const S gx = C<S>(1.f) - fx, gy = C<S>(1.f) - fy;
    const ImgTexel a = img_texel(f, x0, y0), b = img_texel(f, x1, y0), cc = img_texel(f, x0, y1), d = img_texel(f, x1, y1);
    // mix(mix(a, b, fx), mix(c, d, fx), fy), mix(x, y, t) = x (1 - t) + y t
    for (int c = 0; c < 4; ++c) {
        o.v[c] = (gx * a.v[c] + fx * b.v[c]) * gy + (gx * cc.v[c] + fx * d.v[c]) * fy;
        o.nr[c] = (gx * a.nr[c] + fx * b.nr[c]) * gy + (gx * cc.nr[c] + fx * d.nr[c]) * fy;
        o.fr[c] = (gx * a.fr[c] + fx * b.fr[c]) * gy + (gx * cc.fr[c] + fx * d.fr[c]) * fy;
    }
    return o;
}
struct FHeightmap {
    ND_STEP {
        (void)pos;
        const float size = fmaxf(in.p[0], K_EPS);
        const V3<S> q = c.z * sdiv(C<S>(1.f), smax_(sabs(c.dr), C<S>(K_EPS_FINE)) * size);
        const float lift = in.p[1], base = fmaxf(in.p[2], 0.f);
        const bool up = lift >= 0.f;
        float k[4] = {in.p[3], in.p[4], in.p[5], in.p[6]};
        if (in.p[7] > 0.5f) {  // invert: (-w.xyz, dot(w, 1))
            const float sum = k[0] + k[1] + k[2] + k[3];
            k[0] = -k[0]; k[1] = -k[1]; k[2] = -k[2]; k[3] = sum;
        }
        const ImgInfo info = img_info(in);
        float half_x = 1.f, half_y = 1.f, texel = 0.f, h_lo = 0.f, h_hi = 0.f;
        ImgSample<S> sm;
        for (int ch = 0; ch < 4; ++ch) { sm.v[ch] = C<S>(0.f); sm.nr[ch] = C<S>(0.f); sm.fr[ch] = C<S>(0.f); }
        if (info.w > 0u) {
            const float mx = fmaxf((float)info.w, (float)info.h);
            half_x = (float)info.w / mx;
            half_y = (float)info.h / mx;
            texel = 2.f / mx;
            sm = img_sample(info, q.x * (1.f / half_x) * 0.5f + 0.5f, q.z * (1.f / half_y) * 0.5f + 0.5f);
            h_lo = fminf(0.f, k[3] + fminf(k[0], 0.f) + fminf(k[1], 0.f) + fminf(k[2], 0.f));
            h_hi = fmaxf(0.f, k[3] + fmaxf(k[0], 0.f) + fmaxf(k[1], 0.f) + fmaxf(k[2], 0.f));
        }
        const S top = (sm.v[0] * k[0] + sm.v[1] * k[1] + sm.v[2] * k[2] + sm.v[3] * k[3]) * lift;
        const S dy = up ? q.y - top : top - q.y;
        const S flat_side = up ? C<S>(-base) - q.y : q.y - base;
        const float ak[4] = {fabsf(k[0]), fabsf(k[1]), fabsf(k[2]), fabsf(k[3])};
        const float per_texel = fabsf(lift) / fmaxf(texel, K_EPS_FINE);
        const float gx = per_texel * (ak[0] * info.slope[0] + ak[1] * info.slope[1] + ak[2] * info.slope[2] + ak[3] * info.slope[3]);
        const S gn = (sm.nr[0] * ak[0] + sm.nr[1] * ak[1] + sm.nr[2] * ak[2] + sm.nr[3] * ak[3]) * per_texel;
        const S gf = (sm.fr[0] * ak[0] + sm.fr[1] * ak[1] + sm.fr[2] * ak[2] + sm.fr[3] * ak[3]) * per_texel;
        const S cx = dy * (1.f / sqrtf(1.f + gx * gx));
        const S cn = sdiv(dy, ssqrt(gn * gn + 1.f)), cf = sdiv(dy, ssqrt(gf * gf + 1.f));
        const float nearb = 8.f * texel, farb = 64.f * texel;
        const S c_near = sclamp(cn, C<S>(-nearb), C<S>(nearb)), c_far = sclamp(cf, C<S>(-farb), C<S>(farb));
        const S curved = vv(dy) >= 0.f ? smax_(cx, smax_(c_near, c_far)) : smin_(cx, smin_(c_near, c_far));
        const float t_min = fminf(lift * h_lo, lift * h_hi), t_max = fmaxf(lift * h_lo, lift * h_hi);
        const float lo = up ? -base : fminf(t_min, base);
        const float hi = up ? fmaxf(t_max, -base) : base;
        const V3<S> e = mk3(sabs(q.x) - half_x, smax_(C<S>(lo) - q.y, q.y - hi), sabs(q.z) - half_y);
        const S bounds = len3(vmax0(e)) + smin_(smax_(e.x, smax_(e.y, e.z)), C<S>(0.f));
        const S d = smax_(smax_(curved, flat_side), bounds);
        const S elev = sclamp((q.y - lo) * (1.f / fmaxf(hi - lo, K_EPS)), C<S>(0.f), C<S>(1.f));
        ND_RET(mk3(C<S>(0.f), elev, C<S>(0.f)), d * size);
    }
    ND_FIN { return c.dr; }
};

#undef ND_STEP
#undef ND_FIN
#undef ND_RET

using FAmazingSurf = FBoxVariant<0, 1>;
using FMandelboxFoam = FBoxVariant<0, 2>;
using FMandelboxSort = FBoxVariant<1, 0>;
using FMandelboxDiag = FBoxVariant<2, 0>;
using FBuffalo = FSphereFoldMandel<true>;
using FApollonian = FSphereFoldMandel<false>;
using FBuffaloJulia = FSphereFoldJulia<true>;
using FApollonianJulia = FSphereFoldJulia<false>;

// ----------------------------------------------------------------------------- the loop
template <class F, bool TRAP>
NDD float ext_run_f(const DevInst& in, const V3<float>& pos, float* trap) {
    Carry<float> c; c.z = pos; c.dr = 1.f;
    TrapAcc acc;
    if (TRAP) acc = trap_begin(in);
    const int iters = in.iters;
    for (int it = 0; it < iters; ++it) {
        const V3<float> zp = c.z;
        c = F::template step<float>(c, pos, in);
        if (TRAP) trap_update(acc, in, c.z, zp, it);
    }
    if (TRAP) *trap = trap_finish(acc, in, iters);
    return F::template fin<float>(c);
}
template <class F, class S>
NDD S ext_run_d(const DevInst& in, const V3<S>& pos) {
    Carry<S> c; c.z = pos; c.dr = C<S>(1.f);
    for (int it = 0; it < in.iters; ++it) c = F::template step<S>(c, pos, in);
    return F::template fin<S>(c);
}

// X(id, struct) over every formula of this file
#define ND_EXT_FORMULAS(X)                                                 \
    X(ND_FORMULA_SPHERE, FSphere)                                          \
    X(ND_FORMULA_CUBE, FCube)                                              \
    X(ND_FORMULA_CYLINDER, FCylinder)                                      \
    X(ND_FORMULA_TORUS, FTorus)                                            \
    X(ND_FORMULA_PYRAMID, FPyramid)                                        \
    X(ND_FORMULA_PRISM, FPrism)                                            \
    X(ND_FORMULA_SIERPINSKI3, FSierpinski3)                                \
    X(ND_FORMULA_CUBEJULIA, FCubeJulia)                                    \
    X(ND_FORMULA_JULIAKIFS, FJuliaKifs)                                    \
    X(ND_FORMULA_DIAGONALTETRA, FDiagonalTetra)                            \
    X(ND_FORMULA_STARKIFS, FStarKifs)                                      \
    X(ND_FORMULA_KALEIDOJULIA, FKaleidoJulia)                              \
    X(ND_FORMULA_BULBTETRA, FBulbTetra)                                    \
    X(ND_FORMULA_ICOSAHEDRALIFS, FIcosahedralIfs)                          \
    X(ND_FORMULA_ICOSAHEDRALJULIA, FIcosahedralJulia)                      \
    X(ND_FORMULA_ICOSAHEDRALTWIST, FIcosahedralTwist)                      \
    X(ND_FORMULA_DIAGONALICOSA, FDiagonalIcosa)                            \
    X(ND_FORMULA_HYBRIDKIFS, FHybridKifs)                                  \
    X(ND_FORMULA_KALEIDOSCOPICIFS, FKaleidoscopicIfs)                      \
    X(ND_FORMULA_JULIABOX, FJuliaBox)                                      \
    X(ND_FORMULA_AMAZINGSURF, FAmazingSurf)                      \
    X(ND_FORMULA_MANDELBOXFOAM, FMandelboxFoam)                    \
    X(ND_FORMULA_MANDELBOXSORT, FMandelboxSort)                    \
    X(ND_FORMULA_MANDELBOXDIAG, FMandelboxDiag)                    \
    X(ND_FORMULA_MANDELBOXINV, FMandelboxInv)                              \
    X(ND_FORMULA_MANDELBOXTWIST, FMandelboxTwist)                          \
    X(ND_FORMULA_PRISMKIFS, FPrismKifs)                                    \
    X(ND_FORMULA_BUFFALO, FBuffalo)                         \
    X(ND_FORMULA_BUFFALOJULIA, FBuffaloJulia)                     \
    X(ND_FORMULA_BUFFALOTWIST, FBuffaloTwist)                              \
    X(ND_FORMULA_APOLLONIAN, FApollonian)                     \
    X(ND_FORMULA_APOLLONIANJULIA, FApollonianJulia)                 \
    X(ND_FORMULA_KLEINIANJULIA, FKleinianJulia)                            \
    X(ND_FORMULA_PSEUDOKLEINIAN, FPseudoKleinian)                          \
    X(ND_FORMULA_JULIABULB, FJuliaBulb)                                    \
    X(ND_FORMULA_MANDELBULBMIRROR, FMandelbulbMirror)                      \
    X(ND_FORMULA_MANDELBULBTWIST, FMandelbulbTwist)                        \
    X(ND_FORMULA_MANDELBULBDUAL, FMandelbulbDual)                          \
    X(ND_FORMULA_QUATJULIA, FQuatJulia)                                    \
    X(ND_FORMULA_BULBBOX, FBulbBox)                                        \
    X(ND_FORMULA_MANDALAY, FMandalay)                                      \
    X(ND_FORMULA_WOBBLEKIFS, FWobbleKifs)                                  \
    X(ND_FORMULA_RADIOACTIVE, FRadioactive)                                \
    X(ND_FORMULA_RECURSIVESPHERE, FRecursiveSphere)                        \
    X(ND_FORMULA_TREE, FTree)                                              \
    X(ND_FORMULA_KLEINIANLIMIT, FKleinianLimit)                            \
    X(ND_FORMULA_FBMSDF, FFbmSdf)                                          \
    X(ND_FORMULA_SEASCAPE, FSeascape)                                      \
    X(ND_FORMULA_DRAGON, FDragon)                                          \
    X(ND_FORMULA_ROMANESCO, FRomanesco)                                    \
    X(ND_FORMULA_HEIGHTMAP, FHeightmap)

// ----------------------------------------------------------------------------- hybrids
// One de_step of any formula (eval_step, fr.zig:515-539), parameters from an instance or
// a mix-in slot. The tuned four use their general forms here.
template <class S, class I>
NDD Carry<S> any_step(int f, const Carry<S>& c, const V3<S>& pos, const I& in) {
    switch (f) {
        case ND_FORMULA_MANDELBULB: {
            bool esc = false;
            return bulb_step_trig<S>(c, pos, P<S>(in, 0), esc);
        }
        case ND_FORMULA_MANDELBOX: {
            const S m = P<S>(in, 1), F = P<S>(in, 2);
            return box_step_t<S>(c, cj_add(pos, in, 8), P<S>(in, 0), m * m, F * F, P<S>(in, 3), fmaxf(fabsf(in.p[8]), 1.f), in);
        }
        case ND_FORMULA_MENGER3:
            return menger_step<S>(c, in);
        case ND_FORMULA_KLEINIAN:
            return kleinian_step<S>(c, cj_add(pos, in, 7), in);
#define ND_CASE_S(ID, F) case ID: return F::template step<S>(c, pos, in);
        ND_EXT_FORMULAS(ND_CASE_S)
#undef ND_CASE_S
    }
    return c;
}
template <class S>
NDD S any_fin(int f, const Carry<S>& c) {
    switch (f) {
        case ND_FORMULA_MANDELBULB: return bulb_finalize<S>(c);
        case ND_FORMULA_MANDELBOX:
        case ND_FORMULA_KLEINIAN: return fin_scaled(c);
        case ND_FORMULA_MENGER3: return menger_finalize<S>(c);
#define ND_CASE_FI(ID, F) case ID: return F::template fin<S>(c);
        ND_EXT_FORMULAS(ND_CASE_FI)
#undef ND_CASE_FI
    }
    return C<S>(1e6f);  // NO_SURFACE_DE
}
// eval_mixed (template.wgsl:1361-1389) with hybrid_step (:1340-1359): hyb_total steps; step
// it is the base formula's while it % cycle < base_n, else the mix-in whose run of
// iterations it falls in; the trap follows every step; the base formula finalizes.
// One hybrid step through one more call: the loop around it stays small, and the step's
// 54-way switch is compiled once per scalar type.
template <class S>
__device__ __noinline__ Carry<S> hybrid_step_call(Carry<S> c, V3<S> pos, const DevMixin* m) {
    return any_step<S>(m->formula, c, pos, *m);
}
template <class S, bool TRAP>
NDD S hybrid_run(const DevInst& in, const V3<S>& pos, TrapAcc* acc) {
    Carry<S> c; c.z = pos; c.dr = C<S>(1.f);
    for (int it = 0; it < in.hyb_total; ++it) {
        int pic = it % in.hyb_cycle;
        int slot = -1;  // -1: no formula at this cycle position (have = false)
        if (pic < in.hyb_base_n) {
            slot = 0;
        } else {
            pic -= in.hyb_base_n;
            for (int j = 0; j < in.mix_count; ++j) {
                if (pic < in.mix[1 + j].iters) { slot = 1 + j; break; }
                pic -= in.mix[1 + j].iters;
            }
        }
        const V3<S> zp = c.z;
        if (slot >= 0) c = hybrid_step_call<S>(c, pos, &in.mix[slot]);
        if constexpr (TRAP) trap_update(*acc, in, val3(c.z), val3(zp), it);
    }
    return any_fin<S>(in.formula, c);
}

// The dispatch is a real call, not inlined: it keeps the 50 formula loops out of the
// register allocation of the kernels' hot paths, which only ever run the four tuned DEs
// inline.
struct ExtDe { float d, trap; };
template <bool TRAP>
__device__ __noinline__ ExtDe ext_de_call(const DevInst* in, float x, float y, float z) {
    const V3<float> pos = mk3(x, y, z);
    ExtDe r; r.trap = 0.f; r.d = 1e6f;  // NO_SURFACE_DE
    if (in->hybrid) {
        TrapAcc acc;
        if (TRAP) acc = trap_begin(*in);
        r.d = hybrid_run<float, TRAP>(*in, pos, &acc);
        if (TRAP) r.trap = trap_finish(acc, *in, in->hyb_total);
        return r;
    }
    switch (in->formula) {
#define ND_CASE_F(ID, F) case ID: r.d = ext_run_f<F, TRAP>(*in, pos, &r.trap); break;
        ND_EXT_FORMULAS(ND_CASE_F)
#undef ND_CASE_F
    }
    return r;
}
template <bool TRAP>
NDD float ext_de_f(const DevInst& in, const V3<float>& pos, float* trap) {
    const ExtDe r = ext_de_call<TRAP>(&in, pos.x, pos.y, pos.z);
    if (TRAP) *trap = r.trap;
    return r.d;
}
template <class S>
__device__ __noinline__ S ext_de_d_call(const DevInst* in, V3<S> pos) {
    if (in->hybrid) return hybrid_run<S, false>(*in, pos, nullptr);
    switch (in->formula) {
#define ND_CASE_D(ID, F) case ID: return ext_run_d<F, S>(*in, pos);
        ND_EXT_FORMULAS(ND_CASE_D)
#undef ND_CASE_D
    }
    return C<S>(1e6f);
}
template <class S>
NDD S ext_de_d(const DevInst& in, const V3<S>& pos) { return ext_de_d_call<S>(&in, pos); }

}  // namespace nd
