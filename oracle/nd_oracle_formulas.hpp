// nd_oracle_formulas.hpp - TEST INFRASTRUCTURE ONLY. Scalar restatement of the formula
// plug-ins beyond the four of nd_oracle.hpp, one struct per formulas/*.wgsl file, written
// from the WGSL read as text (reference tree). Same scalar templating as nd_oracle.hpp:
// T is the evaluation type (double for the FP64 oracle, float, Dual<N>), P the parameter
// type. Included by nd_oracle.hpp after its helpers.
//
// Parity pins (tests/test_oracle_formulas.py): the primitives against brute-force
// distances to sampled surfaces; every other formula through identities with a formula
// already pinned (a Julia variant equals its Mandelbrot sibling with CScale 0 and the Julia
// offset, a twist of angle a equals a z rotation of a, ...), through the symmetries its
// folds give the set, and through its escape and zero-iteration closed forms. The
// packed-carry formulas (radioactive, recursivesphere, tree, kleinianlimit, fbmsdf) are
// pinned only partly: see the test file for which property covers which.
#pragma once

namespace ndo {

// builtin ids 4.. of nd_builtin_formula, in the order of include/nd_hip.h
enum {
    F_SPHERE = 4, F_CUBE, F_CYLINDER, F_TORUS, F_PYRAMID, F_PRISM,
    F_SIERPINSKI3, F_CUBEJULIA, F_JULIAKIFS, F_DIAGONALTETRA, F_STARKIFS, F_KALEIDOJULIA, F_BULBTETRA,
    F_ICOSAHEDRALIFS, F_ICOSAHEDRALJULIA, F_ICOSAHEDRALTWIST, F_DIAGONALICOSA, F_HYBRIDKIFS, F_KALEIDOSCOPICIFS,
    F_JULIABOX, F_AMAZINGSURF, F_MANDELBOXFOAM, F_MANDELBOXSORT, F_MANDELBOXDIAG, F_MANDELBOXINV,
    F_MANDELBOXTWIST, F_PRISMKIFS, F_BUFFALO, F_BUFFALOJULIA, F_BUFFALOTWIST, F_APOLLONIAN, F_APOLLONIANJULIA,
    F_KLEINIANJULIA, F_PSEUDOKLEINIAN, F_JULIABULB, F_MANDELBULBMIRROR, F_MANDELBULBTWIST, F_MANDELBULBDUAL,
    F_QUATJULIA, F_BULBBOX, F_MANDALAY, F_WOBBLEKIFS, F_RADIOACTIVE, F_RECURSIVESPHERE, F_TREE,
    F_KLEINIANLIMIT, F_FBMSDF, F_SEASCAPE, F_DRAGON, F_ROMANESCO, F_HEIGHTMAP, F_COUNT
};

constexpr double TAU = 6.28318530718;   // template.wgsl:60
constexpr double EPSILON_TINY = 1e-12;  // template.wgsl:68

// ----------------------------------------------------------------------------- helpers
inline double setv(double, double v) { return v; }
inline float setv(float, double v) { return (float)v; }
inline Flop setv(Flop, double v) { return Flop(v); }
template <int N> inline Dual<N> setv(Dual<N> x, double v) { x.v = v; return x; }

template <class T> inline T wsign(T x) { return T(val(x) > 0 ? 1.0 : (val(x) < 0 ? -1.0 : 0.0)); }  // WGSL sign
template <class T> inline T radians_of(T deg) { return deg * T(3.14159265358979323846 / 180.0); }
template <class T> inline T wmix(T a, T b, T t) { return a * (T(1) - t) + b * t; }

// The value as FP32 bits (the packed-carry formulas keep a counter in the low bits of dr).
template <class T> inline uint32_t f32_bits(T x) {
    float f = (float)val(x);
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
// select(m, 1e-30, |m| < 1e-30) as FP32, low bits replaced by code (radioactive_pack & co.)
template <class T> inline T pack_f32(T m, uint32_t mask, uint32_t code) {
    float f = (float)val(m);
    if (std::fabs(f) < 1e-30f) f = 1e-30f;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u = (u & ~mask) | code;
    std::memcpy(&f, &u, 4);
    return setv(m, (double)f);
}

// tetra_fold (template.wgsl:1058-1076)
template <class T> V3<T> tetra_fold(V3<T> v) {
    if (v.x + v.y < T(0)) { T t = -v.y; v.y = -v.x; v.x = t; }
    if (v.x + v.z < T(0)) { T t = -v.z; v.z = -v.x; v.x = t; }
    if (v.y + v.z < T(0)) { T t = -v.z; v.z = -v.y; v.y = t; }
    return v;
}
// wrap_angle (template.wgsl:1080-1082)
template <class T> T wrap_angle(T a, T period) { return a - period * floor(a / period); }
// the polar wedge fold of starkifs / kaleidojulia / prismkifs
template <class T> V3<T> wedge(V3<T> z, double degree) {
    T period = T(TAU / degree);
    T a = wrap_angle(atan2(z.y, z.x), period);
    if (a > period * T(0.5)) a = period - a;
    T r = sqrt(z.x * z.x + z.y * z.y);
    return V3<T>(r * cos(a), r * sin(a), z.z);
}
template <class T> V3<T> swap_if_x_lt_y(V3<T> z) {
    if (z.x < z.y) std::swap(z.x, z.y);
    return z;
}
// the three icosahedral mirrors (formulas/icosahedralifs_ifs.wgsl:16-34), normals built
// and normalised in T like the shader builds them in f32
template <class T> V3<T> icosa_fold(V3<T> z) {
    const double phi = 1.6180339887498949;
    V3<T> n[3] = {normalize(V3<T>(T(-phi), T(phi - 1.0), T(1))), normalize(V3<T>(T(1), T(-phi), T(phi + 1.0))),
                  V3<T>(T(0), T(0), T(-1))};
    for (int k = 0; k < 3; ++k) {
        T t = dot(z, n[k]);
        if (t > T(0)) z = z - n[k] * (T(2) * t);
    }
    return z;
}
template <class T> V3<T> twist_z(V3<T> z, T angle) {
    T ca = cos(angle), sa = sin(angle);
    return V3<T>(z.x * ca - z.y * sa, z.x * sa + z.y * ca, z.z);
}
// the WGSL unit direction of a bulb power step with phases
template <class T> V3<T> power_dir(V3<T> z, T r, T power, T tphase, T pphase) {
    T theta = acos(tclamp(z.z / r, T(-1), T(1))) * power + tphase;
    T phi = atan2(z.y, z.x) * power + pphase;
    return V3<T>(sin(theta) * cos(phi), sin(theta) * sin(phi), cos(theta));
}
template <class T> T de_finalize_plain(const Carry<T>& c) { return length(c.z) / c.dr; }  // template.wgsl:999-1001
template <class T> T unit_box_over_dr(const Carry<T>& c) {
    V3<T> q = vabs(c.z) - V3<T>(T(1), T(1), T(1));
    return length(vmax(q, V3<T>())) / c.dr;
}
// primitive_de_step (template.wgsl:1105-1109)
template <class T> Carry<T> primitive_de_step(const Carry<T>& c, T size) {
    T unit_dr = T(1) / tmax(size, T(EPSILON));
    T f = unit_dr / tmax(abs(c.dr), T(EPSILON_FINE));
    return {c.z * f, unit_dr};
}
template <class T> T over_dr(T d, const Carry<T>& c) { return d / tmax(abs(c.dr), T(EPSILON_FINE)); }
// smin (template.wgsl:1653-1659)
template <class T> T smin_poly(T a, T b, T k) {
    if (k <= T(SMOOTH_K_MIN)) return tmin(a, b);
    T h = tclamp(T(0.5) + T(0.5) * (b - a) / k, T(0), T(1));
    return wmix(b, a, h) - k * h * (T(1) - h);
}

#define O_ITERS(k) template <class P> static int iters(const Params<P>& pp) { return i32(val(pp.p[k])); }
#define O_ONE template <class P> static int iters(const Params<P>&) { return 1; }
#define O_STEP template <class T, class P> static Carry<T> step(const Carry<T>& c, V3<T> pos, const Params<P>& pp)
#define O_FIN template <class T> static T fin(const Carry<T>& c)
#define O_P(k) T(pp.p[k])

// ============================================================================= primitives
struct OSphere {  // formulas/sphere_primitive.wgsl:1-12
    O_ONE
    O_STEP { (void)pos; return primitive_de_step(c, O_P(0)); }
    O_FIN { return over_dr(length(c.z) - T(1), c); }
};
struct OCube {  // formulas/cube_primitive.wgsl:1-15
    O_ONE
    O_STEP { (void)pos; return primitive_de_step(c, O_P(0)); }
    O_FIN {
        V3<T> q = vabs(c.z) - V3<T>(T(1), T(1), T(1));
        T outside = length(vmax(q, V3<T>()));
        T inside = tmin(tmax(q.x, tmax(q.y, q.z)), T(0));
        return over_dr(outside + inside, c);
    }
};
struct OCylinder {  // formulas/cylinder_primitive.wgsl:1-15
    O_ONE
    O_STEP { (void)pos; return primitive_de_step(c, O_P(0)); }
    O_FIN {
        T ex = sqrt(c.z.x * c.z.x + c.z.z * c.z.z) - T(1), ey = abs(c.z.y) - T(1);
        T mx = tmax(ex, T(0)), my = tmax(ey, T(0));
        return over_dr(sqrt(mx * mx + my * my) + tmin(tmax(ex, ey), T(0)), c);
    }
};
struct OTorus {  // formulas/torus_primitive.wgsl:1-16
    O_ONE
    O_STEP {
        (void)c;
        T tube = tmax(O_P(1), T(EPSILON));
        V3<T> s = pos / tube;
        return {V3<T>(sqrt(s.x * s.x + s.z * s.z) - O_P(0) / tube, s.y, T(0)), T(1) / tube};
    }
    O_FIN { return over_dr(sqrt(c.z.x * c.z.x + c.z.y * c.z.y) - T(1), c); }
};
struct OPyramid {  // formulas/pyramid_primitive.wgsl:1-21
    O_ONE
    O_STEP { (void)pos; return primitive_de_step(c, O_P(0)); }
    O_FIN {
        T m = tmax(abs(c.z.x), abs(c.z.z));
        T side = (T(2) * m + c.z.y - T(1)) / sqrt(T(5));
        T base = -(c.z.y + T(1));
        return over_dr(tmax(side, base), c);
    }
};
struct OPrism {  // formulas/prism_primitive.wgsl:1-27
    O_ONE
    O_STEP {
        (void)pos;
        Carry<T> q0 = primitive_de_step(c, O_P(0));
        T half_height = O_P(1) / tmax(O_P(0), T(EPSILON));
        double wedge_a = PI / std::max(std::nearbyint((double)val(pp.p[2])), 3.0);
        T cx = T(std::cos(wedge_a)), cy = T(std::sin(wedge_a));
        T a = wrap_angle(atan2(q0.z.x, q0.z.z), T(2.0 * wedge_a)) - T(wedge_a);
        T l = sqrt(q0.z.x * q0.z.x + q0.z.z * q0.z.z);
        T qx = l * cos(a) - cx, qy = l * abs(sin(a)) - cy;
        qy = qy + tclamp(-qy, T(0), cy);
        T edge = sqrt(qx * qx + qy * qy) * (qx > T(0) ? T(1) : T(-1));
        return {V3<T>(edge, abs(q0.z.y) - half_height, T(0)), q0.dr};
    }
    O_FIN {
        T ex = c.z.x, ey = c.z.y;
        T mx = tmax(ex, T(0)), my = tmax(ey, T(0));
        return over_dr(sqrt(mx * mx + my * my) + tmin(tmax(ex, ey), T(0)), c);
    }
};

// ============================================================================= IFS
// z * scale - offset * (scale - 1) after a fold and the formula rotation
template <class T, class P> Carry<T> ifs_tail(V3<T> z, const Carry<T>& c, const Params<P>& pp, int s, int o0, int o1, int o2) {
    T scale = T(pp.p[s]);
    V3<T> off(T(pp.p[o0]), T(pp.p[o1]), T(pp.p[o2]));
    return {z * scale - off * (scale - T(1)), c.dr * scale};
}
struct OSierpinski3 {  // formulas/sierpinski3_ifs.wgsl:1-25
    O_ITERS(1)
    O_STEP { (void)pos; return ifs_tail(rotate_xyz_deg(tetra_fold(c.z), pp.p[5], pp.p[6], pp.p[7]), c, pp, 0, 2, 3, 4); }
    O_FIN { return de_finalize_plain(c); }
};
struct OCubeJulia {  // formulas/cubejulia_ifs_julia.wgsl:1-25
    O_ITERS(4)
    O_STEP { (void)pos; return ifs_tail(rotate_xyz_deg(sort_desc_abs(c.z), pp.p[5], pp.p[6], pp.p[7]), c, pp, 0, 1, 2, 3); }
    O_FIN { return de_finalize_plain(c); }
};
struct OJuliaKifs {  // formulas/juliakifs_ifs_julia.wgsl:1-25
    O_ITERS(4)
    O_STEP { (void)pos; return ifs_tail(rotate_xyz_deg(tetra_fold(c.z), pp.p[5], pp.p[6], pp.p[7]), c, pp, 0, 1, 2, 3); }
    O_FIN { return de_finalize_plain(c); }
};
struct ODiagonalTetra {  // formulas/diagonaltetra_ifs_fold.wgsl:1-33
    O_ITERS(1)
    O_STEP {
        (void)pos;
        return ifs_tail(rotate_xyz_deg(tetra_fold(swap_if_x_lt_y(c.z)), pp.p[5], pp.p[6], pp.p[7]), c, pp, 0, 2, 3, 4);
    }
    O_FIN { return de_finalize_plain(c); }
};
struct OStarKifs {  // formulas/starkifs_ifs_fold.wgsl:1-40
    O_ITERS(2)
    O_STEP {
        (void)pos;
        double degree = std::max(std::floor((double)val(pp.p[0])), 2.0);
        return ifs_tail(rotate_xyz_deg(tetra_fold(wedge(c.z, degree)), pp.p[6], pp.p[7], pp.p[8]), c, pp, 1, 3, 4, 5);
    }
    O_FIN { return de_finalize_plain(c); }
};
struct OKaleidoJulia {  // formulas/kaleidojulia_ifs_fold_julia.wgsl:1-40
    O_ITERS(5)
    O_STEP {
        (void)pos;
        double degree = std::max(std::floor((double)val(pp.p[0])), 2.0);
        return ifs_tail(rotate_xyz_deg(sort_desc_abs(wedge(c.z, degree)), pp.p[6], pp.p[7], pp.p[8]), c, pp, 1, 2, 3, 4);
    }
    O_FIN { return de_finalize_plain(c); }
};
struct OBulbTetra {  // formulas/bulbtetra_hybrid_bulb_ifs.wgsl:1-38
    O_ITERS(2)
    O_STEP {
        T power = O_P(0);
        T r0 = length(c.z);
        V3<T> z = c.z;
        T dr = c.dr;
        if (r0 <= T(BULB_BAILOUT)) {
            V3<T> d = power_dir(z, r0, power, radians_of(O_P(6)), radians_of(O_P(7)));
            T zr = pow(r0, power);
            dr = pow(r0, power - T(1)) * power * dr;
            z = d * zr + pos;
        }
        z = tetra_fold(z);
        return ifs_tail(z, Carry<T>{z, dr}, pp, 1, 3, 4, 5);
    }
    O_FIN { return de_finalize_plain(c); }
};
struct OHybridKifs {  // formulas/hybridkifs_hybrid_ifs.wgsl:1-35
    O_ITERS(1)
    O_STEP {
        (void)pos;
        T scale = O_P(0);
        V3<T> off(O_P(2), O_P(3), O_P(4));
        V3<T> z = rotate_xyz_deg(sort_desc_abs(tetra_fold(c.z)), pp.p[5], pp.p[6], pp.p[7]);
        z = z * scale - off * (scale - T(1));
        if (z.z < T(-0.5) * off.z * (scale - T(1))) z.z = z.z + off.z * (scale - T(1));
        return {z, c.dr * scale};
    }
    O_FIN { return unit_box_over_dr(c); }
};
struct OKaleidoscopicIfs {  // formulas/kaleidoscopicifs_ifs_twist.wgsl:1-37
    O_ITERS(2)
    O_STEP {
        (void)pos;
        T scale = O_P(0);
        V3<T> off(O_P(3), O_P(4), O_P(5));
        V3<T> z = sort_desc_abs(c.z);
        z = z * scale - off * (scale - T(1));
        if (z.z < T(-0.5) * off.z * (scale - T(1))) z.z = z.z + off.z * (scale - T(1));
        z = twist_z(z, radians_of(O_P(1)));
        z = rotate_xyz_deg(z, pp.p[6], pp.p[7], P(0));
        return {z, c.dr * scale};
    }
    O_FIN { return unit_box_over_dr(c); }
};
// the icosahedral family: abs, [swap], mirrors, then rotation and (z - o) s + o in the
// order of each file
template <class T, class P> V3<T> icosa_scale(V3<T> z, const Params<P>& pp, int s, int o0, int o1, int o2) {
    V3<T> off(T(pp.p[o0]), T(pp.p[o1]), T(pp.p[o2]));
    return (z - off) * T(pp.p[s]) + off;
}
struct OIcosahedralIfs {  // formulas/icosahedralifs_ifs.wgsl:1-44
    O_ITERS(3)
    O_STEP {
        (void)pos;
        V3<T> z = rotate_xyz_deg(icosa_fold(vabs(c.z)), pp.p[5], pp.p[6], pp.p[7]);
        return {icosa_scale(z, pp, 0, 1, 2, 4), c.dr * O_P(0)};
    }
    O_FIN { return de_finalize_plain(c); }
};
struct OIcosahedralJulia {  // formulas/icosahedraljulia_ifs_julia.wgsl:1-48
    O_ITERS(6)
    O_STEP {
        (void)pos;
        V3<T> z = rotate_xyz_deg(icosa_fold(vabs(c.z)), pp.p[8], pp.p[9], pp.p[10]);
        return {icosa_scale(z, pp, 0, 1, 2, 7) + V3<T>(O_P(3), O_P(4), O_P(5)), c.dr * O_P(0)};
    }
    O_FIN { return de_finalize_plain(c); }
};
struct OIcosahedralTwist {  // formulas/icosahedraltwist_ifs_twist.wgsl:1-49
    O_ITERS(4)
    O_STEP {
        (void)pos;
        V3<T> z = icosa_scale(icosa_fold(vabs(c.z)), pp, 0, 1, 2, 5);
        z = twist_z(z, radians_of(O_P(3)));
        z = rotate_xyz_deg(z, pp.p[6], pp.p[7], P(0));
        return {z, c.dr * O_P(0)};
    }
    O_FIN { return de_finalize_plain(c); }
};
struct ODiagonalIcosa {  // formulas/diagonalicosa_ifs_fold.wgsl:1-51
    O_ITERS(3)
    O_STEP {
        (void)pos;
        V3<T> z = rotate_xyz_deg(icosa_fold(swap_if_x_lt_y(vabs(c.z))), pp.p[5], pp.p[6], pp.p[7]);
        return {icosa_scale(z, pp, 0, 1, 2, 4), c.dr * O_P(0)};
    }
    O_FIN { return de_finalize_plain(c); }
};

// ============================================================================= Mandelbox relatives
// z * scale + add, dr * |scale| + dr_add
template <class T> Carry<T> scale_add(V3<T> z, T dr, T scale, V3<T> add, T dr_add) {
    return {z * scale + add, dr * abs(scale) + dr_add};
}
template <class T, class P> V3<T> pos_cs_julia(V3<T> pos, const Params<P>& pp, int cs) {
    return pos * T(pp.p[cs]) + V3<T>(T(pp.p[cs + 1]), T(pp.p[cs + 2]), T(pp.p[cs + 3]));
}
template <class T, class P> T cs_dr(const Params<P>& pp, int cs) { return T(std::max(std::fabs((double)val(pp.p[cs])), 1.0)); }
template <class T, class P> Carry<T> sfold(V3<T> z, T dr, const Params<P>& pp, int m, int f) {
    return sphere_fold(z, dr, T(pp.p[m]) * T(pp.p[m]), T(pp.p[f]) * T(pp.p[f]));
}

struct OJuliaBox {  // formulas/juliabox_box_julia.wgsl:1-37
    O_ITERS(7)
    O_STEP {
        (void)pos;
        Carry<T> s = sfold(box_fold(c.z, O_P(3)), c.dr, pp, 1, 2);
        V3<T> z = rotate_xyz_deg(s.z, pp.p[8], pp.p[9], pp.p[10]);
        return scale_add(z, s.dr, O_P(0), V3<T>(O_P(4), O_P(5), O_P(6)), T(1));
    }
    O_FIN { return de_finalize_scaled(c); }
};
struct OAmazingSurf {  // formulas/amazingsurf_box_fold.wgsl:1-38
    O_ITERS(4)
    O_STEP {
        Carry<T> s = sfold(box_fold(c.z, O_P(3)), c.dr, pp, 1, 2);
        V3<T> z = rotate_xyz_deg(s.z, pp.p[5], pp.p[6], pp.p[7]);
        Carry<T> r = scale_add(z, s.dr, O_P(0), pos_cs_julia(pos, pp, 8), cs_dr<T>(pp, 8));
        r.z.z = abs(r.z.z);
        return r;
    }
    O_FIN { return de_finalize_scaled(c); }
};
struct OMandelboxFoam {  // formulas/mandelboxfoam_box_fold.wgsl:1-39
    O_ITERS(4)
    O_STEP {
        Carry<T> s = sfold(box_fold(c.z, O_P(3)), c.dr, pp, 1, 2);
        V3<T> z = rotate_xyz_deg(s.z, pp.p[5], pp.p[6], pp.p[7]);
        Carry<T> r = scale_add(z, s.dr, O_P(0), pos_cs_julia(pos, pp, 8), cs_dr<T>(pp, 8));
        r.z.x = abs(r.z.x);
        r.z.z = abs(r.z.z);
        return r;
    }
    O_FIN { return de_finalize_scaled(c); }
};
struct OMandelboxSort {  // formulas/mandelboxsort_box_fold.wgsl:1-40
    O_ITERS(4)
    O_STEP {
        Carry<T> s = sfold(box_fold(sort_desc_abs(c.z), O_P(3)), c.dr, pp, 1, 2);
        V3<T> z = rotate_xyz_deg(s.z, pp.p[5], pp.p[6], pp.p[7]);
        return scale_add(z, s.dr, O_P(0), pos_cs_julia(pos, pp, 8), cs_dr<T>(pp, 8));
    }
    O_FIN { return de_finalize_scaled(c); }
};
struct OMandelboxDiag {  // formulas/mandelboxdiag_box_fold.wgsl:1-45
    O_ITERS(4)
    O_STEP {
        Carry<T> s = sfold(box_fold(swap_if_x_lt_y(c.z), O_P(3)), c.dr, pp, 1, 2);
        V3<T> z = rotate_xyz_deg(s.z, pp.p[5], pp.p[6], pp.p[7]);
        return scale_add(z, s.dr, O_P(0), pos_cs_julia(pos, pp, 8), cs_dr<T>(pp, 8));
    }
    O_FIN { return de_finalize_scaled(c); }
};
struct OMandelboxInv {  // formulas/mandelboxinv_box_fold.wgsl:1-46
    O_ITERS(5)
    O_STEP {
        T r2in = tmax(dot(c.z, c.z), T(EPSILON_FINE));
        T tin = O_P(0) * O_P(0) / r2in;
        Carry<T> s = sfold(box_fold(c.z * tin, O_P(4)), c.dr * tin, pp, 2, 3);
        V3<T> z = rotate_xyz_deg(s.z, pp.p[6], pp.p[7], pp.p[8]);
        return scale_add(z, s.dr, O_P(1), pos_cs_julia(pos, pp, 9), cs_dr<T>(pp, 9));
    }
    O_FIN { return de_finalize_scaled(c); }
};
struct OMandelboxTwist {  // formulas/mandelboxtwist_box_twist.wgsl:1-42
    O_ITERS(5)
    O_STEP {
        Carry<T> s = sfold(box_fold(c.z, O_P(3)), c.dr, pp, 1, 2);
        V3<T> z = rotate_xyz_deg(twist_z(s.z, radians_of(O_P(4))), pp.p[6], pp.p[7], P(0));
        return scale_add(z, s.dr, O_P(0), pos_cs_julia(pos, pp, 8), cs_dr<T>(pp, 8));
    }
    O_FIN { return de_finalize_scaled(c); }
};
struct OPrismKifs {  // formulas/prismkifs_box_fold.wgsl:1-52
    O_ITERS(5)
    O_STEP {
        double degree = std::max(std::floor((double)val(pp.p[0])), 2.0);
        Carry<T> s = sfold(box_fold(wedge(c.z, degree), O_P(4)), c.dr, pp, 2, 3);
        V3<T> z = rotate_xyz_deg(s.z, pp.p[6], pp.p[7], pp.p[8]);
        return scale_add(z, s.dr, O_P(1), pos_cs_julia(pos, pp, 9), cs_dr<T>(pp, 9));
    }
    O_FIN { return de_finalize_scaled(c); }
};
struct OBuffalo {  // formulas/buffalo_box.wgsl:1-36
    O_ITERS(3)
    O_STEP {
        Carry<T> s = sfold(vabs(c.z), c.dr, pp, 1, 2);
        V3<T> z = rotate_xyz_deg(s.z, pp.p[4], pp.p[5], pp.p[6]);
        return scale_add(z, s.dr, O_P(0), pos_cs_julia(pos, pp, 7), cs_dr<T>(pp, 7));
    }
    O_FIN { return de_finalize_scaled(c); }
};
struct OBuffaloJulia {  // formulas/buffalojulia_box_julia.wgsl:1-36
    O_ITERS(6)
    O_STEP {
        (void)pos;
        Carry<T> s = sfold(vabs(c.z), c.dr, pp, 1, 2);
        V3<T> z = rotate_xyz_deg(s.z, pp.p[7], pp.p[8], pp.p[9]);
        return scale_add(z, s.dr, O_P(0), V3<T>(O_P(3), O_P(4), O_P(5)), T(1));
    }
    O_FIN { return de_finalize_scaled(c); }
};
struct OBuffaloTwist {  // formulas/buffalotwist_box_twist.wgsl:1-41
    O_ITERS(4)
    O_STEP {
        Carry<T> s = sfold(vabs(c.z), c.dr, pp, 1, 2);
        V3<T> z = rotate_xyz_deg(twist_z(s.z, radians_of(O_P(3))), pp.p[5], pp.p[6], P(0));
        return scale_add(z, s.dr, O_P(0), pos_cs_julia(pos, pp, 7), cs_dr<T>(pp, 7));
    }
    O_FIN { return de_finalize_scaled(c); }
};
struct OApollonian {  // formulas/apollonian_kleinian.wgsl:1-36
    O_ITERS(3)
    O_STEP {
        Carry<T> s = sfold(c.z, c.dr, pp, 1, 2);
        V3<T> z = rotate_xyz_deg(s.z, pp.p[4], pp.p[5], pp.p[6]);
        return scale_add(z, s.dr, O_P(0), pos_cs_julia(pos, pp, 7), cs_dr<T>(pp, 7));
    }
    O_FIN { return de_finalize_scaled(c); }
};
struct OApollonianJulia {  // formulas/apollonianjulia_kleinian_julia.wgsl:1-36
    O_ITERS(6)
    O_STEP {
        (void)pos;
        Carry<T> s = sfold(c.z, c.dr, pp, 1, 2);
        V3<T> z = rotate_xyz_deg(s.z, pp.p[7], pp.p[8], pp.p[9]);
        return scale_add(z, s.dr, O_P(0), V3<T>(O_P(3), O_P(4), O_P(5)), T(1));
    }
    O_FIN { return de_finalize_scaled(c); }
};

// ============================================================================= Kleinian relatives
struct OKleinianJulia {  // formulas/kleinianjulia_kleinian_julia.wgsl:1-34
    O_ITERS(7)
    O_STEP {
        (void)pos;
        Carry<T> si = sphere_invert_offset(lattice_wrap(c.z, O_P(8)), c.dr, V3<T>(O_P(0), O_P(1), O_P(2)), O_P(3) * O_P(3));
        return {si.z + V3<T>(O_P(4), O_P(5), O_P(6)), si.dr + T(1)};
    }
    O_FIN { return de_finalize_scaled(c); }
};
struct OPseudoKleinian {  // formulas/pseudokleinian_kleinian.wgsl:1-34
    O_ITERS(5)
    O_STEP {
        Carry<T> si = sphere_invert_offset(lattice_wrap(c.z, O_P(4)), c.dr, V3<T>(O_P(0), O_P(1), O_P(2)), O_P(3) * O_P(3));
        return {si.z + pos_cs_julia(pos, pp, 6), si.dr + cs_dr<T>(pp, 6)};
    }
    O_FIN { return de_finalize_scaled(c); }
};

// ============================================================================= bulb relatives
struct OJuliaBulb {  // formulas/juliabulb_bulb_julia.wgsl:1-30
    O_ITERS(1)
    O_STEP {
        (void)pos;
        T power = O_P(0);
        T r = length(c.z);
        if (r > T(BULB_BAILOUT)) return c;
        V3<T> d = power_dir(c.z, r, power, radians_of(O_P(5)), radians_of(O_P(6)));
        T rp = pow(r, power - T(1));
        return {d * (rp * r) + V3<T>(O_P(2), O_P(3), O_P(4)), rp * power * c.dr + T(1)};
    }
    O_FIN { return bulb_log_escape(c); }
};
struct OMandelbulbMirror {  // formulas/mandelbulbmirror_bulb_fold.wgsl:1-32
    O_ITERS(1)
    O_STEP {
        T power = O_P(0);
        V3<T> zin(c.z.x, -c.z.y, c.z.z);
        T r = length(zin);
        if (r > T(BULB_BAILOUT)) return c;
        V3<T> d = power_dir(zin, r, power, radians_of(O_P(2)), radians_of(O_P(3)));
        T dr = pow(r, power - T(1)) * power * c.dr + cs_dr<T>(pp, 4);
        return {d * pow(r, power) + pos_cs_julia(pos, pp, 4), dr};
    }
    O_FIN { return bulb_log_escape(c); }
};
struct OMandelbulbTwist {  // formulas/mandelbulbtwist_bulb_twist.wgsl:1-37
    O_ITERS(1)
    O_STEP {
        T power = O_P(0);
        T r = length(c.z);
        if (r > T(BULB_BAILOUT)) return c;
        V3<T> d = power_dir(c.z, r, power, radians_of(O_P(3)), radians_of(O_P(4)));
        T dr = pow(r, power - T(1)) * power * c.dr + cs_dr<T>(pp, 5);
        V3<T> z = twist_z(d * pow(r, power), radians_of(O_P(2)));
        return {z + pos_cs_julia(pos, pp, 5), dr};
    }
    O_FIN { return bulb_log_escape(c); }
};
struct OMandelbulbDual {  // formulas/mandelbulbdual_hybrid_bulb.wgsl:1-45
    O_ITERS(1)
    O_STEP {
        T p1 = O_P(0), p2 = O_P(2);
        T r0 = length(c.z);
        if (r0 > T(BULB_BAILOUT)) return c;
        T tp = radians_of(O_P(3)), ph = radians_of(O_P(4));
        T r1 = length(c.z);
        V3<T> z = power_dir(c.z, r1, p1, tp, ph) * pow(r1, p1);
        T dr = pow(r1, p1 - T(1)) * p1 * c.dr;
        T r2 = length(z);
        V3<T> z2 = power_dir(z, r2, p2, tp, ph) * pow(r2, p2);
        dr = pow(r2, p2 - T(1)) * p2 * dr;
        return {z2 + pos_cs_julia(pos, pp, 5), dr + cs_dr<T>(pp, 5)};
    }
    O_FIN { return bulb_log_escape(c); }
};
struct OQuatJulia {  // formulas/quatjulia_bulb_julia.wgsl:1-38
    O_ITERS(5)
    O_STEP {
        (void)pos;
        V3<T> cv(O_P(1), O_P(2), O_P(3));
        T cv_len = length(cv);
        V3<T> s = c.z;
        T dr = -c.dr;
        if (c.dr > T(0)) {
            V3<T> axis = cv_len > T(EPSILON_FINE) ? cv / cv_len : V3<T>(T(1), T(0), T(0));
            V3<T> v(c.z.y, c.z.z, O_P(4));
            T along = dot(v, axis);
            s = V3<T>(c.z.x, along, length(v - axis * along));
            dr = c.dr;
        }
        T r2 = dot(s, s);
        if (r2 > T(256)) return {s, -dr};
        V3<T> q(s.x * s.x - s.y * s.y - s.z * s.z + O_P(0), T(2) * s.x * s.y + cv_len, T(2) * s.x * s.z);
        return {q, T(-2) * sqrt(r2) * dr};
    }
    O_FIN {
        T r = length(c.z);
        return T(0.5) * log(tmax(r, T(EPSILON_FINE))) * r / tmax(abs(c.dr), T(EPSILON_FINE));
    }
};
struct OBulbBox {  // formulas/bulbbox_hybrid_bulb_box.wgsl:1-52
    O_ITERS(5)
    O_STEP {
        T power = O_P(0);
        T r0 = length(c.z);
        V3<T> z = c.z;
        T dr = c.dr;
        if (r0 <= T(BULB_BAILOUT)) {
            V3<T> d = power_dir(z, r0, power, radians_of(O_P(6)), radians_of(O_P(7)));
            dr = pow(r0, power - T(1)) * power * dr;
            z = d * pow(r0, power);
        }
        Carry<T> s = sfold(box_fold(z, O_P(4)), dr, pp, 2, 3);
        V3<T> zr = rotate_xyz_deg(s.z, pp.p[8], pp.p[9], pp.p[10]);
        return scale_add(zr, s.dr, O_P(1), pos_cs_julia(pos, pp, 11), cs_dr<T>(pp, 11));
    }
    O_FIN { return bulb_log_escape(c); }
};

// ============================================================================= others
template <class T> T mandalay_fold(V3<T> p, T fo, T g) {  // formulas/mandalay_box_fold.wgsl:28-37
    V3<T> q = p;
    if (q.z > q.y) q = V3<T>(q.x, q.z, q.y);
    T vx = q.x - T(2) * fo;
    T vy = q.y - T(4) * fo;
    T v = tmax(abs(vx + fo) - fo, vy);
    T v1 = tmax(vx - g, q.y);
    return tmin(tmin(v, v1), q.x);
}
struct OMandalay {  // formulas/mandalay_box_fold.wgsl:1-74
    O_ITERS(7)
    O_STEP {
        V3<T> cc = pos * T(4) + V3<T>(O_P(12), O_P(13), O_P(14));
        V3<T> z = c.z;
        T dr = c.dr;
        if (dr == T(1)) { z = cc; dr = T(4); }
        if (dot(z, z) > T(1e20)) return {z, dr};
        T L = O_P(8);
        z = z - vclamp(z, -L, L) * T(2);
        V3<T> sg(wsign(z.x), wsign(z.y), wsign(z.z));
        V3<T> a = vabs(z);
        V3<T> gap = V3<T>(O_P(9), O_P(10), O_P(11)) * O_P(5);
        z = V3<T>(mandalay_fold(a, O_P(2), gap.x) * sg.x, mandalay_fold(V3<T>(a.y, a.z, a.x), O_P(3), gap.y) * sg.y,
                  mandalay_fold(V3<T>(a.z, a.x, a.y), O_P(4), gap.z) * sg.z);
        T t = tclamp(T(1) / tmax(dot(z, z), T(EPSILON_TINY)), T(1), T(1) / (O_P(1) * O_P(1)));
        z = z * t;
        dr = dr * t;
        z = z * O_P(0) + cc;
        dr = dr * abs(O_P(0)) + T(4);
        z.z = z.z * O_P(6);
        return {z, dr};
    }
    O_FIN { return (length(c.z) - T(30)) / abs(c.dr); }
};
template <class T> void wobble_rot(T& x, T& y, T a) {  // wobblekifs_rot: cos(a) v + sin(a) (v.y, -v.x)
    T nx = cos(a) * x + sin(a) * y, ny = cos(a) * y - sin(a) * x;
    x = nx;
    y = ny;
}
struct OWobbleKifs {  // formulas/wobblekifs_ifs_twist_julia.wgsl:1-65
    O_ITERS(1)
    O_STEP {
        T scale = O_P(0);
        T t = O_P(7) * T(4);
        T len = length(pos);
        V3<T> z = c.z;
        if (c.dr == T(1)) {
            z = V3<T>(z.y, z.x, z.z);
            wobble_rot(z.x, z.y, O_P(15) + sin(t - T(0.25)) * O_P(14));
        }
        T phase_x = O_P(5) + sin(t + len * O_P(12)) * O_P(8);
        T phase_y = O_P(6) + cos(t - T(0.2) + len * O_P(13)) * O_P(9);
        z = vabs(z) * scale + V3<T>(O_P(2), O_P(3), O_P(4));
        wobble_rot(z.x, z.z, phase_x * T(3.14) + cos(t + len) * O_P(10));
        wobble_rot(z.y, z.z, phase_y * T(3.14) + sin(t + len) * O_P(11));
        return {z, c.dr * scale};
    }
    O_FIN {
        V3<T> a = vabs(c.z);
        T m = tmax(tmax(a.x, a.y), tmax(a.z, T(1e-20)));
        V3<T> q = a / m;
        T q6x = q.x * q.x * q.x * q.x * q.x * q.x, q6y = q.y * q.y * q.y * q.y * q.y * q.y, q6z = q.z * q.z * q.z * q.z * q.z * q.z;
        return m * pow(q6x + q6y + q6z, T(1.0 / 6.0)) / c.dr - T(0.25);
    }
};

// ============================================================================= packed carries
struct ORadioactive {  // formulas/radioactive_ifs_carve_twist.wgsl:1-42
    template <class P> static int iters(const Params<P>& pp) { return i32(val(pp.p[0])); }
    O_STEP {
        (void)pos;
        uint32_t level = f32_bits(c.dr) & 31u;
        V3<T> v = c.z;
        T s = c.dr;
        if (level == 0u) s = v.y - O_P(5);
        T i = O_P(1) * pow(O_P(4), T((double)level));
        T co = cos(O_P(2)), sn = sin(O_P(2));
        v = V3<T>(v.x * co + v.z * sn, v.y, v.z * co - v.x * sn);
        T w = T(2) * i;
        V3<T> m(v.x - w * floor(v.x / w) - i, v.y - w * floor(v.y / w) - i, v.z - w * floor(v.z / w) - i);
        v = V3<T>(i * O_P(3), i * O_P(3), i * O_P(3)) - vabs(m);
        s = tmax(s, tmin(tmin(v.x, v.y), v.z));
        return {v, pack_f32(s, 31u, std::min(level + 1u, 31u))};
    }
    O_FIN { return c.dr; }
};
struct ORecursiveSphere {  // formulas/recursivesphere_ifs_carve_twist.wgsl:1-65
    O_ITERS(0)
    O_STEP {
        (void)pos;
        uint32_t level = f32_bits(c.dr) & 31u;
        T world = O_P(1) / T(5.2);
        V3<T> z = c.z;
        T best = c.dr;
        if (level == 0u) { z = z / world; best = T(INFINITE_DISTANCE); }
        T shrink = pow(O_P(5), T((double)level));
        T radius = T(5.2) / O_P(8) * shrink;
        T fin = O_P(6) * shrink;
        T d = length(z) - radius * O_P(8);
        V3<T> corners = vabs(z) + V3<T>(T(-0.57735027) * radius, T(-0.57735027) * radius, T(-0.57735027) * radius);
        T hole = length(corners) - radius * O_P(2);
        V3<T> ap = vabs(z) * T(0.7071);
        T cut = tmax(hole, fin - abs(ap.x - ap.y));
        cut = tmax(cut, fin - abs(ap.y - ap.z));
        cut = tmax(cut, fin - abs(ap.z - ap.x));
        cut = tmin(cut, hole + O_P(3) * shrink);
        d = -smin_poly(-d, cut, O_P(7) * shrink);
        best = tmin(best, d * world);
        T a = O_P(4) / shrink;
        T co = cos(a), s = sin(a);
        corners = V3<T>(co * corners.x - s * corners.z, corners.y, s * corners.x + co * corners.z);
        z = V3<T>(corners.x, corners.z, -corners.y);
        if (level + 1u >= (uint32_t)std::max((double)val(pp.p[0]), 1.0))
            best = tmin(best, (length(z) - radius * O_P(5) * O_P(8)) * world);
        return {z, pack_f32(best, 31u, std::min(level + 1u, 31u))};
    }
    O_FIN { return c.dr; }
};
template <class T> T tree_limb(V3<T> p, T r1, T r2, T h) {  // formulas/tree_ifs.wgsl:19-31
    T qx = sqrt(p.x * p.x + p.z * p.z), qy = p.y;
    T b = (r1 - r2) / h;
    T a = sqrt(T(1) - b * b);
    T k = -b * qx + a * qy;
    if (k < T(0)) return sqrt(qx * qx + qy * qy) - r1;
    if (k > a * h) return sqrt(qx * qx + (qy - h) * (qy - h)) - r2;
    return a * qx + b * qy - r1;
}
struct OTree {  // formulas/tree_ifs.wgsl:1-75
    O_ITERS(0)
    O_STEP {
        (void)pos;
        uint32_t level = f32_bits(c.dr) & 31u;
        T ratio = O_P(3), thick = O_P(5);
        V3<T> z = c.z;
        T m = c.dr;
        T length_here = T(1);
        if (level == 0u) {
            z = (z + V3<T>(T(0), T(1), T(0))) / T(0.6);
            length_here = O_P(6);
            m = tree_limb(z, thick, thick * ratio, length_here) * T(0.6);
        }
        z.y = z.y - length_here;
        T tw = radians_of(O_P(4));
        z = V3<T>(z.x * cos(tw) - z.z * sin(tw), z.y, z.x * sin(tw) + z.z * cos(tw));
        T sector = T(TAU) / T(std::max(std::nearbyint((double)val(pp.p[1])), 1.0));
        T ang = atan2(z.z, z.x);
        ang = abs(ang - sector * rint(ang / sector));
        T rxz = sqrt(z.x * z.x + z.z * z.z);
        z = V3<T>(rxz * cos(ang), z.y, rxz * sin(ang));
        T th = radians_of(O_P(2));
        z = V3<T>(z.x * cos(th) - z.y * sin(th), z.x * sin(th) + z.y * cos(th), z.z) / ratio;
        uint32_t next = std::min(level + 1u, 31u);
        T wpl = T(0.6) * pow(ratio, T((double)next));
        m = tmin(m, tree_limb(z, thick, thick * ratio, T(1)) * wpl);
        if ((int)next >= i32(val(pp.p[0])) && val(pp.p[7]) > 0)
            m = tmin(m, (length(z - V3<T>(T(0), T(1), T(0))) - O_P(7)) * wpl);
        return {z, pack_f32(m, 31u, next)};
    }
    O_FIN { return c.dr; }
};
struct OKleinianLimit {  // formulas/kleinianlimit_kleinian.wgsl:1-76
    O_ITERS(5)
    O_STEP {
        uint32_t stp = f32_bits(c.dr) & 127u;
        if (stp == 127u) return c;
        T a = O_P(0), b = O_P(1), bx = O_P(2), bz = O_P(3);
        V3<T> z = c.z;
        T df = c.dr;
        if (stp == 0u) { z.y = z.y + T(1); df = T(1); }
        z.x = z.x + b / a * z.y;
        z.x = z.x - T(2) * bx * floor((z.x + bx) / (T(2) * bx));
        z.z = z.z - T(2) * bz * floor((z.z + bz) / (T(2) * bz));
        z.x = z.x - b / a * z.y;
        T xs = z.x + b * T(0.5);
        T sep = a * T(0.5) + wsign(b) * (T(2) * a - T(1.95)) / T(4) * wsign(xs) *
                (T(1) - exp(-(T(7.2) - (T(1.95) - a) * T(15)) * abs(xs)));
        if (z.y >= sep) z = V3<T>(-b, a, T(0)) - z;
        T ir = T(1) / tmax(dot(z, z), T(1e-20));
        z = z * (-ir);
        z.x = -b - z.x;
        z.y = a + z.y;
        df = df * ir;
        if ((int)stp + 1 >= i32(val(pp.p[5]))) {
            T y = tmin(z.y, a - z.y);
            T de = tmin(y, T(0.3)) / tmax(df, T(2));
            T cut = pos.y + T(1) - O_P(4) * a;
            return {z, pack_f32(tmax(de, cut), 127u, 127u)};
        }
        return {z, pack_f32(df, 127u, stp + 1u)};
    }
    O_FIN { return c.dr; }
};
// fbmsdf_hash / lattice / simplex (formulas/fbmsdf_primitive_carve.wgsl:32-70). The hash is
// FP32 arithmetic in the shader and its product reaches ~2.5e5 before the fract, so it is
// restated in FP32, one rounding per operation, at every precision of the oracle.
inline double fbm_hash(double xd, double yd, double zd) {
    auto fr = [](float v) { return v - std::floor(v); };
    const float x = (float)xd, y = (float)yd, z = (float)zd;
    const float hx = x * 0.3183099f, hy = y * 0.3183099f, hz = z * 0.3183099f;
    const float qx = 17.0f * fr(hx + 0.11f), qy = 17.0f * fr(hy + 0.17f), qz = 17.0f * fr(hz + 0.13f);
    const float a = qx * qy, b = a * qz, s = qx + qy, s2 = s + qz;
    return (double)fr(b * s2);
}
template <class T> T fbm_lattice(V3<T> p, double radius) {
    double ix = std::floor(val(p.x)), iy = std::floor(val(p.y)), iz = std::floor(val(p.z));
    V3<T> f(p.x - T(ix), p.y - T(iy), p.z - T(iz));
    T d = T(1e10);
    for (unsigned k = 0; k < 8; ++k) {
        double cx = k & 1u, cy = (k >> 1) & 1u, cz = (k >> 2) & 1u;
        double r = fbm_hash(ix + cx, iy + cy, iz + cz);
        d = tmin(d, length(f - V3<T>(T(cx), T(cy), T(cz))) - T(r * r * radius));
    }
    return d;
}
template <class T> T fbm_simplex(V3<T> p, double radius) {
    const double k1 = 0.333333333, k2 = 0.166666667;
    double px = val(p.x), py = val(p.y), pz = val(p.z);
    double sk = (px + py + pz) * k1;
    double i[3] = {std::floor(px + sk), std::floor(py + sk), std::floor(pz + sk)};
    double un = (i[0] + i[1] + i[2]) * k2;
    V3<T> d0(p.x - T(i[0] - un), p.y - T(i[1] - un), p.z - T(i[2] - un));
    double e[3] = {val(d0.x) >= val(d0.y) ? 1.0 : 0.0, val(d0.y) >= val(d0.z) ? 1.0 : 0.0, val(d0.z) >= val(d0.x) ? 1.0 : 0.0};
    double ezxy[3] = {e[2], e[0], e[1]};
    double i1[3], i2[3];
    for (int k = 0; k < 3; ++k) {
        i1[k] = e[k] * (1.0 - ezxy[k]);
        i2[k] = 1.0 - ezxy[k] * (1.0 - e[k]);
    }
    V3<T> d1 = d0 - V3<T>(T(i1[0] - k2), T(i1[1] - k2), T(i1[2] - k2));
    V3<T> d2 = d0 - V3<T>(T(i2[0] - 2 * k2), T(i2[1] - 2 * k2), T(i2[2] - 2 * k2));
    V3<T> d3 = d0 - V3<T>(T(1 - 3 * k2), T(1 - 3 * k2), T(1 - 3 * k2));
    double r0 = fbm_hash(i[0], i[1], i[2]);
    double r1 = fbm_hash(i[0] + i1[0], i[1] + i1[1], i[2] + i1[2]);
    double r2 = fbm_hash(i[0] + i2[0], i[1] + i2[1], i[2] + i2[2]);
    double r3 = fbm_hash(i[0] + 1, i[1] + 1, i[2] + 1);
    double rs = radius * (0.55 / 0.7);
    return tmin(tmin(length(d0) - T(r0 * r0 * rs), length(d1) - T(r1 * r1 * rs)),
                tmin(length(d2) - T(r2 * r2 * rs), length(d3) - T(r3 * r3 * rs)));
}
struct OFbmSdf {  // formulas/fbmsdf_primitive_carve.wgsl:1-107
    O_ITERS(0)
    O_STEP {
        (void)pos;
        uint32_t level = f32_bits(c.dr) & 31u;
        V3<T> z = c.z;
        T d = c.dr;
        if (level == 0u) {
            V3<T> q = vabs(z) - V3<T>(O_P(1), O_P(1), O_P(1));
            d = tmin(tmax(q.x, tmax(q.y, q.z)), T(0)) + length(vmax(q, V3<T>()));
            z = z + V3<T>(T(0.5) + O_P(7), T(0.5) + O_P(7), T(0.5) + O_P(7));
        }
        T s = pow(O_P(4), T((double)level));
        T base = val(pp.p[6]) > 0.5 ? fbm_simplex(z, val(pp.p[2])) : fbm_lattice(z, val(pp.p[2]));
        T n = s * base;
        T k = tmax(O_P(3) * s, T(EPSILON_FINE));
        T h = tmax(k - abs(d + n), T(0));
        d = tmax(d, -n) + h * h * T(0.25) / k;
        z = V3<T>(-T(0.8) * z.y - T(0.6) * z.z, T(0.8) * z.x + T(0.36) * z.y - T(0.48) * z.z,
                  T(0.6) * z.x - T(0.48) * z.y + T(0.64) * z.z) * O_P(5);
        return {z, pack_f32(d, 31u, std::min(level + 1u, 31u))};
    }
    O_FIN { return c.dr; }
};

// ============================================================================= procedural
// formulas/seascape_primitive.wgsl:14-108 (integer hash with the LCG constants of
// template.wgsl:76-81)
inline double sea_hash(int32_t cx, int32_t cy) {
    uint32_t v[2] = {(uint32_t)cx * 1664525u + 1013904223u, (uint32_t)cy * 1664525u + 1013904223u};
    for (int round = 0; round < 2; ++round) {
        v[0] += v[1] * 1664525u;
        v[1] += v[0] * 1664525u;
        v[0] ^= v[0] >> 16u;
        v[1] ^= v[1] >> 16u;
    }
    return (double)((float)(v[0] ^ v[1]) * 2.3283064365386963e-10f);
}
template <class T> T sea_noise(T px, T py) {
    double ix = std::floor(val(px)), iy = std::floor(val(py));
    T fx = px - T(ix), fy = py - T(iy);
    int32_t cx = (int32_t)ix, cy = (int32_t)iy;
    T ux = fx * fx * (T(3) - T(2) * fx), uy = fy * fy * (T(3) - T(2) * fy);
    T lo = wmix(T(sea_hash(cx, cy)), T(sea_hash(cx + 1, cy)), ux);
    T hi = wmix(T(sea_hash(cx, cy + 1)), T(sea_hash(cx + 1, cy + 1)), ux);
    return T(-1) + T(2) * wmix(lo, hi, uy);
}
template <class T> T sea_octave(T ux, T uy, double choppy) {
    T n = sea_noise(ux, uy);
    ux = ux + n;
    uy = uy + n;
    T wx = T(1) - abs(sin(ux)), wy = T(1) - abs(sin(uy));
    wx = wmix(wx, abs(cos(ux)), wx);
    wy = wmix(wy, abs(cos(uy)), wy);
    T crest = pow(tmax(wx * wy, T(0)), T(0.65));
    return pow(tmax(T(1) - crest, T(0)), T(choppy));
}
struct OSeascape {
    O_ONE
    O_STEP {
        (void)pos;
        auto v = [&](int k) { return (double)val(pp.p[k]); };
        int octaves = std::min(std::max(i32(v(0)), 1), 8);
        double size = std::max(v(1), EPSILON);
        V3<T> q = c.z / (T(size) * tmax(abs(c.dr), T(EPSILON_FINE)));
        double sea = std::max(v(2), 0.0), time = v(5), extent = std::max(v(6), EPSILON), depth = std::max(v(7), 0.0);
        double fgain = std::max(v(8), 0.0), again = std::min(std::max(v(9), 0.0), 0.95), relax = std::min(std::max(v(10), 0.0), 1.0);
        double hmax = 2.0 * sea * (1.0 - std::pow(again, (double)octaves)) / (1.0 - again);
        double reach = hmax;
        T above = q.y - T(reach), below = -q.y;
        T ux = q.x * T(0.75), uy = q.z;
        double freq = v(4), amp = sea, choppy = std::max(v(3), 1.0), stretch = 1.0, slope = 0.0;
        T h = T(0);
        for (int i = 0; i < octaves; ++i) {
            T d = sea_octave((ux + T(time)) * T(freq), (uy + T(time)) * T(freq), choppy) +
                  sea_octave((ux - T(time)) * T(freq), (uy - T(time)) * T(freq), choppy);
            h = h + d * T(amp);
            reach -= 2.0 * amp;
            slope += 2.0 * amp * freq * stretch * (3.5 + 1.05 * choppy);
            double sl = std::sqrt(1.0 + slope * slope);
            above = tmax(above, (q.y - h - T(std::max(reach, 0.0))) / T(sl));
            below = tmax(below, (h - q.y) / T(sl));
            T nx = ux * T(1.6) + uy * T(1.2), ny = ux * T(-1.2) + uy * T(1.6);  // uv * mat2x2f(1.6, 1.2, -1.2, 1.6)
            ux = nx;
            uy = ny;
            freq *= fgain;
            amp *= again;
            stretch *= 2.0;
            choppy = choppy * (1.0 - relax) + relax;
        }
        T wave = q.y >= h ? above : -below;
        V3<T> e(abs(q.x) - T(extent), tmax(-T(depth) - q.y, q.y - T(hmax)), abs(q.z) - T(extent));
        T bounds = length(vmax(e, V3<T>())) + tmin(tmax(e.x, tmax(e.y, e.z)), T(0));
        T d = tmax(wave, bounds);
        T elevation = tclamp(h / T(std::max(hmax, EPSILON)), T(0), T(1));
        return {V3<T>(T(0), elevation, T(0)), d * T(size)};
    }
    O_FIN { return c.dr; }
};

// formulas/heightmap_primitive.wgsl:10-64 with image_info / image_texel / image_sample
// (template.wgsl:1110-1193), reading the slot's packed texel words (the same words the
// library's nd_image_pack makes; bound with ndo_bind_image).
struct OImage {
    std::vector<uint32_t> words;  // 4 per texel: RG, BA (16-bit pairs), near / far slope codes
    uint32_t w = 0, h = 0;
    double slope[4] = {0, 0, 0, 0};
};
inline OImage& oracle_image(int slot) {
    static OImage images[16];
    return images[slot < 0 || slot >= 16 ? 0 : slot];
}
inline double image_slope_decode(uint32_t code) {
    return code == 0 ? 0.0 : std::exp2(((double)code - 255.0) / 16.0 + 1.0);
}
struct OImageTexel { double v[4], nr[4], fr[4]; };
inline OImageTexel image_texel(const OImage& im, uint32_t x, uint32_t y) {
    const uint32_t* t = &im.words[4 * ((size_t)y * im.w + x)];
    OImageTexel o;
    const uint32_t q[4] = {t[0] & 0xffffu, t[0] >> 16, t[1] & 0xffffu, t[1] >> 16};
    for (int c = 0; c < 4; ++c) {
        o.v[c] = (double)q[c] / 65535.0;
        o.nr[c] = image_slope_decode((t[2] >> (8 * c)) & 0xffu);
        o.fr[c] = image_slope_decode((t[3] >> (8 * c)) & 0xffu);
    }
    return o;
}
template <class T> struct OImageSample { T v[4], nr[4], fr[4]; };
template <class T> OImageSample<T> image_sample(const OImage& im, T u, T v) {
    OImageSample<T> o;
    for (int c = 0; c < 4; ++c) { o.v[c] = T(0); o.nr[c] = T(0); o.fr[c] = T(0); }
    if (im.w == 0 || im.h == 0) return o;
    const uint32_t lx = im.w - 1, ly = im.h - 1;
    T tx = tclamp(u * T((double)im.w) - T(0.5), T(0), T((double)lx));
    T ty = tclamp(v * T((double)im.h) - T(0.5), T(0), T((double)ly));
    const uint32_t x0 = std::min((uint32_t)val(tx), lx), y0 = std::min((uint32_t)val(ty), ly);
    const uint32_t x1 = std::min(x0 + 1, lx), y1 = std::min(y0 + 1, ly);
    T fx = tx - T((double)x0), fy = ty - T((double)y0);
    const OImageTexel a = image_texel(im, x0, y0), b = image_texel(im, x1, y0), c = image_texel(im, x0, y1),
                      d = image_texel(im, x1, y1);
    for (int ch = 0; ch < 4; ++ch) {
        o.v[ch] = wmix(wmix(T(a.v[ch]), T(b.v[ch]), fx), wmix(T(c.v[ch]), T(d.v[ch]), fx), fy);
        o.nr[ch] = wmix(wmix(T(a.nr[ch]), T(b.nr[ch]), fx), wmix(T(c.nr[ch]), T(d.nr[ch]), fx), fy);
        o.fr[ch] = wmix(wmix(T(a.fr[ch]), T(b.fr[ch]), fx), wmix(T(c.fr[ch]), T(d.fr[ch]), fx), fy);
    }
    return o;
}
struct OHeightmap {
    O_ONE
    O_STEP {
        (void)pos;
        auto v = [&](int k) { return (double)val(pp.p[k]); };
        const double size = std::max(v(0), EPSILON);
        V3<T> q = c.z / (T(size) * tmax(abs(c.dr), T(EPSILON_FINE)));
        const double lift = v(1), base = std::max(v(2), 0.0);
        const bool up = lift >= 0.0;
        double k[4] = {v(3), v(4), v(5), v(6)};
        if (v(7) > 0.5) {  // select(w, vec4f(-w.xyz, dot(w, vec4f(1.0))), p[7] > 0.5)
            const double sum = k[0] + k[1] + k[2] + k[3];
            k[0] = -k[0]; k[1] = -k[1]; k[2] = -k[2]; k[3] = sum;
        }
        const OImage& im = oracle_image(pp.image);
        double half_x = 1, half_y = 1, texel = 0, h_lo = 0, h_hi = 0;
        OImageSample<T> s;
        for (int ch = 0; ch < 4; ++ch) { s.v[ch] = T(0); s.nr[ch] = T(0); s.fr[ch] = T(0); }
        const double slope[4] = {im.w ? im.slope[0] : 0.0, im.w ? im.slope[1] : 0.0, im.w ? im.slope[2] : 0.0,
                                 im.w ? im.slope[3] : 0.0};
        if (im.w > 0) {
            const double mx = (double)std::max(im.w, im.h);
            half_x = im.w / mx;
            half_y = im.h / mx;
            texel = 2.0 / mx;
            s = image_sample(im, q.x / T(half_x) * T(0.5) + T(0.5), q.z / T(half_y) * T(0.5) + T(0.5));
            h_lo = std::min(0.0, k[3] + std::min(k[0], 0.0) + std::min(k[1], 0.0) + std::min(k[2], 0.0));
            h_hi = std::max(0.0, k[3] + std::max(k[0], 0.0) + std::max(k[1], 0.0) + std::max(k[2], 0.0));
        }
        T top = T(lift) * (T(k[0]) * s.v[0] + T(k[1]) * s.v[1] + T(k[2]) * s.v[2] + T(k[3]) * s.v[3]);
        T dy = up ? q.y - top : top - q.y;
        T flat_side = up ? -T(base) - q.y : q.y - T(base);
        const double ak[4] = {std::fabs(k[0]), std::fabs(k[1]), std::fabs(k[2]), std::fabs(k[3])};
        const double per_texel = std::fabs(lift) / std::max(texel, EPSILON_FINE);
        const double gx = per_texel * (ak[0] * slope[0] + ak[1] * slope[1] + ak[2] * slope[2] + ak[3] * slope[3]);
        T gn = T(per_texel) * (T(ak[0]) * s.nr[0] + T(ak[1]) * s.nr[1] + T(ak[2]) * s.nr[2] + T(ak[3]) * s.nr[3]);
        T gf = T(per_texel) * (T(ak[0]) * s.fr[0] + T(ak[1]) * s.fr[1] + T(ak[2]) * s.fr[2] + T(ak[3]) * s.fr[3]);
        T cx = dy / T(std::sqrt(1.0 + gx * gx));
        T cn = dy / sqrt(T(1) + gn * gn), cf = dy / sqrt(T(1) + gf * gf);
        const double nearb = 8.0 * texel, farb = 64.0 * texel;  // IMAGE_SLOPE_RADIUS_NEAR / _FAR
        T c_near = tclamp(cn, T(-nearb), T(nearb)), c_far = tclamp(cf, T(-farb), T(farb));
        T curved = dy >= T(0) ? tmax(cx, tmax(c_near, c_far)) : tmin(cx, tmin(c_near, c_far));
        const double t_min = std::min(lift * h_lo, lift * h_hi), t_max = std::max(lift * h_lo, lift * h_hi);
        const double lo = up ? -base : std::min(t_min, base);
        const double hi = up ? std::max(t_max, -base) : base;
        V3<T> e(abs(q.x) - T(half_x), tmax(T(lo) - q.y, q.y - T(hi)), abs(q.z) - T(half_y));
        T bounds = length(vmax(e, V3<T>())) + tmin(tmax(e.x, tmax(e.y, e.z)), T(0));
        T d = tmax(tmax(curved, flat_side), bounds);
        T elevation = tclamp((q.y - T(lo)) / T(std::max(hi - lo, EPSILON)), T(0), T(1));
        return {V3<T>(T(0), elevation, T(0)), d * T(size)};
    }
    O_FIN { return c.dr; }
};

// formulas/dragon_ifs_twist.wgsl:7-123 (matrices as column arrays, as WGSL mat3x3f)
struct DMat { double c[3][3]; };  // c[col][row]
inline DMat dm_rz(double a) { return DMat{{{std::cos(a), std::sin(a), 0}, {-std::sin(a), std::cos(a), 0}, {0, 0, 1}}}; }
inline DMat dm_rx(double a) { return DMat{{{1, 0, 0}, {0, std::cos(a), std::sin(a)}, {0, -std::sin(a), std::cos(a)}}}; }
inline DMat dm_mul(const DMat& a, const DMat& b) {
    DMat r{};
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) r.c[j][i] += a.c[k][i] * b.c[j][k];
    return r;
}
inline V3<double> dm_apply(const DMat& a, V3<double> v) {
    return V3<double>(a.c[0][0] * v.x + a.c[1][0] * v.y + a.c[2][0] * v.z, a.c[0][1] * v.x + a.c[1][1] * v.y + a.c[2][1] * v.z,
                      a.c[0][2] * v.x + a.c[1][2] * v.y + a.c[2][2] * v.z);
}
inline DMat dm_inverse(const DMat& m) {  // dragon_inverse
    V3<double> c0(m.c[0][0], m.c[0][1], m.c[0][2]), c1(m.c[1][0], m.c[1][1], m.c[1][2]), c2(m.c[2][0], m.c[2][1], m.c[2][2]);
    V3<double> a = cross(c1, c2), b = cross(c2, c0), d = cross(c0, c1);
    double id = 1.0 / dot(c0, a);
    DMat r;
    V3<double> rows[3] = {a, b, d};
    for (int i = 0; i < 3; ++i) {  // transpose(mat3x3f(a, b, d)): row i is rows[i]
        r.c[0][i] = rows[i].x * id; r.c[1][i] = rows[i].y * id; r.c[2][i] = rows[i].z * id;
    }
    return r;
}
template <class T> V3<T> vec_times_mat(V3<T> z, const DMat& a) {  // WGSL z * a
    return V3<T>(z.x * T(a.c[0][0]) + z.y * T(a.c[0][1]) + z.z * T(a.c[0][2]),
                 z.x * T(a.c[1][0]) + z.y * T(a.c[1][1]) + z.z * T(a.c[1][2]),
                 z.x * T(a.c[2][0]) + z.y * T(a.c[2][1]) + z.z * T(a.c[2][2]));
}
inline double dragon_radius(V3<double> c, V3<double> ob, const DMat& qa, const DMat& qb, double s) {
    return std::max(length(dm_apply(qa, c) * s - c), length(ob + dm_apply(qb, c) * s - c)) / (1.0 - s);
}
inline double dragon_curve_t(uint32_t path, int n, bool levy, double x) {
    double t0 = 0.0, span = 1.0;
    for (int k = 0; k < n; ++k) {
        if (((path >> (uint32_t)k) & 1u) == 1u) {
            t0 += levy ? 0.5 * span : span;
            span *= levy ? 0.5 : -0.5;
        } else {
            span *= 0.5;
        }
    }
    return t0 + span * x;
}
struct ODragon {
    O_ONE
    O_STEP {
        (void)pos;
        auto v = [&](int k) { return (double)val(pp.p[k]); };
        int n = std::min(std::max(i32(v(0)), 1), 20);
        double th = v(1) * (3.14159265358979323846 / 180.0);
        double s = 0.5 / std::cos(th);
        bool levy = v(4) > 0.5;
        double thick = v(5);
        DMat qa = dm_mul(dm_rz(th), dm_rx(v(2) * (3.14159265358979323846 / 180.0)));
        DMat qb = dm_mul(dm_rz(levy ? -th : PI - th), dm_rx(v(3) * (3.14159265358979323846 / 180.0)));
        V3<double> ob = levy ? V3<double>(0.5, 0.5 * std::tan(th), 0.0) : V3<double>(1.0, 0.0, 0.0);
        DMat ma, mb;
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) {
                ma.c[j][i] = (i == j ? 1.0 : 0.0) - (qa.c[j][i] + qb.c[j][i]) * (0.5 * s);
                mb.c[j][i] = (i == j ? 1.0 : 0.0) - qb.c[j][i] * s;
            }
        V3<double> centroid = dm_apply(dm_inverse(ma), ob * 0.5);
        V3<double> mid = dm_apply(dm_inverse(mb), ob) * 0.5;
        double rc = dragon_radius(centroid, ob, qa, qb, s), rm = dragon_radius(mid, ob, qa, qb, s);
        V3<double> cd = rm < rc ? mid : centroid;
        double radius = std::min(rc, rm);
        V3<T> cv(T(cd.x), T(cd.y), T(cd.z));
        V3<T> zs[21];
        zs[0] = c.z / T(2) + cv;
        int depth = 0;
        bool fresh = true;
        uint32_t path = 0u, second = 0u;
        T best = T(INFINITE_DISTANCE);
        double best_t = 0.0;
        for (int visit = 0; visit < 2048 && depth >= 0; ++visit) {
            V3<T> z = zs[depth];
            uint32_t bit = 1u << (uint32_t)depth;
            if (fresh) {
                double here = std::pow(s, (double)depth);
                T bound = (length(z - cv) - T(radius) - T(thick * std::pow(s, (double)(n - depth)))) * T(here);
                if (bound >= best) { depth--; fresh = false; continue; }
                if (depth == n) {
                    T x = tclamp(z.x, T(0), T(1));
                    T d = (length(z - V3<T>(x, T(0), T(0))) - T(thick)) * T(here);
                    if (d < best) { best = d; best_t = dragon_curve_t(path, n, levy, (double)val(x)); }
                    depth--;
                    fresh = false;
                    continue;
                }
            } else if ((second & bit) != 0u) {
                depth--;
                continue;
            }
            V3<T> za = vec_times_mat(z, qa) / T(s);
            V3<T> zb = vec_times_mat(z - V3<T>(T(ob.x), T(ob.y), T(ob.z)), qb) / T(s);
            bool a_first = dot(za - cv, za - cv) <= dot(zb - cv, zb - cv);
            bool pick_b = a_first != fresh;
            second = !fresh ? (second | bit) : (second & ~bit);
            path = pick_b ? (path | bit) : (path & ~bit);
            zs[depth + 1] = pick_b ? zb : za;
            depth++;
            fresh = true;
        }
        for (int k = 0; k <= depth; ++k)
            best = tmin(best, (length(zs[k] - cv) - T(radius) - T(thick * std::pow(s, (double)(n - k)))) * T(std::pow(s, (double)k)));
        return {V3<T>(T(best_t), T(0), T(0)), best * T(2) / tmax(abs(c.dr), T(EPSILON_FINE))};
    }
    O_FIN { return c.dr; }
};

// formulas/romanesco_ifs.wgsl:8-146
template <class T> T rom_cone(V3<T> z, double s, double tip) {
    double r2 = std::min(tip, 0.9 * s);
    return tree_limb(z, T(s), T(r2), T(1.0 - r2 / s));
}
template <class T> T rom_pack(T m, double wpl, uint32_t level) {
    double q = std::min(std::max(std::nearbyint((2.0 - std::log2(wpl)) * 32.0), 1.0), 1023.0);
    return pack_f32(m, 16383u, ((uint32_t)q << 4u) | level);
}
template <class T> V3<T> rom_frame(V3<T> z, double c0, double c1, double u0, double sc, double cc, double lean, double gap,
                                   double sink, double* scale_out) {
    double sn = std::exp(u0 + c0);
    double th = c1 / sc;
    V3<double> er(std::cos(th), 0.0, std::sin(th));
    V3<double> et(-er.z, 0.0, er.x);
    V3<double> normal = er * cc + V3<double>(0.0, sc, 0.0);
    V3<double> up_slope = V3<double>(0.0, cc, 0.0) - er * sc;
    V3<double> axis = normal * std::cos(lean) + up_slope * std::sin(lean);
    double scale = sn * gap;
    V3<double> base = V3<double>(0.0, 1.0, 0.0) + (er * sc - V3<double>(0.0, cc, 0.0)) * sn - axis * (sink * scale);
    V3<double> w = cross(et, axis);
    V3<T> v = z - V3<T>(T(base.x), T(base.y), T(base.z));
    *scale_out = scale;
    return V3<T>(dot(v, V3<T>(T(et.x), T(et.y), T(et.z))), dot(v, V3<T>(T(axis.x), T(axis.y), T(axis.z))),
                 dot(v, V3<T>(T(w.x), T(w.y), T(w.z)))) / T(scale);
}
struct ORomanesco {
    O_ITERS(0)
    O_STEP {
        (void)pos;
        auto v = [&](int k) { return (double)val(pp.p[k]); };
        uint32_t bits = f32_bits(c.dr) & 16383u;
        if (bits != 0u && dot(c.z, c.z) > T(100)) return c;
        uint32_t level = bits & 15u;
        double wpl = std::exp2(2.0 - (double)(bits >> 4u) / 32.0);
        V3<T> z = c.z;
        T m = c.dr;
        const double deg = 3.14159265358979323846 / 180.0;
        double alpha = v(1) * deg, sa = std::sin(alpha), lean = v(4) * deg, sink = v(5), tip = v(6);
        double fi = 3.0, fj = 5.0;
        for (int k = 1; k < i32(v(2)); ++k) { double nx = fi + fj; fi = fj; fj = nx; }
        const double g = 0.381966;
        double ei = fi * g - std::nearbyint(fi * g), ej = fj * g - std::nearbyint(fj * g);
        double lam = std::sqrt((ei * ei - ej * ej) / (fj * fj - fi * fi));
        double kappa = std::sqrt(fi * fi * lam * lam + ei * ei);
        double t = std::tan(alpha) / std::cos(lean);
        double ac = std::atan(t * sa / (PI * kappa * v(3) + t * std::cos(alpha)));
        double sc = std::sin(ac), cc = std::cos(ac);
        if (bits == 0u) {
            wpl = 1.5;
            z = (z + V3<T>(T(0), T(0.5), T(0))) / T(wpl);
            m = tmax(rom_cone(z, sc, tip), -z.y) * T(wpl);
        }
        bool last = (int)level + 1 >= i32(v(0));
        double per = TAU * sc;
        double vax = -fi * lam * per, vay = ei * per, vbx = -fj * lam * per, vby = ej * per;
        double spacing = per * kappa;
        double zx = val(z.x), zy = val(z.y), zz = val(z.z);
        double slant = std::max(std::sqrt(zx * zx + zz * zz) * sc - (zy - 1.0) * cc, 1e-5);
        double u0 = std::log(1.0 / cc) - 0.5 * spacing;
        double dx = std::min(std::log(slant), u0 + spacing) - u0, dy = std::atan2(zz, zx) * sc;
        double det = vax * vby - vay * vbx;
        double x0 = std::nearbyint((dx * vby - dy * vbx) / det), y0 = std::nearbyint((vax * dy - vay * dx) / det);
        double gap = std::sin(alpha - ac) / ((1.0 - sink) * std::cos(lean));
        double bscale;
        V3<T> best = rom_frame(z, 0.0, 0.0, u0, sc, cc, lean, gap, sink, &bscale);
        T best_outer = rom_cone(best, sa, tip) * T(bscale);
        T near = last ? best_outer : rom_cone(best, sc, tip) * T(bscale);
        for (int i = -1; i <= 1; ++i)
            for (int j = -1; j <= 1; ++j) {
                double cx = x0 + i, cy = y0 + j;
                if (cx * fi + cy * fj < -0.5) continue;
                double fs;
                V3<T> f = rom_frame(z, cx * vax + cy * vbx, cx * vay + cy * vby, u0, sc, cc, lean, gap, sink, &fs);
                T outer = rom_cone(f, sa, tip) * T(fs);
                near = tmin(near, last ? outer : rom_cone(f, sc, tip) * T(fs));
                if (outer < best_outer) { best = f; best_outer = outer; bscale = fs; }
            }
        m = tmin(m, near * T(wpl));
        wpl = wpl * bscale;
        return {best, rom_pack(m, wpl, std::min(level + 1u, 15u))};
    }
    O_FIN {
        uint32_t u = f32_bits(c.dr) & ~16383u;
        float f;
        std::memcpy(&f, &u, 4);
        return setv(c.dr, (double)f);
    }
};

#undef O_ITERS
#undef O_ONE
#undef O_STEP
#undef O_FIN
#undef O_P

#define ORACLE_EXT_FORMULAS(X)                                                                                     \
    X(F_SPHERE, OSphere) X(F_CUBE, OCube) X(F_CYLINDER, OCylinder) X(F_TORUS, OTorus) X(F_PYRAMID, OPyramid)       \
    X(F_PRISM, OPrism) X(F_SIERPINSKI3, OSierpinski3) X(F_CUBEJULIA, OCubeJulia) X(F_JULIAKIFS, OJuliaKifs)         \
    X(F_DIAGONALTETRA, ODiagonalTetra) X(F_STARKIFS, OStarKifs) X(F_KALEIDOJULIA, OKaleidoJulia)                    \
    X(F_BULBTETRA, OBulbTetra) X(F_ICOSAHEDRALIFS, OIcosahedralIfs) X(F_ICOSAHEDRALJULIA, OIcosahedralJulia)        \
    X(F_ICOSAHEDRALTWIST, OIcosahedralTwist) X(F_DIAGONALICOSA, ODiagonalIcosa) X(F_HYBRIDKIFS, OHybridKifs)        \
    X(F_KALEIDOSCOPICIFS, OKaleidoscopicIfs) X(F_JULIABOX, OJuliaBox) X(F_AMAZINGSURF, OAmazingSurf)                \
    X(F_MANDELBOXFOAM, OMandelboxFoam) X(F_MANDELBOXSORT, OMandelboxSort) X(F_MANDELBOXDIAG, OMandelboxDiag)        \
    X(F_MANDELBOXINV, OMandelboxInv) X(F_MANDELBOXTWIST, OMandelboxTwist) X(F_PRISMKIFS, OPrismKifs)                \
    X(F_BUFFALO, OBuffalo) X(F_BUFFALOJULIA, OBuffaloJulia) X(F_BUFFALOTWIST, OBuffaloTwist)                        \
    X(F_APOLLONIAN, OApollonian) X(F_APOLLONIANJULIA, OApollonianJulia) X(F_KLEINIANJULIA, OKleinianJulia)          \
    X(F_PSEUDOKLEINIAN, OPseudoKleinian) X(F_JULIABULB, OJuliaBulb) X(F_MANDELBULBMIRROR, OMandelbulbMirror)        \
    X(F_MANDELBULBTWIST, OMandelbulbTwist) X(F_MANDELBULBDUAL, OMandelbulbDual) X(F_QUATJULIA, OQuatJulia)          \
    X(F_BULBBOX, OBulbBox) X(F_MANDALAY, OMandalay) X(F_WOBBLEKIFS, OWobbleKifs) X(F_RADIOACTIVE, ORadioactive)     \
    X(F_RECURSIVESPHERE, ORecursiveSphere) X(F_TREE, OTree) X(F_KLEINIANLIMIT, OKleinianLimit) X(F_FBMSDF, OFbmSdf) \
    X(F_SEASCAPE, OSeascape) X(F_DRAGON, ODragon) X(F_ROMANESCO, ORomanesco) X(F_HEIGHTMAP, OHeightmap)

template <class P> int ext_iterations(int f, const Params<P>& p) {
    switch (f) {
#define X_IT(ID, S) case ID: return S::iters(p);
        ORACLE_EXT_FORMULAS(X_IT)
#undef X_IT
    }
    return 0;
}
template <class T, class P> Carry<T> ext_step(int f, const Carry<T>& c, V3<T> pos, const Params<P>& p) {
    switch (f) {
#define X_ST(ID, S) case ID: return S::template step<T, P>(c, pos, p);
        ORACLE_EXT_FORMULAS(X_ST)
#undef X_ST
    }
    return c;
}
template <class T> T ext_finalize(int f, const Carry<T>& c) {
    switch (f) {
#define X_FI(ID, S) case ID: return S::template fin<T>(c);
        ORACLE_EXT_FORMULAS(X_FI)
#undef X_FI
    }
    return T(NO_SURFACE_DE);
}

}  // namespace ndo
