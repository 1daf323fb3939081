// nd_oracle.hpp - TEST INFRASTRUCTURE ONLY. Scalar CPU restatement of the reference's
// distance-estimator ray march, G-buffer and soft shadow, written from the WGSL read as
// text. It is the parity oracle for the HIP kernels: only tests/, the smoke check and the
// bench's cpu_baseline leg may load it, and never as the thing measured or shipped.
//
// Every function is templated on the scalar type so one restatement serves four uses:
//   double      the FP64 oracle
//   float       the FP32 CPU restatement (CPU baseline, "not the reference WGSL")
//   Flop        a counting scalar that measures ND-FLOP v1 per scene_de evaluation
//   Dual<N>     forward-mode derivatives (analytic normals, FP64 reference gradients)
// P is the type of formula parameters: uniform-only subexpressions are computed in P, so
// with (T = Flop, P = double) they are hoisted out of the count, as ND-FLOP v1 says.
//
// Citations: template.wgsl = src/gpu/shaders/template.wgsl; fr.zig =
// src/gpu/fractal_renderer.zig (reference tree).
//
// Parity pins: every helper is checked in tests/test_oracle_*.py against closed forms,
// invariants and brute force; see the test files for which check pins which function.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <thread>
#include <type_traits>
#include <vector>

#include "nd_uniforms.h"

namespace ndo {

// ----------------------------------------------------------------------------- scalars
inline double val(double x) { return x; }
inline float val(float x) { return x; }

// ND-FLOP v1: add/sub/mul/min/max/div/sqrt = 1 (an FMA written as a*b+c counts 2);
// compare/select/abs/neg = 0; transcendentals counted separately as T-ops.
struct FlopCounter {
    uint64_t flops = 0;
    uint64_t tops = 0;
};
inline thread_local FlopCounter g_flops;

struct Flop {
    double v = 0;
    Flop() = default;
    Flop(double x) : v(x) {}  // NOLINT: implicit from constants and hoisted params
};
inline double val(const Flop& x) { return x.v; }
inline Flop operator+(Flop a, Flop b) { g_flops.flops++; return Flop(a.v + b.v); }
inline Flop operator-(Flop a, Flop b) { g_flops.flops++; return Flop(a.v - b.v); }
inline Flop operator*(Flop a, Flop b) { g_flops.flops++; return Flop(a.v * b.v); }
inline Flop operator/(Flop a, Flop b) { g_flops.flops++; return Flop(a.v / b.v); }
inline Flop operator-(Flop a) { return Flop(-a.v); }
inline Flop& operator+=(Flop& a, Flop b) { a = a + b; return a; }
inline Flop& operator-=(Flop& a, Flop b) { a = a - b; return a; }
inline Flop& operator*=(Flop& a, Flop b) { a = a * b; return a; }
inline bool operator<(Flop a, Flop b) { return a.v < b.v; }
inline bool operator>(Flop a, Flop b) { return a.v > b.v; }
inline bool operator<=(Flop a, Flop b) { return a.v <= b.v; }
inline bool operator>=(Flop a, Flop b) { return a.v >= b.v; }
inline bool operator==(Flop a, Flop b) { return a.v == b.v; }
inline bool operator!=(Flop a, Flop b) { return a.v != b.v; }
inline Flop sqrt(Flop a) { g_flops.flops++; return Flop(std::sqrt(a.v)); }
inline Flop abs(Flop a) { return Flop(std::fabs(a.v)); }
inline Flop fmin(Flop a, Flop b) { g_flops.flops++; return Flop(std::fmin(a.v, b.v)); }
inline Flop fmax(Flop a, Flop b) { g_flops.flops++; return Flop(std::fmax(a.v, b.v)); }
inline Flop floor(Flop a) { return Flop(std::floor(a.v)); }
inline Flop rint(Flop a) { return Flop(std::nearbyint(a.v)); }
inline Flop log(Flop a) { g_flops.tops++; return Flop(std::log(a.v)); }
inline Flop exp(Flop a) { g_flops.tops++; return Flop(std::exp(a.v)); }
inline Flop pow(Flop a, Flop b) { g_flops.tops++; return Flop(std::pow(a.v, b.v)); }
inline Flop acos(Flop a) { g_flops.tops++; return Flop(std::acos(a.v)); }
inline Flop atan2(Flop a, Flop b) { g_flops.tops++; return Flop(std::atan2(a.v, b.v)); }
inline Flop sin(Flop a) { g_flops.tops++; return Flop(std::sin(a.v)); }
inline Flop cos(Flop a) { g_flops.tops++; return Flop(std::cos(a.v)); }

// Forward-mode dual number with N tangent directions (FP64).
template <int N>
struct Dual {
    double v = 0;
    double d[N] = {};
    Dual() = default;
    Dual(double x) : v(x) {}  // NOLINT: constants have zero tangent
    static Dual var(double x, int k) { Dual r(x); r.d[k] = 1.0; return r; }
};
template <int N> inline double val(const Dual<N>& x) { return x.v; }
template <int N> inline Dual<N> chain(const Dual<N>& a, double fv, double dfa) {
    Dual<N> r(fv);
    for (int k = 0; k < N; ++k) r.d[k] = dfa * a.d[k];
    return r;
}
template <int N> inline Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v + b.v); for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k]; return r; }
template <int N> inline Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v - b.v); for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k]; return r; }
template <int N> inline Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v * b.v); for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k]; return r; }
template <int N> inline Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v / b.v);
    for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] * b.v - a.v * b.d[k]) / (b.v * b.v);
    return r; }
template <int N> inline Dual<N> operator-(const Dual<N>& a) { return chain(a, -a.v, -1.0); }
template <int N> inline Dual<N>& operator+=(Dual<N>& a, const Dual<N>& b) { a = a + b; return a; }
template <int N> inline Dual<N>& operator-=(Dual<N>& a, const Dual<N>& b) { a = a - b; return a; }
template <int N> inline Dual<N>& operator*=(Dual<N>& a, const Dual<N>& b) { a = a * b; return a; }
template <int N> inline bool operator<(const Dual<N>& a, const Dual<N>& b) { return a.v < b.v; }
template <int N> inline bool operator>(const Dual<N>& a, const Dual<N>& b) { return a.v > b.v; }
template <int N> inline bool operator<=(const Dual<N>& a, const Dual<N>& b) { return a.v <= b.v; }
template <int N> inline bool operator>=(const Dual<N>& a, const Dual<N>& b) { return a.v >= b.v; }
template <int N> inline bool operator==(const Dual<N>& a, const Dual<N>& b) { return a.v == b.v; }
template <int N> inline bool operator!=(const Dual<N>& a, const Dual<N>& b) { return a.v != b.v; }
template <int N> inline Dual<N> sqrt(const Dual<N>& a) {
    double s = std::sqrt(a.v); return chain(a, s, s > 0 ? 0.5 / s : 0.0); }
template <int N> inline Dual<N> abs(const Dual<N>& a) { return a.v < 0 ? -a : a; }
template <int N> inline Dual<N> fmin(const Dual<N>& a, const Dual<N>& b) { return b.v < a.v ? b : a; }
template <int N> inline Dual<N> fmax(const Dual<N>& a, const Dual<N>& b) { return b.v > a.v ? b : a; }
template <int N> inline Dual<N> floor(const Dual<N>& a) { return Dual<N>(std::floor(a.v)); }
template <int N> inline Dual<N> rint(const Dual<N>& a) { return Dual<N>(std::nearbyint(a.v)); }
template <int N> inline Dual<N> log(const Dual<N>& a) { return chain(a, std::log(a.v), 1.0 / a.v); }
template <int N> inline Dual<N> exp(const Dual<N>& a) { double e = std::exp(a.v); return chain(a, e, e); }
template <int N> inline Dual<N> pow(const Dual<N>& a, const Dual<N>& b) {
    // a > 0 on every call site (r of a Mandelbulb iteration); d(a^b) = b a^(b-1) da + a^b ln a db
    double p = std::pow(a.v, b.v);
    Dual<N> r(p);
    double da = (a.v != 0.0) ? b.v * std::pow(a.v, b.v - 1.0) : 0.0;
    double db = (a.v > 0.0) ? p * std::log(a.v) : 0.0;
    for (int k = 0; k < N; ++k) r.d[k] = da * a.d[k] + db * b.d[k];
    return r; }
template <int N> inline Dual<N> acos(const Dual<N>& a) {
    double s = std::sqrt(std::max(1.0 - a.v * a.v, 0.0));
    return chain(a, std::acos(a.v), s > 0 ? -1.0 / s : 0.0); }
template <int N> inline Dual<N> atan2(const Dual<N>& y, const Dual<N>& x) {
    double den = x.v * x.v + y.v * y.v;
    Dual<N> r(std::atan2(y.v, x.v));
    for (int k = 0; k < N; ++k) r.d[k] = den > 0 ? (x.v * y.d[k] - y.v * x.d[k]) / den : 0.0;
    return r; }
template <int N> inline Dual<N> sin(const Dual<N>& a) { return chain(a, std::sin(a.v), std::cos(a.v)); }
template <int N> inline Dual<N> cos(const Dual<N>& a) { return chain(a, std::cos(a.v), -std::sin(a.v)); }

// Plain-type versions (std:: overloads; named so templates find them unqualified).
using std::abs; using std::acos; using std::atan2; using std::cos; using std::exp; using std::floor;
using std::fmax; using std::fmin; using std::log; using std::pow; using std::rint; using std::sin;
using std::sqrt;

template <class T> inline T tmin(T a, T b) { return fmin(a, b); }
template <class T> inline T tmax(T a, T b) { return fmax(a, b); }
template <class T> inline T tclamp(T x, T lo, T hi) { return tmin(tmax(x, lo), hi); }
template <class T> inline T tmix(T a, T b, T t) { return a * (T(1) - t) + b * t; }  // WGSL mix
template <class T> inline T tsel(T f, T t, bool c) { return c ? t : f; }          // WGSL select(f, t, c)
template <class T> inline T tfract(T x) { return x - floor(x); }

// ----------------------------------------------------------------------------- vec3
template <class T>
struct V3 {
    T x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(T a, T b, T c) : x(a), y(b), z(c) {}
    template <class U> static V3 from(const U* p) { return V3(T(p[0]), T(p[1]), T(p[2])); }
};
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline V3<T> cmul(V3<T> a, V3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> inline V3<T> cdiv(V3<T> a, V3<T> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline T length(V3<T> a) { return sqrt(dot(a, a)); }
template <class T> inline V3<T> normalize(V3<T> a) { return a / length(a); }
template <class T> inline V3<T> vabs(V3<T> a) { return {abs(a.x), abs(a.y), abs(a.z)}; }
template <class T> inline V3<T> vmax(V3<T> a, V3<T> b) { return {tmax(a.x, b.x), tmax(a.y, b.y), tmax(a.z, b.z)}; }
template <class T> inline V3<T> vclamp(V3<T> a, T lo, T hi) {
    return {tclamp(a.x, lo, hi), tclamp(a.y, lo, hi), tclamp(a.z, lo, hi)}; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
template <class T, class U> inline V3<T> vcast(V3<U> a) { return {T(a.x), T(a.y), T(a.z)}; }

// ----------------------------------------------------------------------------- constants
// template.wgsl:61-71, 1013, 1597-1598, 1632-1639, 1193-1209
constexpr double EPSILON = 1e-4;
constexpr double EPSILON_FINE = 1e-6;
constexpr double NO_SURFACE_DE = 1e6;
constexpr double INFINITE_DISTANCE = 1e30;
constexpr double BULB_BAILOUT = 2.0;
constexpr double MIN_INSTANCE_SCALE = 0.001;
constexpr double MIN_STEP_SAFETY = 0.001;
constexpr double SMOOTH_K_MIN = 0.0001;
constexpr double TRAP_INITIAL_DISTANCE = 1e6;
constexpr double TRAP_MIN_SPAN = 1e-4;
constexpr double PI = 3.14159265;  // template.wgsl:61 (the shader's own constant)

enum { TRAP_SPHERE = 1, TRAP_BOX = 2, TRAP_CROSS = 3, TRAP_TORUS = 4 };
enum { TRAP_MAX = 1, TRAP_AVERAGE = 2, TRAP_LAST = 3, TRAP_ITERATION = 4 };
enum { COMBINE_UNION = 0, COMBINE_SMOOTH_MIN = 1, COMBINE_SMOOTH_MAX = 2, COMBINE_SMOOTH_INV_MAX = 3,
       COMBINE_SMOOTH_MIN_LIN = 4, COMBINE_SMOOTH_MIN_NLIN = 5, COMBINE_SMOOTH_MIX = 6 };
enum { F_MANDELBULB = 0, F_MANDELBOX = 1, F_MENGER3 = 2, F_KLEINIAN = 3 };

inline int i32(double x) { return (int)x; }  // WGSL i32(f32): truncation toward zero

// ----------------------------------------------------------------------------- DE helpers
template <class T> struct Carry { V3<T> z; T dr; };

// template.wgsl:947-951
template <class T> V3<T> box_fold(V3<T> z, T limit) {
    return vclamp(z, -limit, limit) * T(2) - z;
}
// template.wgsl:959-969 (radii squared)
template <class T> Carry<T> sphere_fold(V3<T> z, T dr, T min_radius2, T fixed_radius2) {
    T r2 = dot(z, z);
    if (r2 < min_radius2) {
        T t = fixed_radius2 / min_radius2;
        return {z * t, dr * t};
    } else if (r2 < fixed_radius2) {
        T t = fixed_radius2 / r2;
        return {z * t, dr * t};
    }
    return {z, dr};
}
// Mandelbox variant where the constant ratio F/m is a hoisted uniform (ND-FLOP v1).
template <class T, class P> Carry<T> sphere_fold_h(V3<T> z, T dr, P min_radius2, P fixed_radius2) {
    T r2 = dot(z, z);
    if (r2 < T(min_radius2)) {
        T t = T(fixed_radius2 / min_radius2);
        return {z * t, dr * t};
    } else if (r2 < T(fixed_radius2)) {
        T t = T(fixed_radius2) / r2;
        return {z * t, dr * t};
    }
    return {z, dr};
}
// template.wgsl:975-983
template <class T> Carry<T> sphere_invert_offset(V3<T> z, T dr, V3<T> center, T radius2) {
    V3<T> d = z - center;
    T r2 = tmax(dot(d, d), T(EPSILON_FINE));
    if (r2 < radius2) {
        T t = radius2 / r2;
        return {center + d * t, dr * t};
    }
    return {z, dr};
}
// template.wgsl:987-989; WGSL round() is half-to-even, i.e. rint in the default mode.
template <class T> V3<T> lattice_wrap(V3<T> p, T cell) {
    return {p.x - cell * rint(p.x / cell), p.y - cell * rint(p.y / cell), p.z - cell * rint(p.z / cell)};
}
// template.wgsl:993-995
template <class T> T de_finalize_scaled(const Carry<T>& c) { return length(c.z) / abs(c.dr); }
// template.wgsl:1006-1010
template <class T> T bulb_log_escape(const Carry<T>& c) {
    T r = length(c.z);
    T log_est = T(0.5) * log(tmax(r, T(EPSILON_FINE))) * r / tmax(c.dr, T(EPSILON_FINE));
    return tmin(log_est, r);
}
// template.wgsl:1019-1031
template <class T> Carry<T> bulb_power_step(const Carry<T>& c, V3<T> add, T power) {
    T r = length(c.z);
    if (r > T(BULB_BAILOUT)) return c;
    T theta = acos(tclamp(c.z.z / r, T(-1), T(1))) * power;
    T phi = atan2(c.z.y, c.z.x) * power;
    T rp = pow(r, power - T(1));
    T zr = rp * r;
    T dr = rp * power * c.dr + T(1);
    V3<T> z = V3<T>(sin(theta) * cos(phi), sin(theta) * sin(phi), cos(theta)) * zr + add;
    return {z, dr};
}
// template.wgsl:1036-1054
template <class T> V3<T> sort_desc_abs(V3<T> z) {
    V3<T> v = vabs(z);
    if (v.x < v.y) std::swap(v.x, v.y);
    if (v.x < v.z) std::swap(v.x, v.z);
    if (v.y < v.z) std::swap(v.y, v.z);
    return v;
}
// template.wgsl:1087-1093 and rot_x/rot_y/rot_z (:1393-1408): z' = Rz * (Ry * (Rx * z)),
// right-handed, columns as written in the WGSL mat3x3f constructors.
template <class T, class P> V3<T> rotate_xyz_deg(V3<T> z, P dx, P dy, P dz) {
    if (dx == P(0) && dy == P(0) && dz == P(0)) return z;
    const double k = 3.14159265358979323846 / 180.0;  // WGSL radians()
    P ax = dx * P(k), ay = dy * P(k), az = dz * P(k);
    T cx = T(cos(ax)), sx = T(sin(ax)), cy = T(cos(ay)), sy = T(sin(ay)), cz = T(cos(az)), sz = T(sin(az));
    // rot_x columns: (1,0,0), (0,c,s), (0,-s,c)  => v' = (x, c y - s z, s y + c z)
    V3<T> a(z.x, cx * z.y - sx * z.z, sx * z.y + cx * z.z);
    // rot_y columns: (c,0,-s), (0,1,0), (s,0,c)  => v' = (c x + s z, y, -s x + c z)
    V3<T> b(cy * a.x + sy * a.z, a.y, cy * a.z - sy * a.x);
    // rot_z columns: (c,s,0), (-s,c,0), (0,0,1)  => v' = (c x - s y, s x + c y, z)
    return V3<T>(cz * b.x - sz * b.y, sz * b.x + cz * b.y, b.z);
}

// ----------------------------------------------------------------------------- formulas
// Parameters of one formula instance (p = the 16 @param slots, in P).
// p: the 16 @param slots; image: the image slot the formula samples (instance * 4 for its
// own formula, + 1 + j for mix-in j: IMAGE_SLOTS_PER_FRACTAL, template.wgsl:1341-1356)
template <class P> struct Params { P p[16]; int image = -1; };

// mandelbrot.wgsl:1-13
template <class P> int bulb_iterations(const Params<P>& p) { return i32(val(p.p[1])); }
template <class T, class P> Carry<T> bulb_step(const Carry<T>& c, V3<T> pos, const Params<P>& p) {
    return bulb_power_step(c, pos, T(p.p[0]));
}
// formulas/amazingbox_box.wgsl:1-37
template <class P> int box_iterations(const Params<P>& p) { return i32(val(p.p[4])); }
template <class T, class P> Carry<T> box_step(const Carry<T>& c, V3<T> pos, const Params<P>& p) {
    P scale = p.p[0];
    P min_radius2 = p.p[1] * p.p[1];
    P fixed_radius2 = p.p[2] * p.p[2];
    V3<T> z = box_fold(c.z, T(p.p[3]));
    Carry<T> sf = sphere_fold_h(z, c.dr, min_radius2, fixed_radius2);
    z = rotate_xyz_deg(sf.z, p.p[5], p.p[6], p.p[7]);
    // z * scale + pos * p[8] + julia; pos * p[8] + julia is loop-invariant per DE call,
    // but written per iteration in the WGSL, so it is counted per iteration here.
    z = z * T(scale) + pos * T(p.p[8]) + V3<T>(T(p.p[9]), T(p.p[10]), T(p.p[11]));
    T dr = sf.dr * T(abs(scale)) + T(fmax(abs(p.p[8]), P(1)));
    return {z, dr};
}
// formulas/menger3_ifs.wgsl:1-29
template <class P> int menger_iterations(const Params<P>& p) { return i32(val(p.p[0])); }
template <class T, class P> Carry<T> menger_step(const Carry<T>& c, V3<T> pos, const Params<P>& p) {
    (void)pos;
    P scale = p.p[4];
    V3<T> offset(T(p.p[1]), T(p.p[2]), T(p.p[3]));
    V3<T> z = rotate_xyz_deg(sort_desc_abs(c.z), p.p[5], p.p[6], p.p[7]);
    T sm1 = T(scale - P(1));
    z = z * T(scale) - offset * sm1;
    if (z.z < T(-0.5) * offset.z * sm1) z.z = z.z + offset.z * sm1;
    return {z, c.dr * T(scale)};
}
template <class T> T menger_finalize(const Carry<T>& c) {
    V3<T> q = vabs(c.z) - V3<T>(T(1), T(1), T(1));
    return length(vmax(q, V3<T>())) / c.dr;
}
// formulas/kleinian_kleinian.wgsl:1-38
template <class P> int kleinian_iterations(const Params<P>& p) { return i32(val(p.p[6])); }
template <class T, class P> Carry<T> kleinian_step(const Carry<T>& c, V3<T> pos, const Params<P>& p) {
    V3<T> center(T(p.p[1]), T(p.p[2]), T(p.p[3]));
    P radius2 = p.p[4] * p.p[4];
    V3<T> z = box_fold(c.z, T(p.p[0]));
    z = lattice_wrap(z, T(p.p[5]));
    Carry<T> si = sphere_invert_offset(z, c.dr, center, T(radius2));
    z = si.z + pos * T(p.p[7]) + V3<T>(T(p.p[8]), T(p.p[9]), T(p.p[10]));
    T dr = si.dr + T(fmax(abs(p.p[7]), P(1)));
    return {z, dr};
}

}  // namespace ndo

#include "nd_oracle_formulas.hpp"

namespace ndo {

template <class P> int formula_iterations(int f, const Params<P>& p) {
    switch (f) {
        case F_MANDELBULB: return bulb_iterations(p);
        case F_MANDELBOX: return box_iterations(p);
        case F_MENGER3: return menger_iterations(p);
        case F_KLEINIAN: return kleinian_iterations(p);
    }
    return ext_iterations(f, p);
}
template <class T, class P> Carry<T> formula_step(int f, const Carry<T>& c, V3<T> pos, const Params<P>& p) {
    switch (f) {
        case F_MANDELBULB: return bulb_step(c, pos, p);
        case F_MANDELBOX: return box_step(c, pos, p);
        case F_MENGER3: return menger_step(c, pos, p);
        case F_KLEINIAN: return kleinian_step(c, pos, p);
    }
    return ext_step(f, c, pos, p);
}
template <class T> T formula_finalize(int f, const Carry<T>& c) {
    switch (f) {
        case F_MANDELBULB: return bulb_log_escape(c);
        case F_MANDELBOX: return de_finalize_scaled(c);
        case F_MENGER3: return menger_finalize(c);
        case F_KLEINIAN: return de_finalize_scaled(c);
    }
    return ext_finalize(f, c);
}

// ----------------------------------------------------------------------------- orbit traps
// template.wgsl:1207-1289
struct TrapSpec {
    double center[3];
    int shape;
    double box[3];
    int mode;
    double radius, tube;
};
template <class T> T trap_distance(const TrapSpec& t, V3<T> z) {
    V3<T> q = z - V3<T>::from(t.center);
    switch (t.shape) {
        case TRAP_SPHERE: return abs(length(q) - T(t.radius));
        case TRAP_BOX: {
            V3<T> e = vabs(q) - V3<T>::from(t.box);
            return abs(length(vmax(e, V3<T>())) + tmin(tmax(e.x, tmax(e.y, e.z)), T(0)));
        }
        case TRAP_CROSS: return tmin(abs(q.x), tmin(abs(q.y), abs(q.z)));
        case TRAP_TORUS: {
            T a = sqrt(q.x * q.x + q.z * q.z) - T(t.radius);
            return abs(sqrt(a * a + q.y * q.y) - T(t.tube));
        }
        default: return length(q);
    }
}
template <class T> struct TrapAcc { T x, y, z, w; };
template <class T> TrapAcc<T> trap_begin(const TrapSpec& t) {
    bool starts_high = t.mode != TRAP_MAX && t.mode != TRAP_AVERAGE;
    return {T(starts_high ? TRAP_INITIAL_DISTANCE : 0.0), T(0), T(0), T(TRAP_INITIAL_DISTANCE)};
}
template <class T> TrapAcc<T> trap_update(TrapAcc<T> a, const TrapSpec& t, V3<T> z, V3<T> z_prev, int it) {
    T d = trap_distance(t, z);
    a.w = d;
    if (t.mode == TRAP_MAX) {
        a.x = tmax(a.x, d);
    } else if (t.mode == TRAP_AVERAGE) {
        if (z.x != z_prev.x || z.y != z_prev.y || z.z != z_prev.z) { a.x = a.x + d; a.y = a.y + T(1); }
    } else if (t.mode == TRAP_ITERATION) {
        if (d < a.x) { a.x = d; a.z = T((double)it); }
    } else {
        a.x = tmin(a.x, d);
    }
    return a;
}
template <class T> T trap_finish(const TrapAcc<T>& a, const TrapSpec& t, int iters) {
    if (t.mode == TRAP_AVERAGE) return (a.y > T(0.5)) ? a.x / tmax(a.y, T(1)) : a.w;
    if (t.mode == TRAP_LAST) return a.w;
    if (t.mode == TRAP_ITERATION) return (a.z + T(1)) / T((double)std::max(iters, 1));
    return a.x;
}

// ----------------------------------------------------------------------------- scene
// Decoded instance (the fields of FractalInstance the path reads).
template <class P> struct Instance {
    int formula = -1;  // builtin id through the code table; -1 = no DE (NO_SURFACE_DE)
    bool hidden = false;
    int combine = 0;
    double blend_k = 0;
    double offset[3] = {0, 0, 0};
    double scale[3] = {1, 1, 1};
    double frame[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};  // columns
    double step_safety = 1;
    Params<P> p;
    TrapSpec trap;
    // hybrid mix-ins (template.wgsl:1332-1389): the instance runs eval_mixed when hybrid
    bool hybrid = false;
    int mix_count = 0, base_n = 0, total = 0;
    int mix_formula[ND_MAX_MIXINS] = {-1, -1, -1};
    int mix_iters[ND_MAX_MIXINS] = {0, 0, 0};
    Params<P> mix_p[ND_MAX_MIXINS];
};

// Empty-space accelerator grid (built by build_accel below).
struct AccelGrid {
    int res = 0, levels = 0;
    double prm[ND_MAX_CASCADES][4] = {};  // corner xyz, cell
    std::vector<double> safe;             // [level][z][y][x] safe skip distances
};

template <class P> struct Scene {
    int count = 1;  // max(i32(instance_count), 1)
    Instance<P> inst[ND_MAX_INSTANCES];
    double max_dist = 10;
    int carve_count = 0;
    double carves[ND_MAX_CARVES][4];
    double cam_offset[3] = {0, 0, 0};  // translation added to the camera (adjoint parameter)
    std::shared_ptr<const AccelGrid> accel;  // null with the accelerator off
};

// code -> builtin formula. Default: identity for the four built-ins.
struct CodeMap {
    int map[16] = {0, 1, 2, 3, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};
    int lookup(int code) const { return (code >= 0 && code < 16) ? map[code] : -1; }
};

// The same scene with every parameter re-typed as Q (its value only, no tangent): every
// instance field, the hybrid mix-ins included, carries over.
template <class Q, class P> Scene<Q> recast_scene(const Scene<P>& s) {
    Scene<Q> r;
    r.count = s.count;
    r.max_dist = s.max_dist;
    r.carve_count = s.carve_count;
    std::memcpy(r.carves, s.carves, sizeof(s.carves));
    std::memcpy(r.cam_offset, s.cam_offset, sizeof(s.cam_offset));
    r.accel = s.accel;
    for (int i = 0; i < ND_MAX_INSTANCES; ++i) {
        const Instance<P>& a = s.inst[i];
        Instance<Q>& b = r.inst[i];
        b.formula = a.formula; b.hidden = a.hidden; b.combine = a.combine; b.blend_k = a.blend_k;
        std::memcpy(b.offset, a.offset, sizeof(a.offset));
        std::memcpy(b.scale, a.scale, sizeof(a.scale));
        std::memcpy(b.frame, a.frame, sizeof(a.frame));
        b.step_safety = a.step_safety;
        b.trap = a.trap;
        for (int k = 0; k < 16; ++k) b.p.p[k] = Q((double)val(a.p.p[k]));
        b.p.image = a.p.image;
        b.hybrid = a.hybrid; b.mix_count = a.mix_count; b.base_n = a.base_n; b.total = a.total;
        for (int j = 0; j < ND_MAX_MIXINS; ++j) {
            b.mix_formula[j] = a.mix_formula[j];
            b.mix_iters[j] = a.mix_iters[j];
            for (int k = 0; k < 16; ++k) b.mix_p[j].p[k] = Q((double)val(a.mix_p[j].p[k]));
            b.mix_p[j].image = a.mix_p[j].image;
        }
    }
    return r;
}

template <class P> Scene<P> decode_scene(const nd_uniforms& u, const CodeMap& cm) {
    Scene<P> s;
    s.count = std::max(i32(u.instance_count), 1);
    s.max_dist = u.max_dist;
    for (int i = 0; i < ND_MAX_INSTANCES; ++i) {
        const nd_fractal_instance& fi = u.instances[i];
        Instance<P>& in = s.inst[i];
        in.formula = cm.lookup(i32(fi.code));  // instance_code: i32(code) (template.wgsl:1317-1319)
        in.hidden = fi.hidden > 0.5f;
        in.combine = i32(fi.combine_mode + 0.5f);
        in.blend_k = fi.blend_k;
        for (int k = 0; k < 3; ++k) {
            in.offset[k] = fi.offset[k];
            in.scale[k] = fi.scale[k];
            for (int r = 0; r < 3; ++r) in.frame[k][r] = fi.frame[k][r];
        }
        in.step_safety = fi.step_safety;
        for (int k = 0; k < 16; ++k) in.p.p[k] = P(fi.params[k]);
        in.p.image = 4 * i;  // g_image_slot = slot * IMAGE_SLOTS_PER_FRACTAL (fr.zig:494)
        for (int k = 0; k < 3; ++k) { in.trap.center[k] = fi.trap_center[k]; in.trap.box[k] = fi.trap_box[k]; }
        in.trap.shape = i32(fi.trap_shape + 0.5f);  // orbit_trap (template.wgsl:1216-1225)
        in.trap.mode = i32(fi.trap_mode + 0.5f);
        in.trap.radius = fi.trap_radius;
        in.trap.tube = fi.trap_tube;
        in.hybrid = fi.mixin_count > 0.5f;  // instance_local_de (template.wgsl:1600-1605)
        in.mix_count = std::min(std::max(i32(fi.mixin_count), 0), ND_MAX_MIXINS);
        in.base_n = std::max(i32(fi.hybrid_base_iters), 0);
        in.total = std::max(i32(fi.hybrid_total_iters), 0);
        for (int j = 0; j < ND_MAX_MIXINS; ++j) {
            in.mix_formula[j] = cm.lookup(i32(fi.mixins[j].code));  // mixin_code (:1321-1323)
            in.mix_iters[j] = std::max(i32(fi.mixins[j].iterations), 0);
            for (int k = 0; k < 16; ++k) in.mix_p[j].p[k] = P(fi.mixins[j].params[k]);
            in.mix_p[j].image = 4 * i + 1 + j;  // st.image_slot += 1 + j (template.wgsl:1352)
        }
    }
    s.carve_count = std::min(i32(u.carve_count), ND_MAX_CARVES);
    for (int i = 0; i < ND_MAX_CARVES; ++i)
        for (int k = 0; k < 4; ++k) s.carves[i][k] = u.carves[i][k];
    return s;
}

// formula_N wrapper (fr.zig:490-513) with the LOD blend off (approx_flags = 0):
// returns (DE, trap value).
// eval_mixed with hybrid_cycle_len / hybrid_step (template.wgsl:1332-1389), LOD off
template <class T, class P> void eval_mixed(const Instance<P>& in, V3<T> pos, T* de, T* trap) {
    int cycle_len = in.base_n;
    for (int j = 0; j < in.mix_count; ++j) cycle_len += in.mix_iters[j];
    cycle_len = std::max(cycle_len, 1);
    Carry<T> carry{pos, T(1)};
    TrapAcc<T> acc = trap_begin<T>(in.trap);
    for (int it = 0; it < in.total; ++it) {
        int code = in.formula;
        const Params<P>* p = &in.p;
        bool have = true;
        int pos_in_cycle = it % cycle_len;
        if (pos_in_cycle >= in.base_n) {
            pos_in_cycle -= in.base_n;
            have = false;
            for (int j = 0; j < in.mix_count; ++j) {
                if (pos_in_cycle < in.mix_iters[j]) {
                    code = in.mix_formula[j];
                    p = &in.mix_p[j];
                    have = true;
                    break;
                }
                pos_in_cycle -= in.mix_iters[j];
            }
        }
        V3<T> z_prev = carry.z;
        if (have) carry = formula_step(code, carry, pos, *p);
        if (trap) acc = trap_update(acc, in.trap, carry.z, z_prev, it);
    }
    *de = formula_finalize(in.formula, carry);
    if (trap) *trap = trap_finish(acc, in.trap, in.total);
}

template <class T, class P> void eval_formula(const Instance<P>& in, V3<T> pos, T* de, T* trap) {
    if (in.hybrid) { eval_mixed(in, pos, de, trap); return; }
    if (in.formula < 0) { *de = T(NO_SURFACE_DE); *trap = T(0); return; }
    Carry<T> carry{pos, T(1)};
    TrapAcc<T> acc = trap_begin<T>(in.trap);
    int iters = formula_iterations(in.formula, in.p);
    for (int it = 0; it < iters; ++it) {
        V3<T> z_prev = carry.z;
        carry = formula_step(in.formula, carry, pos, in.p);
        if (trap) acc = trap_update(acc, in.trap, carry.z, z_prev, it);
    }
    *de = formula_finalize(in.formula, carry);
    if (trap) *trap = trap_finish(acc, in.trap, iters);
}

// instance_local / instance_de_factor / instance_de_trap (template.wgsl:1600-1630)
template <class T, class P> void instance_de(const Instance<P>& in, V3<T> pos, T* de, T* trap) {
    double mag[3], sgn[3];
    for (int k = 0; k < 3; ++k) {
        mag[k] = std::max(std::fabs(in.scale[k]), MIN_INSTANCE_SCALE);
        sgn[k] = in.scale[k] < 0 ? -mag[k] : mag[k];
    }
    V3<T> q = pos - V3<T>::from(in.offset);
    // transpose(frame) * q: component k is dot(column k, q)
    V3<T> local(T(in.frame[0][0]) * q.x + T(in.frame[0][1]) * q.y + T(in.frame[0][2]) * q.z,
                T(in.frame[1][0]) * q.x + T(in.frame[1][1]) * q.y + T(in.frame[1][2]) * q.z,
                T(in.frame[2][0]) * q.x + T(in.frame[2][1]) * q.y + T(in.frame[2][2]) * q.z);
    local = V3<T>(local.x / T(sgn[0]), local.y / T(sgn[1]), local.z / T(sgn[2]));
    T d, tr;
    eval_formula(in, local, &d, trap ? &tr : (T*)nullptr);
    double factor = std::min(mag[0], std::min(mag[1], mag[2])) *
                    std::min(std::max(in.step_safety, MIN_STEP_SAFETY), 1.0);
    *de = d * T(factor);
    if (trap) *trap = tr;
}

// smooth combine family (template.wgsl:1651-1701)
template <class T> T smin(T a, T b, T k) {
    if (k <= T(SMOOTH_K_MIN)) return tmin(a, b);
    T h = tclamp(T(0.5) + T(0.5) * (b - a) / k, T(0), T(1));
    return tmix(b, a, h) - k * h * (T(1) - h);
}
template <class T> T smax(T a, T b, T k) {
    if (k <= T(SMOOTH_K_MIN)) return tmax(a, b);
    T h = tclamp(T(0.5) - T(0.5) * (b - a) / k, T(0), T(1));
    return tmix(b, a, h) + k * h * (T(1) - h);
}
template <class T> T ssub(T cut, T base, T k) {
    if (k <= T(SMOOTH_K_MIN)) return tmax(-cut, base);
    T h = tclamp(T(0.5) - T(0.5) * (base + cut) / k, T(0), T(1));
    return tmix(base, -cut, h) + k * h * (T(1) - h);
}
template <class T> T smin_lin(T a, T b, T k) {
    if (k <= T(SMOOTH_K_MIN)) return tmin(a, b);
    return tmin(b - tmax(k - a, T(0)), a - tmax(k - b, T(0)));
}
template <class T> T smin_nlin(T a, T b, T k) {
    if (k <= T(SMOOTH_K_MIN)) return tmin(a, b);
    T ra = tmax(k - a, T(0));
    T rb = tmax(k - b, T(0));
    return tmin(b - ra * (k - rb) / k, a - rb * (k - ra) / k);
}
template <class T> T smix(T a, T b, T k) {
    if (k <= T(SMOOTH_K_MIN)) return tmin(a, b);
    T h = tclamp(T(0.5) + T(0.5) * (b - a) / k, T(0), T(1));
    return tmix(b, a, h);
}
// combine_de (template.wgsl:1703-1730): returns the distance and the material weight h.
template <class T> void combine_de(int mode, T d, T di, T k, T* out_d, T* out_h) {
    bool hard = k <= T(SMOOTH_K_MIN);
    T closer = tsel(T(1), T(0), di < d);
    if (mode == COMBINE_UNION) { *out_d = tmin(d, di); *out_h = closer; return; }
    if (mode == COMBINE_SMOOTH_MIN || mode == COMBINE_SMOOTH_MIX) {
        // select() evaluates both arms; a hard k never divides by it in the taken arm.
        T h = hard ? closer : tclamp(T(0.5) + T(0.5) * (di - d) / k, T(0), T(1));
        *out_d = mode == COMBINE_SMOOTH_MIX ? smix(d, di, k) : smin(d, di, k);
        *out_h = h;
        return;
    }
    if (mode == COMBINE_SMOOTH_MAX) {
        T h = hard ? tsel(T(1), T(0), di > d) : tclamp(T(0.5) - T(0.5) * (di - d) / k, T(0), T(1));
        *out_d = smax(d, di, k);
        *out_h = h;
        return;
    }
    if (mode == COMBINE_SMOOTH_INV_MAX) {
        T h = hard ? tsel(T(1), T(0), -di > d) : T(1) - tclamp(T(0.5) - T(0.5) * (d + di) / k, T(0), T(1));
        *out_d = ssub(di, d, k);
        *out_h = h;
        return;
    }
    if (mode == COMBINE_SMOOTH_MIN_LIN) { *out_d = smin_lin(d, di, k); *out_h = closer; return; }
    *out_d = smin_nlin(d, di, k);
    *out_h = closer;
}

// scene_de (template.wgsl:1756-1779) with the identity warp, no particles, carves applied
// (template.wgsl:1744-1754). If trap != null it also returns the trap blended across
// instances with the combine weight the way hit_material blends materials
// (template.wgsl:2128-2155).
template <class T, class P> T scene_de(const Scene<P>& s, V3<T> pos, T* trap = nullptr) {
    T d = T(s.max_dist * 2.0 + 1.0);  // empty_scene_de
    T tr = T(0);
    bool first = true;
    for (int i = 0; i < s.count; ++i) {
        const Instance<P>& in = s.inst[i];
        if (in.hidden) continue;
        T di, ti;
        instance_de(in, pos, &di, trap ? &ti : (T*)nullptr);
        if (first) {
            d = di;
            if (trap) tr = ti;
            first = false;
        } else {
            T nd, h;
            combine_de(in.combine, d, di, T(in.blend_k), &nd, &h);
            if (trap) tr = tmix(ti, tr, h);  // mix(b = new, a = acc, h) as mix_material(a, b, h)
            d = nd;
        }
    }
    if (first) {
        if (trap) *trap = T(0);
        return T(s.max_dist * 2.0 + 1.0);
    }
    for (int i = 0; i < s.carve_count; ++i) {
        const double* c = s.carves[i];
        V3<T> cc{T(c[0]), T(c[1]), T(c[2])};
        d = ssub(length(pos - cc) - T(c[3]), d, T(0.2 * c[3]));
    }
    if (trap) *trap = tr;
    return d;
}

// ----------------------------------------------------------------------------- FLOP probe
// When enabled, every FP64 scene_de evaluated by march_to / visibility_march is evaluated a
// second time on the counting scalar at the same point, so ND-FLOP v1 per DE can be averaged
// over the workload's actual evaluation points.
struct FlopProbe {
    bool on = false;
    uint64_t evals = 0;
    uint64_t flops = 0;
    uint64_t tops = 0;
};
inline thread_local FlopProbe g_probe;

template <class T, class P> T probed_de(const Scene<P>& s, V3<T> pos) {
    if constexpr (std::is_same<T, double>::value && std::is_same<P, double>::value) {
        if (g_probe.on) {
            FlopCounter saved = g_flops;
            g_flops = FlopCounter{};
            (void)scene_de(s, V3<Flop>(Flop(pos.x), Flop(pos.y), Flop(pos.z)));
            g_probe.evals++;
            g_probe.flops += g_flops.flops;
            g_probe.tops += g_flops.tops;
            g_flops = saved;
        }
    }
    return scene_de(s, pos);
}

// ----------------------------------------------------------------------------- march
// ----------------------------------------------------------------------------- accelerator
// placeCascades (accel.zig:34-58): `levels` cubic grids of res^3 cells around the camera,
// level l with cell = (2 max_dist / res) / 2^(levels-1-l) (level 0 finest), each corner
// snapped down to a multiple of its cell.
inline void place_cascades(const double cam[3], double max_dist, int levels, int res, double out[ND_MAX_CASCADES][4]) {
    for (int l = 0; l < ND_MAX_CASCADES; ++l) out[l][0] = out[l][1] = out[l][2] = out[l][3] = 0.0;
    const int n = std::min(std::max(levels, 1), ND_MAX_CASCADES);
    const float res_f = (float)std::max(res, 1);
    const float coarsest = std::max(2.0f * (float)max_dist / res_f, 1e-6f);
    for (int l = 0; l < n; ++l) {
        const float cell = coarsest / (float)(1u << (n - 1 - l));
        const float half = 0.5f * res_f * cell;
        for (int a = 0; a < 3; ++a) out[l][a] = std::floor(((float)cam[a] - half) / cell) * cell;
        out[l][3] = cell;
    }
}

// cs_build_accel (template.wgsl:5409-5431): a voxel stores max(DE(centre) - HALF_SQRT3 * cell,
// 0) * safety, DE evaluated in the scene's parameter precision P.
template <class P> std::shared_ptr<const AccelGrid> build_accel(const nd_uniforms& u, const Scene<P>& s) {
    auto g = std::make_shared<AccelGrid>();
    g->res = i32(u.accel_res);
    g->levels = std::min(i32(u.accel_levels), ND_MAX_CASCADES);
    if (g->res <= 0 || g->levels <= 0) return nullptr;
    for (int l = 0; l < ND_MAX_CASCADES; ++l)
        for (int k = 0; k < 4; ++k) g->prm[l][k] = u.accel_params[l][k];
    const int res = g->res;
    const size_t n = (size_t)res * res * res * g->levels;
    g->safe.assign(n, 0.0);
    const P safety = P(u.accel_safety);
    const int slabs = res * g->levels;
    const int nt = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (int w = 0; w < nt; ++w) {
        pool.emplace_back([&, w] {
            for (int gz = w; gz < slabs; gz += nt) {
                const int level = gz / res, cz = gz % res;
                const P cell = P(g->prm[level][3]);
                for (int y = 0; y < res; ++y)
                    for (int x = 0; x < res; ++x) {
                        V3<P> c(P(g->prm[level][0]) + (P(x) + P(0.5)) * cell, P(g->prm[level][1]) + (P(y) + P(0.5)) * cell,
                                P(g->prm[level][2]) + (P(cz) + P(0.5)) * cell);
                        P d = scene_de(s, c);
                        P v = tmax(d - P(0.8660254) * cell, P(0)) * safety;
                        g->safe[((size_t)gz * res + y) * res + x] = (double)v;
                    }
            }
        });
    }
    for (auto& t : pool) t.join();
    return g;
}

// Attaches the accelerator grid when the Uniforms enable it (the host stamps accel_params
// from place_cascades, fr.zig:3073-3080).
template <class P> void attach_accel(const nd_uniforms& u, Scene<P>& s) {
    if (u.accel_enabled > 0.5f) s.accel = build_accel(u, s);
}

// accel_skip (template.wgsl:428-447): the first level, from the finest, whose grid holds p
// gives the skip distance; 0 outside every grid or with the accelerator off.
template <class T> T accel_skip(const AccelGrid* g, V3<T> p) {
    if (!g) return T(0);
    for (int l = 0; l < g->levels; ++l) {
        const double cell = g->prm[l][3];
        if (cell <= 0.0) continue;
        const T pc[3] = {p.x, p.y, p.z};
        int c[3];
        bool in = true;
        for (int a = 0; a < 3; ++a) {
            c[a] = (int)std::floor((double)val((pc[a] - T(g->prm[l][a])) / T(cell)));
            if (c[a] < 0 || c[a] >= g->res) in = false;
        }
        if (!in) continue;
        return T(g->safe[(((size_t)l * g->res + c[2]) * g->res + c[1]) * g->res + c[0]]);
    }
    return T(0);
}

struct MarchCfg {
    int max_steps;
    double max_dist;
    double eps_c, eps_f;
    bool hq;
    int refine_budget, bisect_budget;
    double footprint_per_unit_t;
};
inline MarchCfg march_cfg(const nd_uniforms& u) {
    // march_to preamble (template.wgsl:1859-1871) with coarse and overstep off.
    MarchCfg m;
    m.max_steps = i32(u.max_steps);
    m.max_dist = u.max_dist;
    m.eps_c = u.epsilon_coefficient;
    m.eps_f = u.epsilon_floor;
    m.hq = u.high_quality > 0.5f;
    int budget = std::max(m.hq ? i32(u.refine_hq) : i32(u.refine_fast), 0);
    bool on_hit = u.march_verify < 0.5f;
    m.refine_budget = on_hit ? 0 : budget;
    m.bisect_budget = on_hit ? budget : 0;
    m.footprint_per_unit_t = (double)u.hq_footprint_budget_px * 2.0 / std::max((double)u.resolution[1], 1.0);
    return m;
}

template <class T> struct MarchResult { bool hit; T dist; int steps; int evals; };

// march_to (template.wgsl:1859-1963) with accel_skip = 0 and warp_step_limit = infinity.
template <class T, class P>
MarchResult<T> march_to(const Scene<P>& s, const MarchCfg& m, V3<T> origin, V3<T> dir, T max_t) {
    const bool on_hit = m.refine_budget == 0;
    T t = T(0);
    int i = 0;
    bool probing = true, bisecting = false, resolved = false;
    T lo = T(0), hi = T(0);
    int refines = 0, evals = 0;
    const int max_samples = m.max_steps * (1 + m.refine_budget) + m.bisect_budget;
    for (int n = 0; n < max_samples; ++n) {
        T sample_t = probing ? t : T(0.5) * (lo + hi);
        T eps = T(m.eps_c) * sample_t + T(m.eps_f);
        if (probing && s.accel) {
            // accelerator skip (template.wgsl:1890-1900): a step without a DE evaluation
            T skip = accel_skip(s.accel.get(), origin + dir * sample_t);
            if (skip > eps) {
                lo = sample_t;
                t = sample_t + skip;
                i++;
                if (t > max_t || i >= m.max_steps) break;
                continue;
            }
        }
        T d = probed_de(s, origin + dir * sample_t);
        evals++;
        if (bisecting) {
            if (d < eps) hi = sample_t; else lo = sample_t;
            refines++;
            if (hi - lo <= eps || refines >= m.bisect_budget) return {true, hi, i, evals};
            continue;
        }
        if (d < eps) {
            if (on_hit && m.bisect_budget > 0 && sample_t - lo > eps) {
                bisecting = true;
                probing = false;
                hi = sample_t;
                refines = 0;
                continue;
            }
            return {true, sample_t, i, evals};
        }
        if (probing) {
            lo = t;
            hi = t + d;  // min(d, step_lim) with step_lim = infinity
            refines = 0;
            resolved = false;
            probing = false;
        } else {
            if (d >= hi - sample_t) resolved = true; else hi = sample_t;
            refines++;
        }
        if (resolved || refines >= m.refine_budget) {
            if (m.hq && !resolved && m.refine_budget > 0)
                hi = tmin(hi, lo + tmax(t, T(1)) * T(m.footprint_per_unit_t));
            t = hi;
            i++;
            if (t > max_t || i >= m.max_steps) break;
            probing = true;
        }
    }
    return {false, t, m.max_steps, evals};
}

// ----------------------------------------------------------------------------- normals
// tetra_gradient / estimate_normal (template.wgsl:1781-1804), e = eps(t)
template <class T, class P> V3<T> normal_tetra(const Scene<P>& s, V3<T> pos, T e) {
    const double K[4][3] = {{1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}, {1, 1, 1}};
    V3<T> g;
    for (int i = 0; i < 4; ++i) {
        V3<T> k(T(K[i][0]), T(K[i][1]), T(K[i][2]));
        g = g + k * scene_de(s, pos + k * e);
    }
    return normalize(g);
}
// 6-tap central difference with step h
template <class T, class P> V3<T> normal_central(const Scene<P>& s, V3<T> pos, T h) {
    V3<T> ex(h, T(0), T(0)), ey(T(0), h, T(0)), ez(T(0), T(0), h);
    V3<T> g(scene_de(s, pos + ex) - scene_de(s, pos - ex), scene_de(s, pos + ey) - scene_de(s, pos - ey),
            scene_de(s, pos + ez) - scene_de(s, pos - ez));
    return normalize(g);
}
// analytic gradient of the DE (forward mode, 3 tangents), normalised
template <class P> V3<double> normal_analytic(const Scene<P>& s, V3<double> pos) {
    using D = Dual<3>;
    V3<D> q(D::var(pos.x, 0), D::var(pos.y, 1), D::var(pos.z, 2));
    const Scene<D> sd = recast_scene<D>(s);
    D d = scene_de(sd, q);
    return normalize(V3<double>(d.d[0], d.d[1], d.d[2]));
}

// ----------------------------------------------------------------------------- shadow
// light_sample for point/global lights (template.wgsl:2178-2190)
template <class T> void light_dir_dist(const nd_light& L, double max_dist, V3<T> surf, V3<T>* dir, T* dist) {
    int type = i32(L.light_type + 0.5f);
    if (type == 0) {
        V3<T> to = V3<T>::from(L.position_or_direction) - surf;
        T dd = length(to);
        *dir = to / tmax(dd, T(EPSILON));
        *dist = dd;
    } else {
        *dir = normalize(V3<T>::from(L.position_or_direction));
        *dist = T(max_dist * 2.0);
    }
}
// visibility_march, opaque-only scene (template.wgsl:2282-2342), shadow_factor constants
// (template.wgsl:2344-2375). Returns clamp(res, 0, 1); counts evals.
template <class T, class P>
T visibility_march_opaque(const Scene<P>& s, V3<T> origin, V3<T> dir, T light_dist, T k, int steps, int* evals) {
    T t = T(0.004);  // SHADOW_BIAS as t0
    T res = T(1);
    for (int i = 0; i < steps; ++i) {
        if (t >= light_dist) break;
        V3<T> p = origin + dir * t;
        if (s.accel) {
            // accelerator skip (template.wgsl:2297-2303): only where it cannot lower res
            T skip = accel_skip(s.accel.get(), p);
            if (skip > T(0.001) && k * skip >= res * t) {
                t = t + skip;
                continue;
            }
        }
        T d = probed_de(s, p);
        if (evals) (*evals)++;
        if (d < T(0.0005)) return T(0);  // SHADOW_HIT_EPS
        res = tmin(res, k * d / t);
        t = t + tmax(d, T(0.001));  // SHADOW_MIN_STEP
    }
    return tclamp(res, T(0), T(1));
}

// ----------------------------------------------------------------------------- adjoint
// Hit depth as a smooth function of the parameters: the march's hit is refined to the
// root of F(t) = DE(o + t d) - (c t + f) (the level set the on-hit bisection converges to,
// template.wgsl:1924-1931) by FP64 bisection on a bracket [lo, hi] with F(lo) >= 0 > F(hi).
template <class P>
bool polished_root(const Scene<P>& s, const MarchCfg& m, V3<double> o, V3<double> d, double* t_out) {
    MarchResult<double> r = march_to(s, m, o, d, m.max_dist);
    if (!r.hit) return false;
    auto F = [&](double t) { return val(scene_de(s, o + d * t)) - (m.eps_c * t + m.eps_f); };
    double hi = r.dist;
    if (F(hi) >= 0) { *t_out = hi; return true; }
    double step = std::max(m.eps_c * hi + m.eps_f, 1e-9), lo = hi;
    int k = 0;
    for (; k < 60; ++k) {
        lo = std::max(hi - step, 0.0);
        if (F(lo) >= 0 || lo == 0.0) break;
        step *= 2;
    }
    if (F(lo) < 0) return false;
    for (int i = 0; i < 200 && hi - lo > 1e-15 * hi; ++i) {
        double mid = 0.5 * (lo + hi);
        if (F(mid) >= 0) lo = mid; else hi = mid;
    }
    *t_out = 0.5 * (lo + hi);
    return true;
}

// ---------------------------------------------------------------------------- camera
// Primary ray of pixel (px, py) of a target of size (tw, th): cs_wf_generate
// (template.wgsl:5119-5127), identical to the fragment path of vs_main + fs_simple.
template <class T> V3<T> primary_dir(const nd_uniforms& u, int px, int py, int tw, int th) {
    T fx = T((double)px + 0.5), fy = T((double)py + 0.5);
    T nx = fx / T((double)tw) * T(2) - T(1);
    T ny = T(1) - fy / T((double)th) * T(2);
    T pfx = nx * T(u.tile_scale[0]) + T(u.tile_bias[0]);
    T pfy = ny * T(u.tile_scale[1]) + T(u.tile_bias[1]);
    T aspect = T(u.resolution[0]) / T(u.resolution[1]);
    T ux = pfx * aspect, uy = -pfy;
    V3<T> f = V3<T>::from(u.camera_forward), r = V3<T>::from(u.camera_right), up = V3<T>::from(u.camera_up);
    return normalize(f + r * ux + up * uy);
}

}  // namespace ndo
