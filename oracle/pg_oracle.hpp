// CPU ORACLE -- TEST INFRASTRUCTURE ONLY.
//
// A from-source restatement of the value-iteration hot path of Peregrine's
// solver (Julia, solver/src/PeregrineSolver.jl and solver/src/Pipeline.jl of
// the reference tree), used as the checker for the HIP library. Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
// product never links, imports or calls it.
//
// Arithmetic contract. Templated on the value type T:
//   T = float  : follows the reference's Float32 CPU arithmetic operation by
//                operation -- same order of every sum and product, no fused
//                multiply-add (build with -ffp-contract=off), IEEE division,
//                ties-to-even rounding, truncating conversions.
//   T = double : the same algorithm carried in double (grid, model and
//                controls promoted from their float values), the FP64 truth
//                used to measure reduced-precision stores.
// sincos: Julia's Base.sincos(Float32) is a Float32 front end over double
// kernels taken from FreeBSD msun (k_sinf.c / k_cosf.c by Bruce D. Evans,
// e_rem_pio2f.c). This file evaluates those published kernels in double with
// the medium-range Cody-Waite reduction; the exact bits of Julia's reduction
// front end are parity-unpinned (no Julia here), the accuracy (<= 1 ulp vs
// the correctly rounded value) is pinned by tests/test_oracle_math.py.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include "../include/peregrine_hip.h"

namespace orc {

// ---------------------------------------------------------------------------
// sincos
// ---------------------------------------------------------------------------

// FreeBSD msun __kernel_sindf / __kernel_cosdf coefficients.
static const double kS1 = -0x15555554cbac77.0p-55;
static const double kS2 = 0x111110896efbb2.0p-59;
static const double kS3 = -0x1a00f9e2cae774.0p-65;
static const double kS4 = 0x16cd878c3b46a7.0p-71;
static const double kC0 = -0x1ffffffd0c5e81.0p-54;
static const double kC1 = 0x155553e1053a42.0p-57;
static const double kC2 = -0x16c087e80f1e27.0p-62;
static const double kC3 = 0x199342e0ee5069.0p-68;
// e_rem_pio2f.c medium-range constants: 2/pi, first 33 bits of pi/2, tail.
static const double kInvPio2 = 6.36619772367581382433e-01;
static const double kPio2_1 = 1.57079631090164184570e+00;
static const double kPio2_1t = 1.58932547735281966916e-08;

inline float ksin(double x) {
    double z = x * x;
    double w = z * z;
    double r = kS3 + z * kS4;
    double s = z * x;
    return (float)((x + s * (kS1 + z * kS2)) + s * w * r);
}

inline float kcos(double x) {
    double z = x * x;
    double w = z * z;
    double r = kC2 + z * kC3;
    return (float)(((1.0 + z * kC0) + w * kC1) + (w * z) * r);
}

// Julia: sincos(x::Float32) (base/special/trig.jl). Non-finite input gives
// NaN (the reference never reaches sincos with one: `diverged` guards it,
// PeregrineSolver.jl:191-192).
inline void sincos_f32(float x, float& s, float& c) {
    if (!std::isfinite(x)) {
        s = c = std::numeric_limits<float>::quiet_NaN();
        return;
    }
    const float pio4f = 3.14159274101257324f / 4.0f;  // Float32(pi)/4
    if (std::fabs(x) < pio4f) {
        if (x == 0.0f) { s = x; c = 1.0f; return; }
        double xd = (double)x;
        s = ksin(xd);
        c = kcos(xd);
        return;
    }
    double xd = (double)x;
    double fn = std::nearbyint(xd * kInvPio2);      // ties-to-even
    double r = xd - fn * kPio2_1;
    double w = fn * kPio2_1t;
    double y = r - w;
    int64_t n = (int64_t)fn;
    float si = ksin(y), co = kcos(y);
    switch (n & 3) {
        case 0: s = si; c = co; break;
        case 1: s = co; c = -si; break;
        case 2: s = -si; c = -co; break;
        default: s = -co; c = si; break;
    }
}

inline void sincos_t(float x, float& s, float& c) { sincos_f32(x, s, c); }
inline void sincos_t(double x, double& s, double& c) { s = std::sin(x); c = std::cos(x); }

// ---------------------------------------------------------------------------
// Julia scalar semantics
// ---------------------------------------------------------------------------

template <class T> inline bool isfin(T v) { return std::isfinite(v); }
// Julia min/max for non-NaN operands (every operand here is non-NaN).
template <class T> inline T jmin(T a, T b) { return (a < b) ? a : b; }
template <class T> inline T jmax(T a, T b) { return (a > b) ? a : b; }
// clamp(x, lo, hi) = ifelse(x > hi, hi, ifelse(x < lo, lo, x))
template <class T> inline T jclamp(T x, T lo, T hi) { return x > hi ? hi : (x < lo ? lo : x); }
// round(x) with RoundNearest: ties to even.
template <class T> inline T jround(T x) { return std::nearbyint(x); }

// ---------------------------------------------------------------------------
// Grid (PeregrineSolver.jl:47-81)
// ---------------------------------------------------------------------------

inline int64_t ncells(const pg_grid& g) {
    int64_t n = 1;
    for (int k = 0; k < 6; ++k) n *= (int64_t)g.n[k];
    return n;
}

// axisvalue(g, k, i) = g.lo[k] + Float32(i) * g.step[k]   (:65)
template <class T> inline T axisvalue(const pg_grid& g, int k, int32_t i) {
    return (T)g.lo[k] + (T)(float)i * (T)g.step[k];
}

// unflatten (:68-76)
inline void unflatten(const pg_grid& g, int64_t idx, int32_t i[6]) {
    int64_t r = idx;
    for (int k = 5; k >= 1; --k) {
        i[k] = (int32_t)(r % (int64_t)g.n[k]);
        r /= (int64_t)g.n[k];
    }
    i[0] = (int32_t)r;
}

// flatten (:78-81)
inline int64_t flatten(const pg_grid& g, int32_t i1, int32_t i2, int32_t i3, int32_t i4,
                       int32_t i5, int32_t i6) {
    return (((((int64_t)i1 * g.n[1] + i2) * g.n[2] + i3) * g.n[3] + i4) * g.n[4] + i5) *
               g.n[5] + i6;
}

// ---------------------------------------------------------------------------
// Drivetrain model (PeregrineSolver.jl:100-171)
// ---------------------------------------------------------------------------

// tanh of a non-negative argument. The reference's CPU and GPU builds take
// tanhf from different libraries (Julia's, CUDA's), which round
// differently, so its float bits are implementation-defined. The oracle
// uses the correctly rounded value: tanh computed in double from exp
// (argument reduction t = k ln2 + r with ln2 in two parts, Taylor series
// of exp(r); expm1 of small arguments from its series), tanh(x) =
// expm1(2x) / (expm1(2x) + 2), rounded to float. Pinned against the C
// library's double tanh, rounded to float, over every float in [1e-6, 30].
inline double tanh_series(double x) {
    static const double inv_fact[18] = {1.0, 1.0, 1.0 / 2.0, 1.0 / 6.0, 1.0 / 24.0, 1.0 / 120.0, 1.0 / 720.0,
                                        1.0 / 5040.0, 1.0 / 40320.0, 1.0 / 362880.0, 1.0 / 3628800.0,
                                        1.0 / 39916800.0, 1.0 / 479001600.0, 1.0 / 6227020800.0,
                                        1.0 / 87178291200.0, 1.0 / 1307674368000.0, 1.0 / 20922789888000.0,
                                        1.0 / 355687428096000.0};
    const double t = 2.0 * std::min(x, 20.0);
    double em;
    if (t < 0.5) {
        double p = inv_fact[17];
        for (int n = 16; n >= 1; --n) p = inv_fact[n] + t * p;
        em = t * p;
    } else {
        const double k = std::rint(t * 1.4426950408889634);
        const double r = (t - k * 6.93147180369123816490e-01) - k * 1.90821492927058770002e-10;
        double p = inv_fact[14];
        for (int n = 13; n >= 0; --n) p = inv_fact[n] + r * p;
        em = std::ldexp(p, (int)k) - 1.0;
    }
    return em / (em + 2.0);
}
inline float tanh_t(float r) { return (float)tanh_series((double)r); }
inline double tanh_t(double r) { return std::tanh(r); }

// traction_gain (:126-136)
template <class T> inline T traction_gain(T u1, T u2, T u3, float knee) {
    if (std::isnan(knee) || knee <= 0.0f) return (T)1;
    T m = std::sqrt(u1 * u1 + u2 * u2 + u3 * u3);
    T r = m / (T)knee;
    if (r < (T)1.0e-6f) return (T)1;
    return tanh_t(r) / r;
}

template <class T> inline void saturate(const pg_model& m, T u1r, T u2r, T u3r, T& u1, T& u2, T& u3) {
    T gn = traction_gain<T>(u1r, u2r, u3r, m.knee);
    u1 = u1r * gn; u2 = u2r * gn; u3 = u3r * gn;
}

// csign(v, e) = clamp(v / e, -1, 1)   (:120)
template <class T> inline T csign(T v, float e) { return jclamp(v / (T)e, (T)-1, (T)1); }

// body_accel_sat (:150-162): each row is one left-to-right sum of 14 terms.
template <class T>
inline void body_accel_sat(const pg_model& m, T u1, T u2, T u3, T vx, T vy, T w, T& ax, T& ay, T& al) {
    const float* B = m.B; const float* A = m.A; const float* q = m.q;
    const float* S = m.S; const float* D = m.D; const float* c = m.c;
    T w2 = w * w;
    T sx = csign(vx, m.eps[0]), sy = csign(vy, m.eps[1]), sw = csign(w, m.eps[2]);
    T dx = std::fabs(vx) * vx, dy = std::fabs(vy) * vy, dw = std::fabs(w) * w;
    for (int r = 0; r < 3; ++r) {
        T acc = (T)B[3 * r] * u1;
        acc = acc + (T)B[3 * r + 1] * u2;
        acc = acc + (T)B[3 * r + 2] * u3;
        acc = acc + (T)A[3 * r] * vx;
        acc = acc + (T)A[3 * r + 1] * vy;
        acc = acc + (T)A[3 * r + 2] * w;
        acc = acc + (T)q[r] * w2;
        acc = acc + (T)S[3 * r] * sx;
        acc = acc + (T)S[3 * r + 1] * sy;
        acc = acc + (T)S[3 * r + 2] * sw;
        acc = acc + (T)D[3 * r] * dx;
        acc = acc + (T)D[3 * r + 1] * dy;
        acc = acc + (T)D[3 * r + 2] * dw;
        acc = acc + (T)c[r];
        if (r == 0) ax = acc; else if (r == 1) ay = acc; else al = acc;
    }
}

// diverged (:191-192)
template <class T> inline bool diverged(T x, T y, T h, T vx, T vy, T w) {
    return !isfin(x + y + h + vx + vy + w);
}

// step_state: forward Euler, all derivatives from the old state (:208-223).
template <class T>
inline void step_state(const pg_model& m, T s[6], T u1r, T u2r, T u3r, T dt, int32_t nsub) {
    T hs = dt / (T)(float)nsub;
    T u1, u2, u3;
    saturate<T>(m, u1r, u2r, u3r, u1, u2, u3);
    T x = s[0], y = s[1], h = s[2], vx = s[3], vy = s[4], w = s[5];
    for (int32_t k = 0; k < nsub; ++k) {
        if (diverged(x, y, h, vx, vy, w)) {
            for (int j = 0; j < 6; ++j) s[j] = std::numeric_limits<T>::quiet_NaN();
            return;
        }
        T ax, ay, al;
        body_accel_sat<T>(m, u1, u2, u3, vx, vy, w, ax, ay, al);
        T dvx = ax + w * vy;
        T dvy = ay - w * vx;
        T sh, ch;
        sincos_t(h, sh, ch);
        x = x + (ch * vx - sh * vy) * hs;
        y = y + (sh * vx + ch * vy) * hs;
        h = h + w * hs;
        vx = vx + dvx * hs;
        vy = vy + dvy * hs;
        w = w + al * hs;
    }
    s[0] = x; s[1] = y; s[2] = h; s[3] = vx; s[4] = vy; s[5] = w;
}

// state_deriv (:226-236)
template <class T>
inline void state_deriv(const pg_model& m, const T s[6], T u1, T u2, T u3, T d[6]) {
    T ax, ay, al;
    body_accel_sat<T>(m, u1, u2, u3, s[3], s[4], s[5], ax, ay, al);
    T sh, ch;
    sincos_t(s[2], sh, ch);
    d[0] = ch * s[3] - sh * s[4];
    d[1] = sh * s[3] + ch * s[4];
    d[2] = s[5];
    d[3] = ax + s[5] * s[4];
    d[4] = ay - s[5] * s[3];
    d[5] = al;
}

// step_state_rk2: explicit midpoint (:259-279).
template <class T>
inline void step_state_rk2(const pg_model& m, T s[6], T u1r, T u2r, T u3r, T dt, int32_t nsub) {
    T hs = dt / (T)(float)nsub;
    T hh = (T)0.5f * hs;
    T u1, u2, u3;
    saturate<T>(m, u1r, u2r, u3r, u1, u2, u3);
    for (int32_t k = 0; k < nsub; ++k) {
        if (diverged(s[0], s[1], s[2], s[3], s[4], s[5])) {
            for (int j = 0; j < 6; ++j) s[j] = std::numeric_limits<T>::quiet_NaN();
            return;
        }
        T d1[6], mid[6], d2[6];
        state_deriv<T>(m, s, u1, u2, u3, d1);
        for (int j = 0; j < 6; ++j) mid[j] = s[j] + d1[j] * hh;
        if (diverged(mid[0], mid[1], mid[2], mid[3], mid[4], mid[5])) {
            for (int j = 0; j < 6; ++j) s[j] = std::numeric_limits<T>::quiet_NaN();
            return;
        }
        state_deriv<T>(m, mid, u1, u2, u3, d2);
        for (int j = 0; j < 6; ++j) s[j] = s[j] + d2[j] * hs;
    }
}

// ---------------------------------------------------------------------------
// Value reads: plain V, or the escape pass's view of it (:756-882)
// ---------------------------------------------------------------------------

template <class T> struct PlainV {
    const T* V;
    inline T operator()(int64_t i) const { return V[i]; }
};

// EscapeView getindex (:836-842): obstacle corners read cap unless the
// updating cell is itself blocked; stored negatives read as their escape
// time, terminal cells as zero, the rest as cap.
template <class T> struct EscV {
    const T* V;
    const uint8_t* occ;
    int64_t blk;
    T cap;
    bool ignore;
    inline T operator()(int64_t i) const {
        if (!ignore && occ[i / blk]) return cap;
        T v = V[i];
        return v < (T)0 ? -v : (v < cap ? (T)0 : cap);
    }
};

// ---------------------------------------------------------------------------
// Interpolation (:399-466, :571-665)
// ---------------------------------------------------------------------------

template <class T, class VR>
inline T interp(const VR& V, const pg_grid& g, T x, T y, T h, T vx, T vy, T w, bool nearest,
                T cap, bool vclamp) {
    T f1 = (x - (T)g.lo[0]) / (T)g.step[0];
    T f2 = (y - (T)g.lo[1]) / (T)g.step[1];
    T n3 = (T)(float)g.n[2];
    T f3 = (h - (T)g.lo[2]) / (T)g.step[2];
    f3 = f3 - std::floor(f3 / n3) * n3;
    T f4 = (vx - (T)g.lo[3]) / (T)g.step[3];
    T f5 = (vy - (T)g.lo[4]) / (T)g.step[4];
    T f6 = (w - (T)g.lo[5]) / (T)g.step[5];
    if (!(isfin(f1) && isfin(f2) && isfin(f3) && isfin(f4) && isfin(f5) && isfin(f6))) return cap;
    // A finite wrapped heading lies in [0, n3] for every heading the
    // reference indexes correctly; outside it the reference would read a
    // corner off the heading axis. Such a lookup rejects.
    if (!(f3 >= (T)0 && f3 <= n3)) return cap;
    if (f1 < (T)0 || f1 > (T)(float)(g.n[0] - 1)) return cap;
    if (f2 < (T)0 || f2 > (T)(float)(g.n[1] - 1)) return cap;
    if (vclamp) {
        f4 = jclamp(f4, (T)0, (T)(float)(g.n[3] - 1));
        f5 = jclamp(f5, (T)0, (T)(float)(g.n[4] - 1));
        f6 = jclamp(f6, (T)0, (T)(float)(g.n[5] - 1));
    } else {
        if (f4 < (T)0 || f4 > (T)(float)(g.n[3] - 1)) return cap;
        if (f5 < (T)0 || f5 > (T)(float)(g.n[4] - 1)) return cap;
        if (f6 < (T)0 || f6 > (T)(float)(g.n[5] - 1)) return cap;
    }
    if (nearest) {
        int32_t i1 = (int32_t)jround(f1), i2 = (int32_t)jround(f2);
        int32_t i3 = (int32_t)jround(f3) % g.n[2];
        int32_t i4 = (int32_t)jround(f4), i5 = (int32_t)jround(f5), i6 = (int32_t)jround(f6);
        return V(flatten(g, i1, i2, i3, i4, i5, i6));
    }
    T b[6] = {std::floor(f1), std::floor(f2), std::floor(f3),
              std::floor(f4), std::floor(f5), std::floor(f6)};
    T f[6] = {f1, f2, f3, f4, f5, f6};
    T t[6];
    int32_t i[6];
    for (int k = 0; k < 6; ++k) { t[k] = f[k] - b[k]; i[k] = (int32_t)b[k]; }
    T acc = (T)0;
    for (int d = 0; d < 64; ++d) {
        int c[6];
        for (int k = 0; k < 6; ++k) c[k] = (d >> k) & 1;
        // weight = product over the six axes, left to right (:448-453)
        T wgt = (c[0] ? t[0] : (T)1 - t[0]);
        for (int k = 1; k < 6; ++k) wgt = wgt * (c[k] ? t[k] : (T)1 - t[k]);
        if (wgt == (T)0) continue;
        int32_t j1 = jmin<int32_t>(i[0] + c[0], g.n[0] - 1);
        int32_t j2 = jmin<int32_t>(i[1] + c[1], g.n[1] - 1);
        int32_t j3 = (i[2] + c[2]) % g.n[2];
        int32_t j4 = jmin<int32_t>(i[3] + c[3], g.n[3] - 1);
        int32_t j5 = jmin<int32_t>(i[4] + c[4], g.n[4] - 1);
        int32_t j6 = jmin<int32_t>(i[5] + c[5], g.n[5] - 1);
        acc = acc + wgt * V(flatten(g, j1, j2, j3, j4, j5, j6));
    }
    return jmin(acc, cap);
}

// Kuhn / Freudenthal simplex interpolation with gradient (:571-665).
template <class T>
inline void cmpx(T& ta, int32_t& ka, T& tb, int32_t& kb) {
    if (!(ta >= tb)) { T t = ta; ta = tb; tb = t; int32_t k = ka; ka = kb; kb = k; }
}

template <class T, class VR>
inline T interp_kuhn(const VR& V, const pg_grid& g, T x, T y, T h, T vx, T vy, T w, T cap,
                     bool vclamp, T grad[6]) {
    for (int k = 0; k < 6; ++k) grad[k] = (T)0;
    T f1 = (x - (T)g.lo[0]) / (T)g.step[0];
    T f2 = (y - (T)g.lo[1]) / (T)g.step[1];
    T n3 = (T)(float)g.n[2];
    T f3 = (h - (T)g.lo[2]) / (T)g.step[2];
    f3 = f3 - std::floor(f3 / n3) * n3;
    T f4 = (vx - (T)g.lo[3]) / (T)g.step[3];
    T f5 = (vy - (T)g.lo[4]) / (T)g.step[4];
    T f6 = (w - (T)g.lo[5]) / (T)g.step[5];
    if (!(isfin(f1) && isfin(f2) && isfin(f3) && isfin(f4) && isfin(f5) && isfin(f6))) return cap;
    if (f1 < (T)0 || f1 > (T)(float)(g.n[0] - 1)) return cap;
    if (f2 < (T)0 || f2 > (T)(float)(g.n[1] - 1)) return cap;
    if (vclamp) {
        f4 = jclamp(f4, (T)0, (T)(float)(g.n[3] - 1));
        f5 = jclamp(f5, (T)0, (T)(float)(g.n[4] - 1));
        f6 = jclamp(f6, (T)0, (T)(float)(g.n[5] - 1));
    } else {
        if (f4 < (T)0 || f4 > (T)(float)(g.n[3] - 1)) return cap;
        if (f5 < (T)0 || f5 > (T)(float)(g.n[4] - 1)) return cap;
        if (f6 < (T)0 || f6 > (T)(float)(g.n[5] - 1)) return cap;
    }
    T f[6] = {f1, f2, f3, f4, f5, f6};
    T b[6];
    int32_t ii[6];
    for (int k = 0; k < 6; ++k) { b[k] = std::floor(f[k]); ii[k] = (int32_t)b[k]; }
    ii[2] = ii[2] % g.n[2];
    T s[6];
    int32_t a[6];
    for (int k = 0; k < 6; ++k) { s[k] = f[k] - b[k]; a[k] = k + 1; }
    // 12-comparator network, in the reference's order (:475-489)
    cmpx(s[0], a[0], s[5], a[5]); cmpx(s[1], a[1], s[3], a[3]); cmpx(s[2], a[2], s[4], a[4]);
    cmpx(s[1], a[1], s[2], a[2]); cmpx(s[3], a[3], s[4], a[4]); cmpx(s[0], a[0], s[3], a[3]);
    cmpx(s[2], a[2], s[5], a[5]); cmpx(s[0], a[0], s[1], a[1]); cmpx(s[2], a[2], s[3], a[3]);
    cmpx(s[4], a[4], s[5], a[5]); cmpx(s[1], a[1], s[2], a[2]); cmpx(s[3], a[3], s[4], a[4]);
    T v0 = V(flatten(g, ii[0], ii[1], ii[2], ii[3], ii[4], ii[5]));
    T acc = ((T)1 - s[0]) * v0;
    T prev = v0;
    T e[6];
    for (int k = 0; k < 6; ++k) {
        int ax = a[k] - 1;
        if (ax == 2) ii[2] = (ii[2] + 1) % g.n[2];
        else ii[ax] = jmin<int32_t>(ii[ax] + 1, g.n[ax] - 1);
        T v = V(flatten(g, ii[0], ii[1], ii[2], ii[3], ii[4], ii[5]));
        T wk = (k < 5) ? (s[k] - s[k + 1]) : s[5];
        acc = acc + wk * v;
        e[k] = (v - prev) / (T)g.step[ax];
        prev = v;
    }
    for (int ax = 0; ax < 6; ++ax) {
        // first k in order a1..a5 matching the axis, else e6 (:640-651)
        T d = e[5];
        for (int k = 4; k >= 0; --k) if (a[k] == ax + 1) d = e[k];
        grad[ax] = d;
    }
    return jmin(acc, cap);
}

template <class T, class VR>
inline T value_at(const VR& V, const pg_grid& g, T x, T y, T h, T vx, T vy, T w, const pg_params& p) {
    if (p.simplex) {
        T gr[6];
        return interp_kuhn<T>(V, g, x, y, h, vx, vy, w, (T)p.cap, p.vclamp != 0, gr);
    }
    return interp<T>(V, g, x, y, h, vx, vy, w, p.nearest != 0, (T)p.cap, p.vclamp != 0);
}

// ---------------------------------------------------------------------------
// Horizon and substeps (:954-989)
// ---------------------------------------------------------------------------

template <class T> inline T cfl_tau(const pg_grid& g, const pg_model& m, T vfx, T vfy, T w, const pg_params& p) {
    if (p.cfl <= 0.0f) return (T)p.dt;
    const float* B = m.B;
    T ax = jmax(jmax(std::fabs((T)B[0]), std::fabs((T)B[1])), std::fabs((T)B[2]));
    T ay = jmax(jmax(std::fabs((T)B[3]), std::fabs((T)B[4])), std::fabs((T)B[5]));
    T aw = jmax(jmax(std::fabs((T)B[6]), std::fabs((T)B[7])), std::fabs((T)B[8]));
    T t = (T)p.tau_max;
    if (ax > (T)1.0e-6f) t = jmin(t, (T)p.cfl * (T)g.step[3] / ax);
    if (ay > (T)1.0e-6f) t = jmin(t, (T)p.cfl * (T)g.step[4] / ay);
    if (aw > (T)1.0e-6f) t = jmin(t, (T)p.cfl * (T)g.step[5] / aw);
    T sp = std::sqrt(vfx * vfx + vfy * vfy);
    if (sp > (T)1.0e-3f) t = jmin(t, (T)p.cfl * jmin((T)g.step[0], (T)g.step[1]) / sp);
    T aw2 = std::fabs(w);
    if (aw2 > (T)1.0e-4f) t = jmin(t, (T)p.cfl * (T)g.step[2] / aw2);
    return jclamp(t, (T)p.tau_min, (T)p.tau_max);
}

template <class T> inline int32_t substeps_for(T tau, const pg_params& p) {
    if (p.hmax <= 0.0f) return p.nsub;
    int32_t k = (int32_t)jmin(tau / (T)p.hmax, (T)1.0e4f) + 1;   // unsafe_trunc: toward zero
    return jmax<int32_t>(1, jmin<int32_t>(k, 64));
}

// ---------------------------------------------------------------------------
// lookahead (:1100-1153)
// ---------------------------------------------------------------------------

template <class T, class VR>
inline T lookahead(const VR& V, const uint8_t* occ, bool occ_ignore, const pg_grid& g,
                   const pg_model& m, T x, T y, T h, T vx, T vy, T w, T u1, T u2, T u3,
                   T tau, int32_t nsub, const pg_params& p) {
    T s[6] = {x, y, h, vx, vy, w};
    if (p.rk2) step_state_rk2<T>(m, s, u1, u2, u3, tau, nsub);
    else step_state<T>(m, s, u1, u2, u3, tau, nsub);
    T nx = s[0], ny = s[1], nh = s[2], nvx = s[3], nvy = s[4], nw = s[5];
    T dx = nx - x, dy = ny - y, dh = nh - h;
    const T cap = (T)p.cap;
    if (!(isfin(dx) && isfin(dy) && isfin(dh) && isfin(nvx) && isfin(nvy) && isfin(nw))) return cap;

    int32_t nchk = p.checks;
    if (p.adaptive_checks) {
        T cells_pos = std::sqrt(dx * dx + dy * dy) / ((T)0.5f * jmin((T)g.step[0], (T)g.step[1]));
        T cells_rot = std::fabs(dh) / ((T)0.5f * (T)g.step[2]);
        T cells = jmin(jmax(cells_pos, cells_rot), (T)1.0e4f);
        int32_t need = (int32_t)cells + 1;
        nchk = jmax<int32_t>(p.checks, jmin<int32_t>(need, 24));
    }
    for (int32_t sc = 1; sc <= nchk; ++sc) {
        T a = (T)(float)sc / (T)(float)nchk;
        T px = x + dx * a;
        T py = y + dy * a;
        T ph = h + dh * a;
        T gx = (px - (T)g.lo[0]) / (T)g.step[0];
        T gy = (py - (T)g.lo[1]) / (T)g.step[1];
        if (gx < (T)0 || gy < (T)0 || gx > (T)(float)(g.n[0] - 1) || gy > (T)(float)(g.n[1] - 1))
            return cap;
        T n3f = (T)(float)g.n[2];
        T gh = (ph - (T)g.lo[2]) / (T)g.step[2];
        gh = gh - std::floor(gh / n3f) * n3f;
        // The wrapped heading lies in [0, n3] for every heading the reference
        // can index; outside it (non-finite or astronomically large headings)
        // the reference raises at Int32(round(gh)). Such a probe rejects.
        if (!(gh >= (T)0 && gh <= n3f)) return cap;
        int32_t ci = (int32_t)jround(gx), cj = (int32_t)jround(gy);
        int32_t ck = (int32_t)jround(gh) % g.n[2];
        if (!occ_ignore && occ[((int64_t)ci * g.n[1] + cj) * g.n[2] + ck]) return cap;
    }
    T snh, cnh;
    sincos_t(nh, snh, cnh);
    T nvfx = cnh * nvx - snh * nvy;
    T nvfy = snh * nvx + cnh * nvy;
    T nv = value_at<T>(V, g, nx, ny, nh, nvfx, nvfy, nw, p);
    return jmin(tau + nv, cap);
}

// ---------------------------------------------------------------------------
// cell_update (:1205-1344)
// ---------------------------------------------------------------------------
//
// proj_oct (:1165-1170): scale a command onto the octahedron surface
// |u1| + |u2| + |u3| = 1; the zero command stays zero.
template <class T> inline void proj_oct(T& u1, T& u2, T& u3) {
    T s = std::fabs(u1) + std::fabs(u2) + std::fabs(u3);
    if (s < (T)1.0e-6f) { u1 = u2 = u3 = (T)0; return; }
    T r = (T)1 / s;
    u1 = u1 * r; u2 = u2 * r; u3 = u3 * r;
}

// The warm-start policy store, Int8 interleaved per cell (pol_get / pol_set!,
// :728-744): a command component v is kept as round(clamp(v, -1, 1) * 127),
// ties to even, and read back as Float32(q) * Float32(1/127).
inline float pol_read(int8_t q) { return (float)q * (1.0f / 127.0f); }
inline int8_t pol_quant(float v) {
    float c = v < -1.0f ? -1.0f : (v > 1.0f ? 1.0f : v);
    return (int8_t)(int32_t)std::nearbyint(c * 127.0f);
}

// Returns the backed-up value, the index of the winning lattice control in
// *best_k (0 = coast; -1 when the warm-start or a pattern-search command
// won) and the winning command in bout[3] when bout is given. Each sweep
// evaluates coast, the cell's warm-start command (wpol, when given and
// non-zero), the lattice (whole, or this phase's rotating slice), `rounds`
// of pattern search around the best so far (six axis moves of +-delta
// projected onto the octahedron, delta halving each round), then the
// horizon ladder (ntau > 1) and the dt rung (cfl > 0) with the best
// command. Strict < throughout: the earlier candidate wins ties.
template <class T>
inline T cell_update(int64_t idx, const T* V, const uint8_t* occ, const pg_grid& g,
                     const pg_model& m, const float* ctl, int32_t nctl, const pg_params& p,
                     int32_t phase, bool esc, int32_t* best_k, const int8_t* wpol = nullptr,
                     T* bout = nullptr) {
    int32_t i[6];
    unflatten(g, idx, i);
    bool blocked = occ[((int64_t)i[0] * g.n[1] + i[1]) * g.n[2] + i[2]] != 0;
    const T cap = (T)p.cap;
    *best_k = 0;
    if (bout) bout[0] = bout[1] = bout[2] = (T)0;
    if (!esc && blocked) return cap;
    T x = axisvalue<T>(g, 0, i[0]), y = axisvalue<T>(g, 1, i[1]);
    T h = axisvalue<T>(g, 2, i[2]), vfx = axisvalue<T>(g, 3, i[3]);
    T vfy = axisvalue<T>(g, 4, i[4]), w = axisvalue<T>(g, 5, i[5]);
    T sh0, ch0;
    sincos_t(h, sh0, ch0);
    T vx = ch0 * vfx + sh0 * vfy;
    T vy = -sh0 * vfx + ch0 * vfy;
    T tau0 = cfl_tau<T>(g, m, vfx, vfy, w, p);
    int32_t ns0 = substeps_for<T>(tau0, p);

    int64_t blk = (int64_t)g.n[3] * g.n[4] * g.n[5];
    auto run = [&](T u1, T u2, T u3, T tau, int32_t ns) -> T {
        if (esc) {
            EscV<T> ev{V, occ, blk, cap, blocked};
            return lookahead<T>(ev, occ, blocked, g, m, x, y, h, vx, vy, w, u1, u2, u3, tau, ns, p);
        }
        PlainV<T> pv{V};
        return lookahead<T>(pv, occ, false, g, m, x, y, h, vx, vy, w, u1, u2, u3, tau, ns, p);
    };

    T best = run((T)0, (T)0, (T)0, tau0, ns0);
    T b1 = (T)0, b2 = (T)0, b3 = (T)0;
    if (wpol) {
        const int8_t* q = wpol + 3 * idx;
        T q1 = (T)pol_read(q[0]), q2 = (T)pol_read(q[1]), q3 = (T)pol_read(q[2]);
        if (std::fabs(q1) + std::fabs(q2) + std::fabs(q3) > (T)1.0e-6f) {
            T c = run(q1, q2, q3, tau0, ns0);
            if (c < best) { best = c; b1 = q1; b2 = q2; b3 = q3; *best_k = -1; }
        }
    }
    int32_t nnc = jmax<int32_t>(nctl - 1, 1);
    int32_t span = p.ncoarse <= 0 ? nnc : jmin<int32_t>(p.ncoarse, nnc);
    int32_t off = p.ncoarse <= 0 ? 0 : (int32_t)(((int64_t)phase * span) % nnc + nnc) % nnc;
    for (int32_t j = 0; j < span; ++j) {
        int32_t k = (int32_t)(((int64_t)off + j) % nnc) + 1;   // 0-based, skips coast
        T u1 = (T)ctl[k], u2 = (T)ctl[k + nctl], u3 = (T)ctl[k + 2 * nctl];
        T c = run(u1, u2, u3, tau0, ns0);
        if (c < best) { best = c; *best_k = k; b1 = u1; b2 = u2; b3 = u3; }
    }
    T delta = (T)p.delta0;
    for (int32_t rd = 0; rd < p.rounds; ++rd) {
        for (int32_t d = 1; d <= 6; ++d) {
            T sg = (d % 2 == 0) ? -delta : delta;
            T e1 = d <= 2 ? sg : (T)0;
            T e2 = (d == 3 || d == 4) ? sg : (T)0;
            T e3 = d >= 5 ? sg : (T)0;
            T c1 = b1 + e1, c2 = b2 + e2, c3 = b3 + e3;
            proj_oct<T>(c1, c2, c3);
            if (std::fabs(c1) + std::fabs(c2) + std::fabs(c3) < (T)1.0e-6f) continue;
            T c = run(c1, c2, c3, tau0, ns0);
            if (c < best) { best = c; b1 = c1; b2 = c2; b3 = c3; *best_k = -1; }
        }
        delta = delta * (T)0.5f;
    }
    if (p.ntau > 1) {
        T tau = tau0 * (T)0.5f;
        for (int32_t k = 1; k <= p.ntau; ++k) {
            if (k != 2) {   // rung 2 is tau0 at ratio 2, already evaluated
                T t = jmin(jmax(tau, (T)p.tau_min), (T)p.tau_max);
                T c = run(b1, b2, b3, t, substeps_for<T>(t, p));
                if (c < best) best = c;
            }
            tau = tau * (T)p.tau_ratio;
        }
        if (p.cfl > 0.0f) {
            T t = jmin(jmax((T)p.dt, (T)p.tau_min), (T)p.tau_max);
            T c = run(b1, b2, b3, t, substeps_for<T>(t, p));
            if (c < best) best = c;
        }
    }
    if (bout) { bout[0] = b1; bout[1] = b2; bout[2] = b3; }
    return best;
}

// ---------------------------------------------------------------------------
// Adjoint support (no counterpart in the reference): the successor's
// derivative with respect to one drivetrain parameter, and the gradient of
// the multilinear interpolant. Parameter index j of theta (30 entries):
// B (0-8, row-major), A (9-17), q (18-20), S diagonal (21-23), D diagonal
// (24-26), c (27-29), as in body_accel_sat (:150-162).
// ---------------------------------------------------------------------------

// d a_r / d theta_j holding the state fixed.
inline double accel_param_partial(int j, int r, double u1, double u2, double u3, double vx, double vy,
                                  double w, const pg_model& m) {
    const double v[3] = {vx, vy, w};
    const double u[3] = {u1, u2, u3};
    if (j < 9) return (j / 3 == r) ? u[j % 3] : 0.0;
    if (j < 18) return ((j - 9) / 3 == r) ? v[(j - 9) % 3] : 0.0;
    if (j < 21) return (j - 18 == r) ? w * w : 0.0;
    if (j < 24) return (j - 21 == r) ? csign<double>(v[r], m.eps[r]) : 0.0;
    if (j < 27) return (j - 24 == r) ? std::fabs(v[r]) * v[r] : 0.0;
    return (j - 27 == r) ? 1.0 : 0.0;
}

// Forward Euler (:208-223) in double from s[6] under the raw command u,
// carrying the tangent ds[6] = d state / d theta_j (ds starts at 0).
inline void step_state_tangent(const pg_model& m, double s[6], double ds[6], double u1r, double u2r,
                               double u3r, double dt, int32_t nsub, int j) {
    const double hs = dt / (double)(float)nsub;
    double u1, u2, u3;
    saturate<double>(m, u1r, u2r, u3r, u1, u2, u3);
    for (int k = 0; k < 6; ++k) ds[k] = 0.0;
    for (int32_t k = 0; k < nsub; ++k) {
        const double h = s[2], vx = s[3], vy = s[4], w = s[5];
        const double dh = ds[2], dvx = ds[3], dvy = ds[4], dw = ds[5];
        double ax, ay, al;
        body_accel_sat<double>(m, u1, u2, u3, vx, vy, w, ax, ay, al);
        // tangent of the acceleration rows: state part + explicit part
        const double v[3] = {vx, vy, w}, dv[3] = {dvx, dvy, dw};
        double dcs[3], dqd[3];
        for (int c = 0; c < 3; ++c) {
            const double e = (double)m.eps[c];
            dcs[c] = std::fabs(v[c] / e) < 1.0 ? dv[c] / e : 0.0;   // d csign
            dqd[c] = 2.0 * std::fabs(v[c]) * dv[c];                   // d |v| v
        }
        double da[3];
        for (int r = 0; r < 3; ++r) {
            double t = 0.0;
            for (int c = 0; c < 3; ++c)
                t += (double)m.A[3 * r + c] * dv[c] + (double)m.S[3 * r + c] * dcs[c] +
                     (double)m.D[3 * r + c] * dqd[c];
            t += 2.0 * (double)m.q[r] * w * dw;
            da[r] = t + accel_param_partial(j, r, u1, u2, u3, vx, vy, w, m);
        }
        double sh, ch;
        sincos_t(h, sh, ch);
        const double a1 = dvx - vy * dh, a2 = dvy + vx * dh;
        const double dF[6] = {ch * a1 - sh * a2, sh * a1 + ch * a2, dw,
                              da[0] + dw * vy + w * dvy, da[1] - dw * vx - w * dvx, da[2]};
        const double F[6] = {ch * vx - sh * vy, sh * vx + ch * vy, w, ax + w * vy, ay - w * vx, al};
        for (int q = 0; q < 6; ++q) {
            s[q] = s[q] + F[q] * hs;
            ds[q] = ds[q] + dF[q] * hs;
        }
    }
}

// Multilinear interpolant's value and gradient with respect to the
// fractional coordinates f (cells), corners clamped / periodic as in interp.
template <class VR>
inline double interp_grad(const VR& V, const pg_grid& g, const double f[6], double grad[6]) {
    double t[6];
    int32_t i[6];
    for (int k = 0; k < 6; ++k) {
        const double b = std::floor(f[k]);
        t[k] = f[k] - b;
        i[k] = (int32_t)b;
    }
    double acc = 0.0;
    for (int k = 0; k < 6; ++k) grad[k] = 0.0;
    for (int d = 0; d < 64; ++d) {
        int32_t j[6];
        for (int k = 0; k < 6; ++k) {
            const int c = (d >> k) & 1;
            j[k] = (k == 2) ? (i[k] + c) % g.n[2] : jmin<int32_t>(i[k] + c, g.n[k] - 1);
        }
        const double v = V(flatten(g, j[0], j[1], j[2], j[3], j[4], j[5]));
        double wall = 1.0;
        for (int k = 0; k < 6; ++k) wall *= ((d >> k) & 1) ? t[k] : 1.0 - t[k];
        acc += wall * v;
        for (int a = 0; a < 6; ++a) {
            double wa = ((d >> a) & 1) ? 1.0 : -1.0;
            for (int k = 0; k < 6; ++k)
                if (k != a) wa *= ((d >> k) & 1) ? t[k] : 1.0 - t[k];
            grad[a] += wa * v;
        }
    }
    return acc;
}

}  // namespace orc
