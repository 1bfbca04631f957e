// Device arithmetic of the Bellman backup for gfx950.
//
// Every function here reproduces the reference's Float32 CPU arithmetic
// (solver/src/PeregrineSolver.jl) operation for operation: the same order
// of every sum and product and no fused multiply-add (the library is built
// with -ffp-contract=off; fmaf is never written here). Two substitutions
// keep the results bit-identical while avoiding the slow IEEE sequences:
//   * a float division a / b by a grid or model constant b is computed as
//     RN32(RN64(a) * RN64(1/b)); this equals IEEE a / b for every float a
//     (verified exhaustively per divisor by the test suite);
//   * sincos is the FreeBSD-msun double-evaluated float kernel pair with
//     Cody-Waite reduction, the same deterministic function the CPU oracle
//     evaluates (<= 1 ulp from the correctly rounded value).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pgk {

// Launch-uniform constants of one solve, passed by value.
struct Consts {
    int32_t n[6];
    float lo[6];
    float step[6];
    double rstep[6];   // RN64(1 / step[k])
    float lim[6];      // Float32(n[k] - 1); lim[2] = Float32(n[2]) (wrapped heading bound)
    float n3f;         // Float32(n[2])
    double rn3;        // RN64(1 / Float32(n[2]))
    int64_t ncell;
    int64_t stride[6]; // flat-index strides
    float A[9], q[3], S[9], D[9], c[3];
    double reps[3];    // RN64(1 / eps[k])
    float hs;          // tau / Float32(nsub), IEEE
    float tau;
    float cap;
    int32_t nsub;
    int32_t nchk;
    int32_t nctl;
    float probe_a[24]; // Float32(sc) / Float32(nchk), IEEE
};

__device__ __forceinline__ float qdiv(float a, double rb) {
    return (float)((double)a * rb);
}

// ---------------------------------------------------------------------------
// sincos: FreeBSD msun k_sinf.c / k_cosf.c kernels, e_rem_pio2f.c reduction.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void sincos_ref(float x, float& s, float& c) {
    const double S1 = -0x15555554cbac77.0p-55, S2 = 0x111110896efbb2.0p-59;
    const double S3 = -0x1a00f9e2cae774.0p-65, S4 = 0x16cd878c3b46a7.0p-71;
    const double C0 = -0x1ffffffd0c5e81.0p-54, C1 = 0x155553e1053a42.0p-57;
    const double C2 = -0x16c087e80f1e27.0p-62, C3 = 0x199342e0ee5069.0p-68;
    const double invpio2 = 6.36619772367581382433e-01;
    const double pio2_1 = 1.57079631090164184570e+00;
    const double pio2_1t = 1.58932547735281966916e-08;
    const double xd = (double)x;
    const bool small = __builtin_fabsf(x) < (3.14159274101257324f / 4.0f);
    const double fn = small ? 0.0 : __builtin_rint(xd * invpio2);
    const double r = xd - fn * pio2_1;
    const double wt = fn * pio2_1t;
    const double y = r - wt;
    const int n = (int)fn;
    const double z = y * y;
    const double w = z * z;
    const double rs = S3 + z * S4;
    const double sz = z * y;
    const float si = (float)((y + sz * (S1 + z * S2)) + sz * w * rs);
    const double rc = C2 + z * C3;
    const float co = (float)(((1.0 + z * C0) + w * C1) + (w * z) * rc);
    const int q = n & 3;
    float ss = (q == 0) ? si : (q == 1) ? co : (q == 2) ? -si : -co;
    float cc = (q == 0) ? co : (q == 1) ? -si : (q == 2) ? -co : si;
    s = (x == 0.0f) ? x : ss;
    c = (x == 0.0f) ? 1.0f : cc;
}

__device__ __forceinline__ float clamp1(float v) {
    // clamp(v, -1, 1) for non-NaN v
    return __builtin_amdgcn_fmed3f(v, -1.0f, 1.0f);
}

__device__ __forceinline__ bool finite_f(float v) {
    return __builtin_isfinite(v);
}

// ---------------------------------------------------------------------------
// Forward Euler over nsub substeps (PeregrineSolver.jl:208-223). The command
// enters through bu = B*u_sat, the first three terms of each acceleration
// row's left-to-right sum, which the host folds once per control.
// Returns false if the integration diverged (the reference returns a NaN
// state, PeregrineSolver.jl:191-192,213).
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool euler(const Consts& K, float bux, float buy, float buw,
                                      float& x, float& y, float& h, float& vx, float& vy,
                                      float& w) {
    bool ok = true;
    const float hs = K.hs;
    for (int k = 0; k < K.nsub; ++k) {
        float sum = x + y;
        sum = sum + h;
        sum = sum + vx;
        sum = sum + vy;
        sum = sum + w;
        ok = ok && finite_f(sum);
        const float w2 = w * w;
        const float sx = clamp1(qdiv(vx, K.reps[0]));
        const float sy = clamp1(qdiv(vy, K.reps[1]));
        const float sw = clamp1(qdiv(w, K.reps[2]));
        const float dx = __builtin_fabsf(vx) * vx;
        const float dy = __builtin_fabsf(vy) * vy;
        const float dw = __builtin_fabsf(w) * w;
        float ax = bux;
        ax = ax + K.A[0] * vx; ax = ax + K.A[1] * vy; ax = ax + K.A[2] * w;
        ax = ax + K.q[0] * w2;
        ax = ax + K.S[0] * sx; ax = ax + K.S[1] * sy; ax = ax + K.S[2] * sw;
        ax = ax + K.D[0] * dx; ax = ax + K.D[1] * dy; ax = ax + K.D[2] * dw;
        ax = ax + K.c[0];
        float ay = buy;
        ay = ay + K.A[3] * vx; ay = ay + K.A[4] * vy; ay = ay + K.A[5] * w;
        ay = ay + K.q[1] * w2;
        ay = ay + K.S[3] * sx; ay = ay + K.S[4] * sy; ay = ay + K.S[5] * sw;
        ay = ay + K.D[3] * dx; ay = ay + K.D[4] * dy; ay = ay + K.D[5] * dw;
        ay = ay + K.c[1];
        float al = buw;
        al = al + K.A[6] * vx; al = al + K.A[7] * vy; al = al + K.A[8] * w;
        al = al + K.q[2] * w2;
        al = al + K.S[6] * sx; al = al + K.S[7] * sy; al = al + K.S[8] * sw;
        al = al + K.D[6] * dx; al = al + K.D[7] * dy; al = al + K.D[8] * dw;
        al = al + K.c[2];
        const float dvx = ax + w * vy;
        const float dvy = ay - w * vx;
        float sh, ch;
        sincos_ref(h, sh, ch);
        const float px = ch * vx - sh * vy;
        const float py = sh * vx + ch * vy;
        x = x + px * hs;
        y = y + py * hs;
        h = h + w * hs;
        vx = vx + dvx * hs;
        vy = vy + dvy * hs;
        w = w + al * hs;
    }
    return ok;
}

// Wrapped heading coordinate f - floor(f / n3) * n3 (PeregrineSolver.jl:407-408).
__device__ __forceinline__ float wrap_h(const Consts& K, float f) {
    return f - __builtin_floorf(qdiv(f, K.rn3)) * K.n3f;
}

// Swept collision probes along the chord (PeregrineSolver.jl:1127-1144).
// Returns true if every probe is on the field and in free space. A wrapped
// heading outside [0, n3] (only reachable from a non-finite or astronomically
// large heading, where the reference raises) is rejected.
template <class OccRead>
__device__ __forceinline__ bool probes_free(const Consts& K, const OccRead& occ, float x,
                                            float y, float h, float dx, float dy, float dh) {
    bool ok = true;
    for (int sc = 0; sc < K.nchk; ++sc) {
        const float a = K.probe_a[sc];
        const float px = x + dx * a;
        const float py = y + dy * a;
        const float ph = h + dh * a;
        const float gx = qdiv(px - K.lo[0], K.rstep[0]);
        const float gy = qdiv(py - K.lo[1], K.rstep[1]);
        const float gh = wrap_h(K, qdiv(ph - K.lo[2], K.rstep[2]));
        const bool in = (gx >= 0.0f) & (gy >= 0.0f) & (gx <= K.lim[0]) & (gy <= K.lim[1]) &
                        (gh >= 0.0f) & (gh <= K.n3f);
        if (!in) { ok = false; break; }
        const int ci = (int)__builtin_rintf(gx);
        const int cj = (int)__builtin_rintf(gy);
        int ck = (int)__builtin_rintf(gh);
        ck = (ck >= K.n[2]) ? ck - K.n[2] : ck;
        if (occ(ci, cj, ck)) { ok = false; break; }
    }
    return ok;
}

}  // namespace pgk
