// FP64 arithmetic of the Bellman backup, for the FP64 tables the precision
// form is measured against and for FP64 residual checks.
//
// The algorithm is the plain scheme of solver/src/PeregrineSolver.jl carried
// in double: grid values lo + Float32(i) * step, forward Euler with every
// derivative from the old state (:208-223), the traction gain and B * u
// folded per control on the host in double, swept probes with
// round-to-nearest-even cells (:1127-1144), 64-corner multilinear lookup with
// clamped / periodic corners (:399-466), value min(tau + V, cap). Divisions are IEEE divisions
// and sin/cos the device library's double functions; no result here is meant
// to match the FP32 path bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pgk {

struct Consts64 {
    int32_t n[6];
    double lo[6];
    double step[6];
    double lim[6];     // n[k] - 1; lim[2] = n[2] (wrapped heading bound)
    double n3;
    int64_t ncell;
    int64_t stride[6];
    double A[9], q[3], S[9], D[9], c[3], eps[3];
    double hs;         // tau / nsub
    double tau;
    double cap;
    int32_t nsub;
    int32_t nchk;
    int32_t nctl;
    double probe_a[24];
};

struct Start64 {
    double x, y, h, vx, vy, w;
};

__device__ __forceinline__ double clamp1_64(double v) { return v > 1.0 ? 1.0 : (v < -1.0 ? -1.0 : v); }

__device__ __forceinline__ void cell_start64(const Consts64& K, const int i[6], Start64& c) {
    c.x = K.lo[0] + (double)(float)i[0] * K.step[0];
    c.y = K.lo[1] + (double)(float)i[1] * K.step[1];
    c.h = K.lo[2] + (double)(float)i[2] * K.step[2];
    const double vfx = K.lo[3] + (double)(float)i[3] * K.step[3];
    const double vfy = K.lo[4] + (double)(float)i[4] * K.step[4];
    c.w = K.lo[5] + (double)(float)i[5] * K.step[5];
    double sh, ch;
    sincos(c.h, &sh, &ch);
    c.vx = ch * vfx + sh * vfy;
    c.vy = -sh * vfx + ch * vfy;
}

// Left-to-right acceleration rows (:150-162) after the command terms bu.
__device__ __forceinline__ void accel64(const Consts64& K, double bux, double buy, double buw, double vx,
                                        double vy, double w, double& ax, double& ay, double& al) {
    const double w2 = w * w;
    const double sx = clamp1_64(vx / K.eps[0]);
    const double sy = clamp1_64(vy / K.eps[1]);
    const double sw = clamp1_64(w / K.eps[2]);
    const double dx = fabs(vx) * vx, dy = fabs(vy) * vy, dw = fabs(w) * w;
    const double bu[3] = {bux, buy, buw};
    double out[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        double a = bu[r];
        a = a + K.A[3 * r] * vx;
        a = a + K.A[3 * r + 1] * vy;
        a = a + K.A[3 * r + 2] * w;
        a = a + K.q[r] * w2;
        a = a + K.S[3 * r] * sx;
        a = a + K.S[3 * r + 1] * sy;
        a = a + K.S[3 * r + 2] * sw;
        a = a + K.D[3 * r] * dx;
        a = a + K.D[3 * r + 1] * dy;
        a = a + K.D[3 * r + 2] * dw;
        a = a + K.c[r];
        out[r] = a;
    }
    ax = out[0]; ay = out[1]; al = out[2];
}

// One lookahead to the successor's fractional coordinates (:1100-1153).
// Returns false when the control is rejected (value cap).
template <class OccRead>
__device__ __forceinline__ bool successor64(const Consts64& K, const OccRead& occ, const Start64& c,
                                            double bux, double buy, double buw, double f[6]) {
    double x = c.x, y = c.y, h = c.h, vx = c.vx, vy = c.vy, w = c.w;
    const double hs = K.hs;
    for (int k = 0; k < K.nsub; ++k) {
        double sum = x + y;
        sum = sum + h;
        sum = sum + vx;
        sum = sum + vy;
        sum = sum + w;
        if (!__builtin_isfinite(sum)) return false;
        double ax, ay, al;
        accel64(K, bux, buy, buw, vx, vy, w, ax, ay, al);
        const double dvx = ax + w * vy;
        const double dvy = ay - w * vx;
        double sh, ch;
        sincos(h, &sh, &ch);
        x = x + (ch * vx - sh * vy) * hs;
        y = y + (sh * vx + ch * vy) * hs;
        h = h + w * hs;
        vx = vx + dvx * hs;
        vy = vy + dvy * hs;
        w = w + al * hs;
    }
    const double dx = x - c.x, dy = y - c.y, dh = h - c.h;
    if (!(__builtin_isfinite(dx) & __builtin_isfinite(dy) & __builtin_isfinite(dh) &
          __builtin_isfinite(vx) & __builtin_isfinite(vy) & __builtin_isfinite(w)))
        return false;
    for (int sc = 0; sc < K.nchk; ++sc) {
        const double a = K.probe_a[sc];
        const double gx = (c.x + dx * a - K.lo[0]) / K.step[0];
        const double gy = (c.y + dy * a - K.lo[1]) / K.step[1];
        if (gx < 0.0 || gy < 0.0 || gx > K.lim[0] || gy > K.lim[1]) return false;
        double gh = (c.h + dh * a - K.lo[2]) / K.step[2];
        gh = gh - floor(gh / K.n3) * K.n3;
        if (!(gh >= 0.0 && gh <= K.n3)) return false;
        int ck = (int)rint(gh);
        ck = ck >= K.n[2] ? ck - K.n[2] : ck;
        if (occ((int)rint(gx), (int)rint(gy), ck)) return false;
    }
    double snh, cnh;
    sincos(h, &snh, &cnh);
    const double nvfx = cnh * vx - snh * vy;
    const double nvfy = snh * vx + cnh * vy;
    f[0] = (x - K.lo[0]) / K.step[0];
    f[1] = (y - K.lo[1]) / K.step[1];
    double f3 = (h - K.lo[2]) / K.step[2];
    f[2] = f3 - floor(f3 / K.n3) * K.n3;
    f[3] = (nvfx - K.lo[3]) / K.step[3];
    f[4] = (nvfy - K.lo[4]) / K.step[4];
    f[5] = (w - K.lo[5]) / K.step[5];
    bool in = true;
#pragma unroll
    for (int k = 0; k < 6; ++k) in = in & (f[k] >= 0.0) & (f[k] <= K.lim[k]);
    return in;
}

// Weights of the 64 corners (d = 0..63, bit 0 = x ... bit 5 = w) as the
// product lo[d & 7] * hi[d >> 3] of the first three axes' factors and the
// last three's (the FP64 tables need the value, not the FP32 path's exact
// left-to-right product order).
__device__ __forceinline__ void corner_weights64(const double t[6], double lo[8], double hi[8]) {
#pragma unroll
    for (int d = 0; d < 8; ++d) {
        lo[d] = ((d & 1) ? t[0] : 1.0 - t[0]) * ((d & 2) ? t[1] : 1.0 - t[1]) * ((d & 4) ? t[2] : 1.0 - t[2]);
        hi[d] = ((d & 1) ? t[3] : 1.0 - t[3]) * ((d & 2) ? t[4] : 1.0 - t[4]) * ((d & 4) ? t[5] : 1.0 - t[5]);
    }
}

}  // namespace pgk
