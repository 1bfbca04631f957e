// Contracted FP32 arithmetic of the Bellman backup, for the precision form.
//
// Same scheme as pg_device.hpp (forward Euler, swept probes, 64-corner
// multilinear lookup, the reference's rejection rules) evaluated the fast
// way: fused multiply-adds, reciprocal multiplies in place of divisions by
// grid constants, the hardware sin/cos, and the multilinear interpolant as a
// tensor product of 63 lerps (identical weights in exact arithmetic,
// PeregrineSolver.jl:399-466). Results agree with the exact path to FP32
// rounding, not bit for bit; the precision form is measured against the
// FP64 restatement instead. Every fused multiply-add is written out
// (fmaf), never left to the compiler's contraction, so that each kernel
// built from these functions (the in-HBM and the out-of-core instantiations)
// computes the same bits.
#pragma once
#include "pg_device.hpp"

namespace pgk {

// Launch-uniform reciprocals of the fast path.
struct FastConsts {
    float rstep[6];   // 1 / step[k]
    float rn3;        // 1 / n3
    float reps[3];    // 1 / eps[k]
};

__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

__device__ __forceinline__ void sincos_fast(float x, float& s, float& c) {
    // v_sin_f32 / v_cos_f32 take the angle in revolutions
    const float r = x * 0.15915494309189535f;
    s = __builtin_amdgcn_sinf(r);
    c = __builtin_amdgcn_cosf(r);
}

template <bool DIAG>
__device__ __forceinline__ void accel_fast(const Consts& K, const FastConsts& F, float bux, float buy,
                                           float buw, float vx, float vy, float w, float& ax,
                                           float& ay, float& al) {
    const float sx = clamp1(vx * F.reps[0]);
    const float sy = clamp1(vy * F.reps[1]);
    const float sw = clamp1(w * F.reps[2]);
    const float dx = __builtin_fabsf(vx) * vx;
    const float dy = __builtin_fabsf(vy) * vy;
    const float dw = __builtin_fabsf(w) * w;
    if (DIAG) {
        ax = fma_(K.D[0], dx, fma_(K.S[0], sx, fma_(K.A[0], vx, bux))) + K.c[0];
        ay = fma_(K.D[4], dy, fma_(K.S[4], sy, fma_(K.A[4], vy, buy))) + K.c[1];
        al = fma_(K.D[8], dw, fma_(K.S[8], sw, fma_(K.A[8], w, buw))) + K.c[2];
    } else {
        const float w2 = w * w;
        const float z[11] = {vx, vy, w, w2, sx, sy, sw, dx, dy, dw, 0.0f};
        float r[3] = {bux, buy, buw};
#pragma unroll
        for (int row = 0; row < 3; ++row) {
            const float co[10] = {K.A[3 * row], K.A[3 * row + 1], K.A[3 * row + 2], K.q[row], K.S[3 * row],
                                  K.S[3 * row + 1], K.S[3 * row + 2], K.D[3 * row], K.D[3 * row + 1],
                                  K.D[3 * row + 2]};
#pragma unroll
            for (int j = 0; j < 10; ++j) r[row] = fma_(co[j], z[j], r[row]);
            r[row] = r[row] + K.c[row];
        }
        ax = r[0]; ay = r[1]; al = r[2];
    }
}

// The state-only part of the integration. With its B*u_sat term set apart,
// the first Euler substep depends on the state alone: positions and heading
// after it (with sin/cos of that heading) and the velocities less hs*B*u_sat.
struct FastStart {
    float x0, y0, h0;          // the grid state
    float x1, y1, h1, sh1, ch1;
    float vx1, vy1, w1;
};

template <bool DIAG>
__device__ __forceinline__ void fast_start(const Consts& K, const FastConsts& F, const CellStart& c,
                                           FastStart& S) {
    const float hs = K.hs;
    float ax, ay, al;
    accel_fast<DIAG>(K, F, 0.0f, 0.0f, 0.0f, c.vx, c.vy, c.w, ax, ay, al);
    S.x0 = c.x;
    S.y0 = c.y;
    S.h0 = c.h;
    S.x1 = fma_(fma_(c.ch, c.vx, -(c.sh * c.vy)), hs, c.x);
    S.y1 = fma_(fma_(c.sh, c.vx, c.ch * c.vy), hs, c.y);
    S.h1 = fma_(c.w, hs, c.h);
    S.vx1 = fma_(fma_(c.w, c.vy, ax), hs, c.vx);
    S.vy1 = fma_(fma_(-c.w, c.vx, ay), hs, c.vy);
    S.w1 = fma_(al, hs, c.w);
    sincos_fast(S.h1, S.sh1, S.ch1);
}

// Forward Euler, all derivatives from the old state (PeregrineSolver.jl:208-221),
// from the state-only first substep. Divergence is caught by the caller's
// finiteness test on the successor.
template <bool DIAG>
__device__ __forceinline__ void euler_fast(const Consts& K, const FastConsts& F, const FastStart& S,
                                           float bux, float buy, float buw, float& x, float& y,
                                           float& h, float& vx, float& vy, float& w) {
    const float hs = K.hs;
    x = S.x1; y = S.y1; h = S.h1;
    vx = fma_(bux, hs, S.vx1);
    vy = fma_(buy, hs, S.vy1);
    w = fma_(buw, hs, S.w1);
    float sh = S.sh1, ch = S.ch1;
    for (int k = 1; k < K.nsub; ++k) {
        if (k > 1) sincos_fast(h, sh, ch);
        float ax, ay, al;
        accel_fast<DIAG>(K, F, bux, buy, buw, vx, vy, w, ax, ay, al);
        const float dvx = fma_(w, vy, ax);
        const float dvy = fma_(-w, vx, ay);
        const float px = fma_(ch, vx, -(sh * vy));
        const float py = fma_(sh, vx, ch * vy);
        x = fma_(px, hs, x);
        y = fma_(py, hs, y);
        h = fma_(w, hs, h);
        vx = fma_(dvx, hs, vx);
        vy = fma_(dvy, hs, vy);
        w = fma_(al, hs, w);
    }
}

__device__ __forceinline__ float wrap_fast(const Consts& K, const FastConsts& F, float f) {
    return fma_(-__builtin_floorf(f * F.rn3), K.n3f, f);
}

template <class OccRead>
__device__ __forceinline__ bool probes_fast(const Consts& K, const FastConsts& F, const OccRead& occ,
                                            float x, float y, float h, float dx, float dy, float dh) {
    for (int sc = 0; sc < K.nchk; ++sc) {
        const float a = K.probe_a[sc];
        const float gx = (fma_(dx, a, x) - K.lo[0]) * F.rstep[0];
        const float gy = (fma_(dy, a, y) - K.lo[1]) * F.rstep[1];
        const float gh = wrap_fast(K, F, (fma_(dh, a, h) - K.lo[2]) * F.rstep[2]);
        const bool in = (gx >= 0.0f) & (gy >= 0.0f) & (gx <= K.lim[0]) & (gy <= K.lim[1]) &
                        (gh >= 0.0f) & (gh <= K.n3f);
        if (!in) return false;
        int ck = (int)__builtin_rintf(gh);
        ck = (ck >= K.n[2]) ? ck - K.n[2] : ck;
        if (occ((int)__builtin_rintf(gx), (int)__builtin_rintf(gy), ck)) return false;
    }
    return true;
}

// The successor's fractional grid coordinates; false if rejected. With
// probe false the swept probes are skipped: a box whose probes can only
// reach free cells (the plan's flag) keeps just the final bounds test, which
// holds the probes' own (the swept points lie between the in-grid start and
// the successor).
template <bool DIAG, class OccRead>
__device__ __forceinline__ bool successor_fast(const Consts& K, const FastConsts& F,
                                               const OccRead& og, const FastStart& S, float bux,
                                               float buy, float buw, bool probe, float f[6]) {
    float nx, ny, nh, nvx, nvy, nw;
    euler_fast<DIAG>(K, F, S, bux, buy, buw, nx, ny, nh, nvx, nvy, nw);
    const float dx = nx - S.x0, dy = ny - S.y0, dh = nh - S.h0;
    if (!(finite_f(dx) & finite_f(dy) & finite_f(dh) & finite_f(nvx) & finite_f(nvy) & finite_f(nw)))
        return false;
    if (probe && !probes_fast(K, F, og, S.x0, S.y0, S.h0, dx, dy, dh)) return false;
    float snh, cnh;
    sincos_fast(nh, snh, cnh);
    const float nvfx = fma_(cnh, nvx, -(snh * nvy));
    const float nvfy = fma_(snh, nvx, cnh * nvy);
    f[0] = (nx - K.lo[0]) * F.rstep[0];
    f[1] = (ny - K.lo[1]) * F.rstep[1];
    f[2] = wrap_fast(K, F, (nh - K.lo[2]) * F.rstep[2]);
    f[3] = (nvfx - K.lo[3]) * F.rstep[3];
    f[4] = (nvfy - K.lo[4]) * F.rstep[4];
    f[5] = (nw - K.lo[5]) * F.rstep[5];
    bool in = true;
#pragma unroll
    for (int k = 0; k < 6; ++k) in = in & (f[k] >= 0.0f) & (f[k] <= K.lim[k]);
    return in;
}

// The x, y, heading grid coordinates of the successor's position (heading
// unwrapped), for every finite successor whatever the tests decide; false
// if not finite. The plan bounds the cells a box's probes can touch by them.
template <bool DIAG>
__device__ __forceinline__ bool successor_reach_fast(const Consts& K, const FastConsts& F, const FastStart& S,
                                                     float bux, float buy, float buw, float g[3]) {
    float nx, ny, nh, nvx, nvy, nw;
    euler_fast<DIAG>(K, F, S, bux, buy, buw, nx, ny, nh, nvx, nvy, nw);
    g[0] = (nx - K.lo[0]) * F.rstep[0];
    g[1] = (ny - K.lo[1]) * F.rstep[1];
    g[2] = (nh - K.lo[2]) * F.rstep[2];
    return finite_f(g[0]) & finite_f(g[1]) & finite_f(g[2]) & finite_f(nvx) & finite_f(nvy) & finite_f(nw);
}

__device__ __forceinline__ float lerp_fast(float a, float b, float t) { return fma_(t, b - a, a); }

// Multilinear interpolation as a tensor product: w first (the corner pairs
// are adjacent in memory), then vy, vx, h, y, x. vals[d] holds corner d with
// bit 0 = x ... bit 5 = w.
__device__ __forceinline__ float tensor_interp(const float t[6], const float vals[64]) {
    float l5[32];
#pragma unroll
    for (int d = 0; d < 32; ++d) l5[d] = lerp_fast(vals[d], vals[d + 32], t[5]);
    float l4[16];
#pragma unroll
    for (int d = 0; d < 16; ++d) l4[d] = lerp_fast(l5[d], l5[d + 16], t[4]);
    float l3[8];
#pragma unroll
    for (int d = 0; d < 8; ++d) l3[d] = lerp_fast(l4[d], l4[d + 8], t[3]);
    float l2[4];
#pragma unroll
    for (int d = 0; d < 4; ++d) l2[d] = lerp_fast(l3[d], l3[d + 4], t[2]);
    const float l1a = lerp_fast(l2[0], l2[2], t[1]);
    const float l1b = lerp_fast(l2[1], l2[3], t[1]);
    return lerp_fast(l1a, l1b, t[0]);
}

// The same interpolant with the corners fetched on demand: fetch(d) returns
// corner d (bit 0 = x ... bit 5 = w). Eight groups of eight corners (x, y,
// h fixed within a group) are each reduced over omega, vy, vx as they are
// loaded, then the eight results over h, y, x. The lerp tree is
// tensor_interp's, so the value is identical; only eight corner values are
// live at a time.
template <class Fetch>
__device__ __forceinline__ float tensor_interp_grouped(const Fetch& fetch, const float t[6]) {
    float r[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const float a0 = lerp_fast(fetch(c), fetch(c | 32), t[5]);
        const float a1 = lerp_fast(fetch(c | 16), fetch(c | 48), t[5]);
        const float a2 = lerp_fast(fetch(c | 8), fetch(c | 40), t[5]);
        const float a3 = lerp_fast(fetch(c | 24), fetch(c | 56), t[5]);
        const float b0 = lerp_fast(a0, a1, t[4]);
        const float b1 = lerp_fast(a2, a3, t[4]);
        r[c] = lerp_fast(b0, b1, t[3]);
    }
    const float h0 = lerp_fast(r[0], r[4], t[2]), h1 = lerp_fast(r[1], r[5], t[2]);
    const float h2 = lerp_fast(r[2], r[6], t[2]), h3 = lerp_fast(r[3], r[7], t[2]);
    const float y0 = lerp_fast(h0, h2, t[1]), y1 = lerp_fast(h1, h3, t[1]);
    return lerp_fast(y0, y1, t[0]);
}

// The same lerp tree as a loop that is not unrolled, for the rare backups
// whose corners lie outside the LDS window: eight passes, one group of
// eight corners each (x, y, h fixed), in the order c = 0, 4, 2, 6, 1, 5,
// 3, 7 (c is the bit reversal of the pass number), so that h-pairs (0, 4),
// (2, 6), (1, 5), (3, 7) and then y-pairs (0, 2), (1, 3) complete in turn.
// Every lerp has the operands and order of tensor_interp_grouped's, so the
// value is identical, while at most eight corner reads are in flight --
// unrolled, the addressed and decoded reads of all 64 corners would raise
// the register pressure of the whole backup.
template <class Fetch>
__device__ __forceinline__ float tensor_interp_rolled(const Fetch& fetch, const float t[6]) {
    float rp = 0.0f, hf = 0.0f, y0 = 0.0f, y1 = 0.0f;
#pragma unroll 1
    for (int i = 0; i < 8; ++i) {
        const int c = ((i & 1) << 2) | (i & 2) | ((i >> 2) & 1);
        const float a0 = lerp_fast(fetch(c), fetch(c | 32), t[5]);
        const float a1 = lerp_fast(fetch(c | 16), fetch(c | 48), t[5]);
        const float a2 = lerp_fast(fetch(c | 8), fetch(c | 40), t[5]);
        const float a3 = lerp_fast(fetch(c | 24), fetch(c | 56), t[5]);
        const float b0 = lerp_fast(a0, a1, t[4]);
        const float b1 = lerp_fast(a2, a3, t[4]);
        const float r = lerp_fast(b0, b1, t[3]);
        if ((i & 1) == 0) {
            rp = r;                                   // r[c] of an h-pair's first group
        } else {
            const float h = lerp_fast(rp, r, t[2]);   // h0, h2, h1, h3 in turn
            if ((i & 2) == 0) {
                hf = h;
            } else {
                const float y = lerp_fast(hf, h, t[1]);
                if (i == 3) y0 = y; else y1 = y;
            }
        }
    }
    return lerp_fast(y0, y1, t[0]);
}

}  // namespace pgk
