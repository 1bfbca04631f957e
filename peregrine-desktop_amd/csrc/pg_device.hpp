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
    int32_t rk2;       // explicit midpoint substeps instead of forward Euler
    int32_t nearest;   // read the nearest node instead of the 64-corner interpolant
    int32_t vclamp;    // clamp velocity lookups into the grid instead of rejecting them
    int32_t adaptive;  // probe count from the step length (adaptive_checks)
    float half_xy;     // 0.5 * min(step x, step y), IEEE (adaptive probe count)
    float half_h;      // 0.5 * step h, IEEE
    // the rotating slice of the control lattice (ncoarse > 0): controls
    // ((off + j) mod nnc) + 1 for j < span after coast; off from the sweep's phase
    int32_t nnc, span, off;
    int32_t simplex;   // Kuhn (7-vertex simplex) lookups instead of 64-corner multilinear
    // per-cell horizon (cfl > 0, PeregrineSolver.jl:954-989): tau from the
    // cell's speed and spin, bounded by cfl_tb (tau_max and the acceleration
    // terms, host-folded); substeps from tau / hmax
    int32_t cfl;
    float cfl_tb, cfl_xy, cfl_h, tau_min, tau_max, hmax;
    float probe_a[24]; // Float32(sc) / Float32(nchk), IEEE
    // commands evaluated on the device (warm start, pattern search, ladder;
    // PeregrineSolver.jl:1240-1343): B and the traction knee for B * u_sat,
    // the search rounds and first radius, the ladder's rungs and ratio
    float B[9];
    float knee;
    int32_t rounds, ntau;
    float delta0, tau_ratio;
};

// Control of scan position r (0 = coast, then the slice in scan order).
__device__ __forceinline__ int slice_control(const Consts& K, int r) {
    if (r == 0) return 0;
    int k = K.off + r - 1;
    k = k >= K.nnc ? k - K.nnc : k;
    return k + 1;
}

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
__device__ __forceinline__ bool euler(const Consts& K, float hs, int nsub, float bux, float buy, float buw,
                                      float& x, float& y, float& h, float& vx, float& vy,
                                      float& w) {
    bool ok = true;
    for (int k = 0; k < nsub; ++k) {
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
// GC: Consts, or a GridC copy of its grid constants.
template <class GC>
__device__ __forceinline__ float wrap_h(const GC& K, float f) {
    return f - __builtin_floorf(qdiv(f, K.rn3)) * K.n3f;
}

// Swept collision probes along the chord (PeregrineSolver.jl:1127-1144).
// Returns true if every probe is on the field and in free space. A wrapped
// heading outside [0, n3] (only reachable from a non-finite or astronomically
// large heading, where the reference raises) is rejected.
// GEN = false compiles the default scheme only (fixed probe count); the
// kernels pick it when no option is set.
template <class OccRead, bool GEN = true>
__device__ __forceinline__ bool probes_free(const Consts& K, const OccRead& occ, float x,
                                            float y, float h, float dx, float dy, float dh) {
    int nchk = K.nchk;
    const bool adaptive = GEN && K.adaptive;
    if (adaptive) {
        // enough probes that none is more than half a cell from the next
        // (PeregrineSolver.jl:1127-1133)
        const float cells_pos = sqrtf(dx * dx + dy * dy) / K.half_xy;
        const float cells_rot = __builtin_fabsf(dh) / K.half_h;
        float cells = cells_pos > cells_rot ? cells_pos : cells_rot;
        cells = cells < 1.0e4f ? cells : 1.0e4f;
        const int need = (int)cells + 1;
        const int cap24 = need < 24 ? need : 24;
        nchk = K.nchk > cap24 ? K.nchk : cap24;
    }
    bool ok = true;
    for (int sc = 0; sc < nchk; ++sc) {
        const float a = adaptive ? (float)(sc + 1) / (float)nchk : K.probe_a[sc];
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


// ---------------------------------------------------------------------------
// Acceleration rows with the state-dependent terms precomputed. DIAG = true
// drops every term whose coefficient is exactly +-0 (off-diagonal A, S, D and
// q): adding a +-0 product to a partial sum changes at most the sign of a zero
// partial sum, and no zero's sign reaches any V value (divisions are by
// nonzero constants; comparisons, floor and rounding ignore the sign of
// zero), so the V bits are those of the full left-to-right sum.
// ---------------------------------------------------------------------------
template <bool DIAG>
__device__ __forceinline__ void accel_rows(const Consts& K, float bux, float buy, float buw,
                                           float vx, float vy, float w, float w2, float sx,
                                           float sy, float sw, float dx, float dy, float dw,
                                           float& ax, float& ay, float& al) {
    if (DIAG) {
        ax = bux; ax = ax + K.A[0] * vx; ax = ax + K.S[0] * sx; ax = ax + K.D[0] * dx; ax = ax + K.c[0];
        ay = buy; ay = ay + K.A[4] * vy; ay = ay + K.S[4] * sy; ay = ay + K.D[4] * dy; ay = ay + K.c[1];
        al = buw; al = al + K.A[8] * w;  al = al + K.S[8] * sw; al = al + K.D[8] * dw; al = al + K.c[2];
    } else {
        ax = bux;
        ax = ax + K.A[0] * vx; ax = ax + K.A[1] * vy; ax = ax + K.A[2] * w;
        ax = ax + K.q[0] * w2;
        ax = ax + K.S[0] * sx; ax = ax + K.S[1] * sy; ax = ax + K.S[2] * sw;
        ax = ax + K.D[0] * dx; ax = ax + K.D[1] * dy; ax = ax + K.D[2] * dw;
        ax = ax + K.c[0];
        ay = buy;
        ay = ay + K.A[3] * vx; ay = ay + K.A[4] * vy; ay = ay + K.A[5] * w;
        ay = ay + K.q[1] * w2;
        ay = ay + K.S[3] * sx; ay = ay + K.S[4] * sy; ay = ay + K.S[5] * sw;
        ay = ay + K.D[3] * dx; ay = ay + K.D[4] * dy; ay = ay + K.D[5] * dw;
        ay = ay + K.c[1];
        al = buw;
        al = al + K.A[6] * vx; al = al + K.A[7] * vy; al = al + K.A[8] * w;
        al = al + K.q[2] * w2;
        al = al + K.S[6] * sx; al = al + K.S[7] * sy; al = al + K.S[8] * sw;
        al = al + K.D[6] * dx; al = al + K.D[7] * dy; al = al + K.D[8] * dw;
        al = al + K.c[2];
    }
}

// The per-cell quantities every lookahead of a cell starts from: the body
// state and the state-only terms of the first substep (the heading's sincos
// doubles as the field->body rotation, PeregrineSolver.jl:1234-1236).
struct CellStart {
    float x, y, h, vx, vy, w;
    float sh, ch;
    float sx, sy, sw, dx, dy, dw, w2;
    float tau, hs;     // the cell's horizon and substep
    int32_t nsub;
    // the control-independent part of the first Euler substep: position and
    // heading after it, their partial left-fold sum and the heading's sincos
    float x1, y1, h1, s1, sh1, ch1;
};

// cfl_tau (PeregrineSolver.jl:954-980) of a cell with field velocity (vfx,
// vfy) and spin w: the host-folded bound, then the speed and spin terms,
// clamped to [tau_min, tau_max].
__device__ __forceinline__ float cfl_tau_dev(const Consts& K, float vfx, float vfy, float w) {
    float t = K.cfl_tb;
    const float sp = sqrtf(vfx * vfx + vfy * vfy);
    if (sp > 1.0e-3f) {
        const float c = K.cfl_xy / sp;
        t = t < c ? t : c;
    }
    const float aw = __builtin_fabsf(w);
    if (aw > 1.0e-4f) {
        const float c = K.cfl_h / aw;
        t = t < c ? t : c;
    }
    return t > K.tau_max ? K.tau_max : (t < K.tau_min ? K.tau_min : t);
}

// substeps_for (PeregrineSolver.jl:985-989) of a horizon.
__device__ __forceinline__ int substeps_dev(const Consts& K, float tau) {
    if (K.hmax <= 0.0f) return K.nsub;
    float q = tau / K.hmax;
    q = q < 1.0e4f ? q : 1.0e4f;
    int k = (int)q + 1;
    k = k < 64 ? k : 64;
    return k > 1 ? k : 1;
}

// Position and heading after the first Euler substep, and their partial
// left-fold sum (the control does not enter them).
__device__ __forceinline__ void cell_sub1(CellStart& c) {
    const float px = c.ch * c.vx - c.sh * c.vy;
    const float py = c.sh * c.vx + c.ch * c.vy;
    c.x1 = c.x + px * c.hs;
    c.y1 = c.y + py * c.hs;
    c.h1 = c.h + c.w * c.hs;
    const float sum = c.x1 + c.y1;
    c.s1 = sum + c.h1;
}

__device__ __forceinline__ void cell_start(const Consts& K, int i0, int i1, int i2, int i3, int i4,
                                           int i5, CellStart& c) {
    c.x = K.lo[0] + (float)i0 * K.step[0];
    c.y = K.lo[1] + (float)i1 * K.step[1];
    c.h = K.lo[2] + (float)i2 * K.step[2];
    const float vfx = K.lo[3] + (float)i3 * K.step[3];
    const float vfy = K.lo[4] + (float)i4 * K.step[4];
    c.w = K.lo[5] + (float)i5 * K.step[5];
    sincos_ref(c.h, c.sh, c.ch);
    c.vx = c.ch * vfx + c.sh * vfy;
    c.vy = -c.sh * vfx + c.ch * vfy;
    c.w2 = c.w * c.w;
    c.sx = clamp1(qdiv(c.vx, K.reps[0]));
    c.sy = clamp1(qdiv(c.vy, K.reps[1]));
    c.sw = clamp1(qdiv(c.w, K.reps[2]));
    c.dx = __builtin_fabsf(c.vx) * c.vx;
    c.dy = __builtin_fabsf(c.vy) * c.vy;
    c.dw = __builtin_fabsf(c.w) * c.w;
    if (K.cfl) {
        c.tau = cfl_tau_dev(K, vfx, vfy, c.w);
        c.nsub = substeps_dev(K, c.tau);
        c.hs = c.tau / (float)c.nsub;
    } else {
        c.tau = K.tau;
        c.nsub = K.nsub;
        c.hs = K.hs;
    }
    cell_sub1(c);
    sincos_ref(c.h1, c.sh1, c.ch1);
}

// Kuhn / Freudenthal simplex interpolation (PeregrineSolver.jl:571-665): the
// fractional parts sorted, largest first, by the reference's fixed network of
// 12 comparators (:475-489); the simplex's 7 vertices are walked from the
// cell's base corner adding one axis at a time, weights (1 - s1), s1 - s2,
// ..., s6, summed in that order. corner(mask) reads the vertex whose axes in
// mask (bit a = axis a) take their upper corner.
__device__ __forceinline__ void kuhn_cmpx(float& ta, int& ka, float& tb, int& kb) {
    if (!(ta >= tb)) {
        const float t = ta; ta = tb; tb = t;
        const int k = ka; ka = kb; kb = k;
    }
}

template <class Corner>
__device__ __forceinline__ float kuhn_interp(const float f[6], const Corner& corner) {
    float s[6];
    int a[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        s[k] = f[k] - __builtin_floorf(f[k]);
        a[k] = k;
    }
    kuhn_cmpx(s[0], a[0], s[5], a[5]); kuhn_cmpx(s[1], a[1], s[3], a[3]); kuhn_cmpx(s[2], a[2], s[4], a[4]);
    kuhn_cmpx(s[1], a[1], s[2], a[2]); kuhn_cmpx(s[3], a[3], s[4], a[4]); kuhn_cmpx(s[0], a[0], s[3], a[3]);
    kuhn_cmpx(s[2], a[2], s[5], a[5]); kuhn_cmpx(s[0], a[0], s[1], a[1]); kuhn_cmpx(s[2], a[2], s[3], a[3]);
    kuhn_cmpx(s[4], a[4], s[5], a[5]); kuhn_cmpx(s[1], a[1], s[2], a[2]); kuhn_cmpx(s[3], a[3], s[4], a[4]);
    const float v0 = corner(0);
    float acc = (1.0f - s[0]) * v0;
    int mask = 0;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        mask |= 1 << a[k];
        const float v = corner(mask);
        const float wk = (k < 5) ? (s[k] - s[k + 1]) : s[5];
        acc = acc + wk * v;
    }
    return acc;
}

// Forward Euler from a cell start: substep 1 reuses the cell's state-only
// terms (a grid state is finite, so the divergence test of substep 1 always
// passes), substeps 2..nsub are PeregrineSolver.jl:208-221 verbatim.
// One Euler substep (step_state's loop body, PeregrineSolver.jl:208-223)
// from a state whose finite-check prefix (x + y) + h and heading sincos are
// given.
template <bool DIAG>
__device__ __forceinline__ void euler_substep(const Consts& K, float bux, float buy, float buw, float hs, float pre,
                                              float sh, float ch, bool& ok, float& x, float& y, float& h, float& vx,
                                              float& vy, float& w) {
    float sum = pre + vx;
    sum = sum + vy;
    sum = sum + w;
    ok = ok && finite_f(sum);
    const float w2 = DIAG ? 0.0f : w * w;
    const float sx = clamp1(qdiv(vx, K.reps[0]));
    const float sy = clamp1(qdiv(vy, K.reps[1]));
    const float sw = clamp1(qdiv(w, K.reps[2]));
    const float dx = __builtin_fabsf(vx) * vx;
    const float dy = __builtin_fabsf(vy) * vy;
    const float dw = __builtin_fabsf(w) * w;
    float ax, ay, al;
    accel_rows<DIAG>(K, bux, buy, buw, vx, vy, w, w2, sx, sy, sw, dx, dy, dw, ax, ay, al);
    const float dvx = ax + w * vy;
    const float dvy = ay - w * vx;
    const float px = ch * vx - sh * vy;
    const float py = sh * vx + ch * vy;
    x = x + px * hs;
    y = y + py * hs;
    h = h + w * hs;
    vx = vx + dvx * hs;
    vy = vy + dvy * hs;
    w = w + al * hs;
}

// Forward Euler from a cell start: the first substep's position, heading and
// the second substep's sincos are the start's (they do not depend on the
// control), the rest as step_state does them.
template <bool DIAG>
__device__ __forceinline__ bool euler_from(const Consts& K, const CellStart& c, float bux, float buy,
                                           float buw, float& x, float& y, float& h, float& vx,
                                           float& vy, float& w) {
    const float hs = c.hs;
    bool ok = true;
    {
        float ax, ay, al;
        accel_rows<DIAG>(K, bux, buy, buw, c.vx, c.vy, c.w, c.w2, c.sx, c.sy, c.sw, c.dx, c.dy, c.dw,
                         ax, ay, al);
        const float dvx = ax + c.w * c.vy;
        const float dvy = ay - c.w * c.vx;
        x = c.x1;
        y = c.y1;
        h = c.h1;
        vx = c.vx + dvx * hs;
        vy = c.vy + dvy * hs;
        w = c.w + al * hs;
    }
    if (c.nsub > 1) euler_substep<DIAG>(K, bux, buy, buw, hs, c.s1, c.sh1, c.ch1, ok, x, y, h, vx, vy, w);
    for (int k = 2; k < c.nsub; ++k) {
        float pre = x + y;
        pre = pre + h;
        float sh, ch;
        sincos_ref(h, sh, ch);
        euler_substep<DIAG>(K, bux, buy, buw, hs, pre, sh, ch, ok, x, y, h, vx, vy, w);
    }
    return ok;
}

// Left-fold sum x + y + h + vx + vy + w of a state is finite (diverged,
// PeregrineSolver.jl:191-192, negated).
__device__ __forceinline__ bool state_finite(const float s[6]) {
    float sum = s[0] + s[1];
    sum = sum + s[2];
    sum = sum + s[3];
    sum = sum + s[4];
    sum = sum + s[5];
    return finite_f(sum);
}

// state_deriv (PeregrineSolver.jl:226-236) of state s under B*u_sat.
template <bool DIAG>
__device__ __forceinline__ void state_deriv(const Consts& K, float bux, float buy, float buw, const float s[6],
                                            float d[6]) {
    const float vx = s[3], vy = s[4], w = s[5];
    const float w2 = DIAG ? 0.0f : w * w;
    const float sx = clamp1(qdiv(vx, K.reps[0]));
    const float sy = clamp1(qdiv(vy, K.reps[1]));
    const float sw = clamp1(qdiv(w, K.reps[2]));
    const float dx = __builtin_fabsf(vx) * vx;
    const float dy = __builtin_fabsf(vy) * vy;
    const float dw = __builtin_fabsf(w) * w;
    float ax, ay, al;
    accel_rows<DIAG>(K, bux, buy, buw, vx, vy, w, w2, sx, sy, sw, dx, dy, dw, ax, ay, al);
    float sh, ch;
    sincos_ref(s[2], sh, ch);
    d[0] = ch * vx - sh * vy;
    d[1] = sh * vx + ch * vy;
    d[2] = w;
    d[3] = ax + w * vy;
    d[4] = ay - w * vx;
    d[5] = al;
}

// Explicit midpoint from a cell start (step_state_rk2, PeregrineSolver.jl:259-279):
// per substep d1 = f(s), mid = s + d1*hs/2, d2 = f(mid), s = s + d2*hs, with
// the divergence test on s and on mid. The first substep's f(s) is the
// cell's own terms (a grid state is finite).
template <bool DIAG>
__device__ __forceinline__ bool rk2_from(const Consts& K, const CellStart& c, float bux, float buy, float buw,
                                         float& x, float& y, float& h, float& vx, float& vy, float& w) {
    const float hs = c.hs;
    const float hh = 0.5f * hs;
    float s[6] = {c.x, c.y, c.h, c.vx, c.vy, c.w};
    bool ok = true;
    for (int k = 0; k < c.nsub; ++k) {
        float d[6];
        if (k == 0) {
            float ax, ay, al;
            accel_rows<DIAG>(K, bux, buy, buw, c.vx, c.vy, c.w, c.w2, c.sx, c.sy, c.sw, c.dx, c.dy, c.dw, ax, ay,
                             al);
            d[0] = c.ch * c.vx - c.sh * c.vy;
            d[1] = c.sh * c.vx + c.ch * c.vy;
            d[2] = c.w;
            d[3] = ax + c.w * c.vy;
            d[4] = ay - c.w * c.vx;
            d[5] = al;
        } else {
            ok = ok && state_finite(s);
            state_deriv<DIAG>(K, bux, buy, buw, s, d);
        }
        float mid[6];
#pragma unroll
        for (int j = 0; j < 6; ++j) mid[j] = s[j] + d[j] * hh;
        ok = ok && state_finite(mid);
        state_deriv<DIAG>(K, bux, buy, buw, mid, d);
#pragma unroll
        for (int j = 0; j < 6; ++j) s[j] = s[j] + d[j] * hs;
    }
    x = s[0]; y = s[1]; h = s[2]; vx = s[3]; vy = s[4]; w = s[5];
    return ok;
}

// The same from an arbitrary start state (the global-gather kernel).
__device__ __forceinline__ bool rk2_state(const Consts& K, float hs, int nsub, float bux, float buy, float buw,
                                          float& x, float& y, float& h, float& vx, float& vy, float& w) {
    const float hh = 0.5f * hs;
    float s[6] = {x, y, h, vx, vy, w};
    bool ok = true;
    for (int k = 0; k < nsub; ++k) {
        float d[6];
        ok = ok && state_finite(s);
        state_deriv<false>(K, bux, buy, buw, s, d);
        float mid[6];
#pragma unroll
        for (int j = 0; j < 6; ++j) mid[j] = s[j] + d[j] * hh;
        ok = ok && state_finite(mid);
        state_deriv<false>(K, bux, buy, buw, mid, d);
#pragma unroll
        for (int j = 0; j < 6; ++j) s[j] = s[j] + d[j] * hs;
    }
    x = s[0]; y = s[1]; h = s[2]; vx = s[3]; vy = s[4]; w = s[5];
    return ok;
}

// The grid constants of a lookup's coordinates, as Consts holds them. A
// kernel short of scalar registers keeps this copy in LDS and reads it once
// per lookahead instead of holding 27 more registers across its loop.
struct GridC {
    double rstep[6];
    double rn3;
    float lo[6];
    float lim[6];
    float n3f;
};

__device__ __forceinline__ void gridc_fill(const Consts& K, GridC& g) {
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        g.rstep[a] = K.rstep[a];
        g.lo[a] = K.lo[a];
        g.lim[a] = K.lim[a];
    }
    g.rn3 = K.rn3;
    g.n3f = K.n3f;
}

// Fractional coordinates of a successor, with the range tests of interp
// (PeregrineSolver.jl:401-428). Returns false when the lookup rejects. The
// grid constants come from gc: K itself, or a GridC copy of them.
template <bool GEN = true, class GC>
__device__ __forceinline__ bool interp_coords(const Consts& K, const GC& gc, float x, float y, float h, float vx,
                                              float vy, float w, float f[6]) {
    f[0] = qdiv(x - gc.lo[0], gc.rstep[0]);
    f[1] = qdiv(y - gc.lo[1], gc.rstep[1]);
    f[2] = wrap_h(gc, qdiv(h - gc.lo[2], gc.rstep[2]));
    f[3] = qdiv(vx - gc.lo[3], gc.rstep[3]);
    f[4] = qdiv(vy - gc.lo[4], gc.rstep[4]);
    f[5] = qdiv(w - gc.lo[5], gc.rstep[5]);
    bool in = true;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        if (k >= 3 && GEN && K.vclamp) {
            // velocity_clamp: a finite velocity off the grid reads the edge
            in = in & finite_f(f[k]);
            f[k] = f[k] > gc.lim[k] ? gc.lim[k] : (f[k] < 0.0f ? 0.0f : f[k]);
        } else {
            in = in & (f[k] >= 0.0f) & (f[k] <= gc.lim[k]);
        }
    }
    return in;
}

template <bool GEN = true>
__device__ __forceinline__ bool interp_coords(const Consts& K, float x, float y, float h, float vx,
                                              float vy, float w, float f[6]) {
    return interp_coords<GEN>(K, K, x, y, h, vx, vy, w, f);
}

// Level-by-level corner weights of the first five axes; weight d of the
// 64-corner sum is p5[d & 31] * (bit 5 of d ? t6 : u6), which is the
// reference's left-to-right six-factor product.
struct Weights {
    float p5[32];
    float t6, u6;
};

__device__ __forceinline__ void corner_weights(const float t[6], const float u[6], Weights& W) {
    float p2[4], p3[8], p4[16];
#pragma unroll
    for (int d = 0; d < 4; ++d) p2[d] = ((d & 1) ? t[0] : u[0]) * ((d & 2) ? t[1] : u[1]);
#pragma unroll
    for (int d = 0; d < 8; ++d) p3[d] = p2[d & 3] * ((d & 4) ? t[2] : u[2]);
#pragma unroll
    for (int d = 0; d < 16; ++d) p4[d] = p3[d & 7] * ((d & 8) ? t[3] : u[3]);
#pragma unroll
    for (int d = 0; d < 32; ++d) W.p5[d] = p4[d & 15] * ((d & 16) ? t[4] : u[4]);
    W.t6 = t[5];
    W.u6 = u[5];
}

// The reference's accumulation order d = 0..63 (bit 0 = x ... bit 5 = w).
__device__ __forceinline__ float corner_sum(const Weights& W, const float vals[64]) {
    float acc = 0.0f;
#pragma unroll
    for (int d = 0; d < 64; ++d) {
        const float wgt = W.p5[d & 31] * ((d & 32) ? W.t6 : W.u6);
        acc = acc + wgt * vals[d];
    }
    return acc;
}

// Global-memory corner reads with the reference's clamped / periodic
// indices (PeregrineSolver.jl:455-463).
__device__ __forceinline__ void fetch_global(const Consts& K, const float* __restrict__ V,
                                             const int i[6], float vals[64]) {
    int64_t o0[6], o1[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        int j0, j1;
        if (k == 2) {
            j0 = (i[k] >= K.n[2]) ? i[k] - K.n[2] : i[k];
            j1 = i[k] + 1;
            j1 = (j1 >= K.n[2]) ? j1 - K.n[2] : j1;
        } else {
            j0 = i[k];
            j1 = (i[k] + 1 < K.n[k] - 1) ? i[k] + 1 : K.n[k] - 1;
        }
        o0[k] = (int64_t)j0 * K.stride[k];
        o1[k] = (int64_t)j1 * K.stride[k];
    }
#pragma unroll
    for (int d = 0; d < 64; ++d) {
        int64_t a = 0;
#pragma unroll
        for (int k = 0; k < 6; ++k) a += ((d >> k) & 1) ? o1[k] : o0[k];
        vals[d] = V[a];
    }
}

// Value stores of the precision-form kernels.
#define PG_STORE_F32 0
#define PG_STORE_F16 1   // IEEE binary16 codes of (v - off) / scale
#define PG_STORE_U16 2   // 16-bit fixed-point codes of (v - off) / scale
// Quantization tiles of the 16-bit stores: 2^kQuantShift cells per axis (a
// multiple of every box edge, so each box lies in one tile).
constexpr int kQuantShift = 2;
struct QuantGeom {
    int32_t nt[6];    // tiles per axis, ceil(n / 2^kQuantShift)
    int64_t ntile;
};

// Box decomposition of the grid for the LDS-staged kernels.
// Boxes of a region of the grid (the whole grid, or one tile of an
// out-of-core sweep). The plan's per-box arrays are numbered over the whole
// grid's boxes, so a region carries that numbering too.
struct BoxGeom {
    int32_t b[6];     // box dims
    int32_t nb[6];    // boxes per axis over the region
    int64_t nbox;     // boxes of the region
    int32_t org[6];   // first cell of the region per axis (a multiple of b)
    int32_t rn[6];    // cells of the region per axis
    int32_t nbg[6];   // boxes per axis over the whole grid
    int64_t blk0;     // first workgroup of this launch (launches are chunked)
    int32_t rev;      // walk the boxes in reverse order (in-place sweeps alternate)
};

// Workgroups of a launch over nbox boxes (one box each): whole rounds of 8
// chunks of kBoxChunk boxes, the tail's workgroups idle.
constexpr int64_t kBoxChunk = 32;
__host__ __device__ inline int64_t box_blocks(int64_t nbox) {
    const int64_t chunks = (nbox + kBoxChunk - 1) / kBoxChunk;
    return ((chunks + 7) / 8) * 8 * kBoxChunk;
}

// Active boxes of a Jacobi sweep. One bit per box over the whole grid's
// boxes, numbered as BoxGeom's (omega fastest): the nbg[5] <= 32 boxes of a
// row along omega share a word (word gbox / nbg[5], bit gbox % nbg[5]).
// A box whose read range (every cell any of its backups can read, from a
// plan pass with the sweep's own arithmetic) holds no box that changed in
// the previous sweep, and which did not change itself, is skipped: its
// backups would see the inputs of the last sweep that computed them, so
// V' = min(V, TV) = V, and both Jacobi buffers already hold V.
struct ActiveG {
    const uint32_t* prev;   // boxes changed by the previous sweep; null: every box is computed
    uint32_t* next;         // this sweep marks the boxes it changes (FP32 stores; null: not marked)
    const int16_t* rorig;   // per box: first cell of the read range per axis (heading modulo n3)
    const uint8_t* rext;    //          cells of the read range per axis
    unsigned* nwork;        // counts the boxes computed (may be null)
};

// Block-wide sum (FP64) and max of the per-cell decreases: wave shuffles,
// then one LDS slot per wave; thread 0 writes the block's partial sum and
// folds the max into *linf_bits (non-negative floats order as their bits).
// Returns the block's max to thread 0 (0 to the others).
__device__ __forceinline__ float block_reduce_sum_max(double dsum, float dmax, double* part_sum,
                                                      unsigned* linf_bits) {
    __shared__ double ssum[16];
    __shared__ float smax[16];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        dsum += __shfl_xor(dsum, off, 64);
        dmax = fmaxf(dmax, __shfl_xor(dmax, off, 64));
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) { ssum[wv] = dsum; smax[wv] = dmax; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = ssum[0];
        float m = smax[0];
        const int nw = (int)(blockDim.x >> 6);
        for (int i = 1; i < nw; ++i) { s += ssum[i]; m = fmaxf(m, smax[i]); }
        part_sum[blockIdx.x] = s;
        if (m > 0.0f) atomicMax(linf_bits, __float_as_uint(m));
        return m;
    }
    return 0.0f;
}

}  // namespace pgk
