// Bellman sweep kernels for gfx950 (plain scheme: fixed horizon, full
// control scan, forward Euler or explicit midpoint, fixed swept probes,
// 64-corner multilinear), and the escape pass over the same backup.
//
// One thread per cell; the control table is scanned in order with a strict
// `<`, coast first, so ties resolve to the lowest index exactly as
// cell_update does (solver/src/PeregrineSolver.jl:1245-1275). A sweep is
// Jacobi: every backup reads Vin and writes min(Vin, TVin) to Vout, which
// makes the result independent of scheduling and equal bit for bit to the
// CPU restatement.
#include "pg_box.hpp"
#include "pg_tanh.hpp"
#include "pg_kernels.h"

namespace pgk {

// 64-corner multilinear interpolation of V at a continuous state
// (PeregrineSolver.jl:399-466). Rejections return cap. The corner weight is
// the left-to-right product over the six axes and the corners are summed in
// the reference's order d = 0..63 (bit 0 = x ... bit 5 = w); a zero weight
// adds +0 to a non-negative running sum, so the reference's skip of those
// corners needs no branch here.
//
// ESC reads V as the escape pass does (EscapeView, PeregrineSolver.jl:
// 836-842): a corner whose (x, y, h) cell is an obstacle reads cap unless
// `ignore` (the updating cell is itself blocked); otherwise a negative value
// reads as its escape time, a real route (0 <= v < cap) as 0, the rest as cap.
// Where a backup reads V: the whole table (VGlobal), or an out-of-core
// tile's window (VWindow). off() is a corner index's offset along axis k,
// load() the value at a summed offset.
struct VGlobal {
    const float* V;
    __device__ __forceinline__ int64_t off(const Consts& K, int k, int j) const { return (int64_t)j * K.stride[k]; }
    __device__ __forceinline__ float load(const Consts&, int64_t a) const { return V[a]; }
    // one corner of a 64-corner gather (load_acc), and the gather's end
    // (report): the whole table holds every offset
    __device__ __forceinline__ float load_acc(const Consts&, int64_t a, bool&) const { return V[a]; }
    __device__ __forceinline__ void report(bool) const {}
    // the value of cell i (flat index idx on the whole grid)
    __device__ __forceinline__ float self(const Consts&, int64_t idx, const int32_t*) const { return V[idx]; }
};

// A window: cells x in [o0, o0 + L0), y in [o1, o1 + L1), heading in
// [o2, o2 + L2) modulo n3, every velocity cell; omega fastest. An index
// outside it has an offset so negative that every sum holding it stays
// negative; load() reads only offsets inside the window, and any other
// returns cap and raises *miss (the window was built to hold every read).
constexpr int64_t kOutsideWindow = -((int64_t)1 << 60);

struct VWindow {
    const float* __restrict__ W;
    int32_t o[3], L[3];
    int64_t ws[6];
    int64_t size;
    unsigned* miss;
    __device__ __forceinline__ int64_t off(const Consts& K, int k, int j) const {
        if (k < 3) {
            int l = j - o[k];
            if (k == 2 && l < 0) l += K.n[2];
            if ((unsigned)l >= (unsigned)L[k]) return kOutsideWindow;
            return (int64_t)l * ws[k];
        }
        return (int64_t)j * ws[k];
    }
    __device__ __forceinline__ float load(const Consts& K, int64_t a) const {
        if ((uint64_t)a >= (uint64_t)size) {
            atomicOr(miss, 1u);
            return K.cap;
        }
        return W[a];
    }
    // One corner of a 64-corner gather without the defect check's branch:
    // cap outside the window, noted in bad; report raises *miss once for the
    // gather (a branch and an atomic per corner run the kernel out of scalar
    // registers).
    __device__ __forceinline__ float load_acc(const Consts& K, int64_t a, bool& bad) const {
        const bool out = (uint64_t)a >= (uint64_t)size;
        bad |= out;
        const float v = W[out ? 0 : a];
        return out ? K.cap : v;
    }
    __device__ __forceinline__ void report(bool bad) const {
        if (bad) atomicOr(miss, 1u);
    }
    __device__ __forceinline__ float self(const Consts& K, int64_t, const int32_t* i) const {
        int64_t a = 0;
#pragma unroll
        for (int k = 0; k < 6; ++k) a += off(K, k, i[k]);
        return load(K, a);
    }
};

// The cells a launch backs up: x, y, heading from t0 over tn, every velocity
// cell; omega fastest; ncell of them (the whole grid, or an out-of-core
// tile's interior).
struct TileCells {
    int32_t t0[3], tn[3];
    int64_t ncell;
};

template <bool ESC, class VR>
__device__ __forceinline__ float interp_global(const Consts& K, const VR& V,
                                               const uint8_t* __restrict__ occ, bool ignore,
                                               float x, float y, float h, float vx, float vy,
                                               float w) {
    float f[6];
    f[0] = qdiv(x - K.lo[0], K.rstep[0]);
    f[1] = qdiv(y - K.lo[1], K.rstep[1]);
    f[2] = wrap_h(K, qdiv(h - K.lo[2], K.rstep[2]));
    f[3] = qdiv(vx - K.lo[3], K.rstep[3]);
    f[4] = qdiv(vy - K.lo[4], K.rstep[4]);
    f[5] = qdiv(w - K.lo[5], K.rstep[5]);
    bool in = true;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        if (k >= 3 && K.vclamp) {
            in = in & finite_f(f[k]);
            f[k] = f[k] > K.lim[k] ? K.lim[k] : (f[k] < 0.0f ? 0.0f : f[k]);
        } else {
            in = in & (f[k] >= 0.0f) & (f[k] <= K.lim[k]);
        }
    }
    if (!in) return K.cap;
    float t[6], u[6];
    int64_t o0[6], o1[6];
    int jc[3][2];   // corner cells of x, y, h (the escape reads' occupancy)
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const float b = __builtin_floorf(f[k]);
        t[k] = f[k] - b;
        u[k] = 1.0f - t[k];
        const int i = (int)b;
        int j0, j1;
        if (k == 2) {
            j0 = (i >= K.n[2]) ? i - K.n[2] : i;
            j1 = i + 1;
            j1 = (j1 >= K.n[2]) ? j1 - K.n[2] : j1;
        } else {
            j0 = i;
            j1 = (i + 1 < K.n[k] - 1) ? i + 1 : K.n[k] - 1;
        }
        o0[k] = V.off(K, k, j0);
        o1[k] = V.off(K, k, j1);
        if (k < 3) { jc[k][0] = j0; jc[k][1] = j1; }
    }
    if (K.nearest) {
        // the node nearest the successor (PeregrineSolver.jl:424-428): per
        // axis the lower or upper corner, as rint(f) falls
        int64_t a = 0;
        int c = 0;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const bool up = __builtin_rintf(f[k]) != __builtin_floorf(f[k]);
            a += up ? o1[k] : o0[k];
            if (k < 3) c |= (up ? 1 : 0) << k;
        }
        float v = V.load(K, a);
        if (ESC) {
            const bool wall = !ignore && occ[((int64_t)jc[0][c & 1] * K.n[1] + jc[1][(c >> 1) & 1]) * K.n[2] +
                                             jc[2][(c >> 2) & 1]] != 0;
            v = wall ? K.cap : (v < 0.0f ? -v : (v < K.cap ? 0.0f : K.cap));
        }
        return v;
    }
    if (K.simplex) {
        const float acc = kuhn_interp(f, [&](int mask) {
            int64_t a = 0;
            int c = 0;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const bool up = (mask >> k) & 1;
                a += up ? o1[k] : o0[k];
                if (k < 3) c |= (up ? 1 : 0) << k;
            }
            float v = V.load(K, a);
            if (ESC) {
                const bool wall = !ignore && occ[((int64_t)jc[0][c & 1] * K.n[1] + jc[1][(c >> 1) & 1]) * K.n[2] +
                                                 jc[2][(c >> 2) & 1]] != 0;
                v = wall ? K.cap : (v < 0.0f ? -v : (v < K.cap ? 0.0f : K.cap));
            }
            return v;
        });
        return (acc < K.cap) ? acc : K.cap;
    }
    // weights: p[d] for the first five axes, built level by level so that
    // every weight is ((((w1*w2)*w3)*w4)*w5)*w6 in the reference's order
    float p2[4], p3[8], p4[16], p5[32];
    int64_t a2[4], a3[8], a4[16], a5[32];
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        p2[d] = ((d & 1) ? t[0] : u[0]) * ((d & 2) ? t[1] : u[1]);
        a2[d] = ((d & 1) ? o1[0] : o0[0]) + ((d & 2) ? o1[1] : o0[1]);
    }
#pragma unroll
    for (int d = 0; d < 8; ++d) {
        p3[d] = p2[d & 3] * ((d & 4) ? t[2] : u[2]);
        a3[d] = a2[d & 3] + ((d & 4) ? o1[2] : o0[2]);
    }
#pragma unroll
    for (int d = 0; d < 16; ++d) {
        p4[d] = p3[d & 7] * ((d & 8) ? t[3] : u[3]);
        a4[d] = a3[d & 7] + ((d & 8) ? o1[3] : o0[3]);
    }
#pragma unroll
    for (int d = 0; d < 32; ++d) {
        p5[d] = p4[d & 15] * ((d & 16) ? t[4] : u[4]);
        a5[d] = a4[d & 15] + ((d & 16) ? o1[4] : o0[4]);
    }
    float vals[64];
    bool bad = false;
#pragma unroll
    for (int d = 0; d < 64; ++d) vals[d] = V.load_acc(K, a5[d & 31] + ((d & 32) ? o1[5] : o0[5]), bad);
    V.report(bad);
    if (ESC) {
        bool wall[8];
#pragma unroll
        for (int c = 0; c < 8; ++c)
            wall[c] = !ignore && occ[((int64_t)jc[0][c & 1] * K.n[1] + jc[1][(c >> 1) & 1]) * K.n[2] +
                                     jc[2][(c >> 2) & 1]] != 0;
#pragma unroll
        for (int d = 0; d < 64; ++d) {
            const float v = vals[d];
            vals[d] = wall[d & 7] ? K.cap : (v < 0.0f ? -v : (v < K.cap ? 0.0f : K.cap));
        }
    }
    float acc = 0.0f;
#pragma unroll
    for (int d = 0; d < 64; ++d) {
        const float wgt = p5[d & 31] * ((d & 32) ? t[5] : u[5]);
        acc = acc + wgt * vals[d];
    }
    return (acc < K.cap) ? acc : K.cap;
}

// Occupancy as the probes read it; `ignore` waives it (the escape pass's
// OccView, PeregrineSolver.jl:864-874: a blocked cell may move through
// obstacle space).
struct OccGlobal {
    const uint8_t* __restrict__ occ;
    int32_t n2, n3;
    bool ignore;
    __device__ __forceinline__ bool operator()(int ci, int cj, int ck) const {
        return !ignore && occ[((int64_t)ci * n2 + cj) * n3 + ck] != 0;
    }
};

// lookahead (PeregrineSolver.jl:1100-1153) for one control.
template <bool ESC, class VR>
__device__ __forceinline__ float lookahead_global(const Consts& K, const VR& V,
                                                  const OccGlobal& occ, float tau, float hs, int nsub,
                                                  float x, float y,
                                                  float h, float vx, float vy, float w,
                                                  float bux, float buy, float buw) {
    float nx = x, ny = y, nh = h, nvx = vx, nvy = vy, nw = w;
    const bool ok = K.rk2 ? rk2_state(K, hs, nsub, bux, buy, buw, nx, ny, nh, nvx, nvy, nw)
                          : euler(K, hs, nsub, bux, buy, buw, nx, ny, nh, nvx, nvy, nw);
    const float dx = nx - x, dy = ny - y, dh = nh - h;
    if (!(ok & finite_f(dx) & finite_f(dy) & finite_f(dh) & finite_f(nvx) & finite_f(nvy) &
          finite_f(nw)))
        return K.cap;
    if (!probes_free(K, occ, x, y, h, dx, dy, dh)) return K.cap;
    float snh, cnh;
    sincos_ref(nh, snh, cnh);
    const float nvfx = cnh * nvx - snh * nvy;
    const float nvfy = snh * nvx + cnh * nvy;
    const float nv = interp_global<ESC>(K, V, occ.occ, occ.ignore, nx, ny, nh, nvfx, nvfy, nw);
    const float c = tau + nv;
    return (c < K.cap) ? c : K.cap;
}

// B * u_sat of a raw command (body_accel_sat's first three terms per row,
// PeregrineSolver.jl:139-162): the host precomputes these for the lattice,
// the device for the warm-start and pattern-search commands, with the same
// arithmetic.
__device__ __forceinline__ void bu_of(const Consts& K, float u1r, float u2r, float u3r, float& bx, float& by,
                                      float& bw) {
    const float gn = pgt::traction_gain_f(u1r, u2r, u3r, K.knee);
    const float u1 = u1r * gn, u2 = u2r * gn, u3 = u3r * gn;
    float a[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        float t = K.B[3 * r] * u1;
        t = t + K.B[3 * r + 1] * u2;
        t = t + K.B[3 * r + 2] * u3;
        a[r] = t;
    }
    bx = a[0]; by = a[1]; bw = a[2];
}

// proj_oct (PeregrineSolver.jl:1165-1170): onto |u1| + |u2| + |u3| = 1.
__device__ __forceinline__ void proj_oct(float& u1, float& u2, float& u3) {
    const float s = __builtin_fabsf(u1) + __builtin_fabsf(u2) + __builtin_fabsf(u3);
    if (s < 1.0e-6f) { u1 = u2 = u3 = 0.0f; return; }
    const float r = 1.0f / s;
    u1 = u1 * r; u2 = u2 * r; u3 = u3 * r;
}

// Warm-start store entry (pol_set!, PeregrineSolver.jl:736-744).
__device__ __forceinline__ int8_t pol_quant(float v) {
    const float c = v < -1.0f ? -1.0f : (v > 1.0f ? 1.0f : v);
    return (int8_t)(int32_t)__builtin_rintf(c * 127.0f);
}

// One thread per cell of T (the whole grid, or an out-of-core tile's interior:
// then Vout, pol and wpol are interior-shaped and V is read through the
// tile's window). ESC: the escape pass (gpu_escape_kernel!, Gpu.jl:51-67)
// in Jacobi form: a cell with a real route (0 <= v < cap) is the boundary and
// keeps its value; any other cell, blocked ones included (through obstacle
// space, PeregrineSolver.jl:864-874), takes -c when its backup c beats the
// escape time it holds (-v if v < 0, else cap).
//
// The candidates, in the order cell_update takes them (PeregrineSolver.jl:
// 1240-1343), the earlier winning ties: coast; with PROD, the cell's
// warm-start command (wpol, Int8 x 1/127, when non-zero); the lattice (whole
// or this phase's slice); with PROD, K.rounds rounds of pattern search around
// the best command (six axis moves of +-delta projected onto the octahedron,
// delta halving per round), the horizon ladder (K.ntau > 1) and the dt rung
// (cfl > 0) with the best command. One loop walks the candidate slots so
// that the lookahead is inlined once. When the value improves the winning
// command is written back to the cell's own wpol entry (only the cell's own
// entry is read, so the in-place update is the Jacobi one).
// Three waves per SIMD (at most 168 VGPRs): the production value sweep then
// spills 6 VGPRs (28 B/lane of scratch) and still sweeps 20% faster than at
// two waves (188 VGPRs, no spill).
template <bool WRITE_V, bool WRITE_POL, bool ESC, bool PROD, class VR>
__global__ __launch_bounds__(256, 3) void sweep_global_kernel(
    VR Vin, float* Vout, TileCells T, const uint8_t* __restrict__ occ,
    const float* __restrict__ bu, const float* __restrict__ uraw, Consts K, uint8_t* __restrict__ pol,
    int8_t* __restrict__ wpol, double* __restrict__ part_sum, unsigned* __restrict__ linf_bits) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float dec = 0.0f;
    if (tid < T.ncell) {
        // With a per-cell horizon the substep counts follow the velocity
        // (cfl_tau), so the threads of a wave take one velocity at 64
        // consecutive positions (x, y, heading): every lane runs the same
        // number of substeps. The positions go in blocks of 64 whose every
        // velocity is taken by consecutive waves, so that the block's reads
        // of V, which share cache lines across velocities, are served while
        // those lines are still cached. Otherwise consecutive threads take
        // consecutive cells.
        int64_t idx = tid;
        if (K.cfl) {
            constexpr int64_t P = 64;
            const int64_t nvel = (int64_t)K.n[3] * K.n[4] * K.n[5];
            const int64_t npos = T.ncell / nvel;
            const int64_t pb = tid / (P * nvel);               // the last block may hold fewer positions
            const int64_t pp = npos - pb * P < P ? npos - pb * P : P;
            const int64_t rem = tid - pb * P * nvel;
            const int64_t vel = rem / pp;
            idx = (pb * P + (rem - vel * pp)) * nvel + vel;
        }
        int64_t r = idx;
        int32_t i[6];
#pragma unroll
        for (int k = 5; k >= 3; --k) {
            i[k] = (int32_t)(r % K.n[k]);
            r /= K.n[k];
        }
        i[2] = T.t0[2] + (int32_t)(r % T.tn[2]);
        r /= T.tn[2];
        i[1] = T.t0[1] + (int32_t)(r % T.tn[1]);
        i[0] = T.t0[0] + (int32_t)(r / T.tn[1]);
        const bool blocked = occ[((int64_t)i[0] * K.n[1] + i[1]) * K.n[2] + i[2]] != 0;
        const float old = Vin.self(K, idx, i);
        const bool terminal = old >= 0.0f && old < K.cap;
        float best = K.cap;
        int bk = 0;
        float b1 = 0.0f, b2 = 0.0f, b3 = 0.0f;
        if (ESC ? !terminal : !blocked) {
            const float x = K.lo[0] + (float)i[0] * K.step[0];
            const float y = K.lo[1] + (float)i[1] * K.step[1];
            const float h = K.lo[2] + (float)i[2] * K.step[2];
            const float vfx = K.lo[3] + (float)i[3] * K.step[3];
            const float vfy = K.lo[4] + (float)i[4] * K.step[4];
            const float w = K.lo[5] + (float)i[5] * K.step[5];
            float sh0, ch0;
            sincos_ref(h, sh0, ch0);
            const float vx = ch0 * vfx + sh0 * vfy;
            const float vy = -sh0 * vfx + ch0 * vfy;
            const OccGlobal og{occ, K.n[1], K.n[2], ESC && blocked};
            // the cell's horizon: dt, or from its speed and spin (cfl > 0)
            float tau = K.tau, hs = K.hs;
            int nsub = K.nsub;
            if (K.cfl) {
                tau = cfl_tau_dev(K, vfx, vfy, w);
                nsub = substeps_dev(K, tau);
                hs = tau / (float)nsub;
            }
            // slots: 0 coast, 1 warm start, 2 .. 1+span lattice, then 6 per
            // search round, then ntau ladder rungs and the dt rung
            const int s_lat = 2, s_pat = 2 + K.span;
            const int s_lad = PROD ? s_pat + 6 * K.rounds : s_pat;
            const int s_end = PROD ? (K.ntau > 1 ? s_lad + K.ntau + 1 : s_lad) : s_pat;
            float tl = tau * 0.5f;        // the ladder's running horizon
            float lx = 0.0f, ly = 0.0f, lw = 0.0f;   // B * u_sat of the ladder's command
            for (int slot = 0; slot < s_end; ++slot) {
                float c1 = 0.0f, c2 = 0.0f, c3 = 0.0f, bx, by, bw, t = tau, ths = hs;
                int tn = nsub, kk = -1;
                bool ladder = false;
                if (slot == 0) {
                    bx = bu[0]; by = bu[1]; bw = bu[2];
                    kk = 0;
                } else if (slot < s_lat) {
                    if (!PROD || ESC || !wpol) continue;
                    c1 = (float)wpol[3 * idx] * (1.0f / 127.0f);
                    c2 = (float)wpol[3 * idx + 1] * (1.0f / 127.0f);
                    c3 = (float)wpol[3 * idx + 2] * (1.0f / 127.0f);
                    if (!(__builtin_fabsf(c1) + __builtin_fabsf(c2) + __builtin_fabsf(c3) > 1.0e-6f)) continue;
                    bu_of(K, c1, c2, c3, bx, by, bw);
                } else if (slot < s_pat) {
                    kk = slice_control(K, slot - 1);
                    bx = bu[3 * kk]; by = bu[3 * kk + 1]; bw = bu[3 * kk + 2];
                    if (PROD) { c1 = uraw[3 * kk]; c2 = uraw[3 * kk + 1]; c3 = uraw[3 * kk + 2]; }
                } else if (slot < s_lad) {
                    const int j = slot - s_pat, d = j % 6 + 1;
                    // delta0 halved once per finished round (exact: a power of two)
                    const float delta = __builtin_ldexpf(K.delta0, -(j / 6));
                    const float sg = (d % 2 == 0) ? -delta : delta;
                    c1 = b1 + (d <= 2 ? sg : 0.0f);
                    c2 = b2 + ((d == 3 || d == 4) ? sg : 0.0f);
                    c3 = b3 + (d >= 5 ? sg : 0.0f);
                    proj_oct(c1, c2, c3);
                    if (__builtin_fabsf(c1) + __builtin_fabsf(c2) + __builtin_fabsf(c3) < 1.0e-6f) continue;
                    bu_of(K, c1, c2, c3, bx, by, bw);
                } else {
                    const int k = slot - s_lad + 1;   // rung k, or ntau + 1 for the dt rung
                    if (k == 1) bu_of(K, b1, b2, b3, lx, ly, lw);
                    if (k <= K.ntau) {
                        t = tl;
                        tl = tl * K.tau_ratio;
                        if (k == 2) continue;
                    } else {
                        if (!K.cfl) continue;
                        t = K.tau;
                    }
                    t = t > K.tau_min ? t : K.tau_min;
                    t = t < K.tau_max ? t : K.tau_max;
                    tn = substeps_dev(K, t);
                    ths = t / (float)tn;
                    bx = lx; by = ly; bw = lw;
                    ladder = true;
                }
                const float c = lookahead_global<ESC>(K, Vin, og, t, ths, tn, x, y, h, vx, vy, w, bx, by, bw);
                if (c < best) {
                    best = c;
                    if (!ladder) { bk = kk; b1 = c1; b2 = c2; b3 = c3; }
                }
            }
        }
        if (WRITE_V) {
            float out = old;
            if (ESC) {
                const float e_old = old < 0.0f ? -old : K.cap;
                if (!terminal && best < e_old) { out = -best; dec = e_old - best; }
            } else if (best < old) {
                out = best;
                dec = old - best;
                if (PROD && wpol) {
                    wpol[3 * idx] = pol_quant(b1);
                    wpol[3 * idx + 1] = pol_quant(b2);
                    wpol[3 * idx + 2] = pol_quant(b3);
                }
            }
            Vout[idx] = out;
        }
        if (WRITE_POL) pol[idx] = (uint8_t)bk;
    }
    if (WRITE_V) block_reduce_sum_max((double)dec, dec, part_sum, linf_bits);
}

// Reach of the production scheme, for out-of-core windows: the successor's
// displacement from every (heading, velocity, spin) grid state -- position
// does not enter the dynamics -- under every lattice command and `ns` raw
// commands sampled over the octahedron (the pattern search and the warm
// start evaluate commands on it), at the cell's horizon, every ladder rung
// and the dt rung. Probes are not applied (a superset). Out: the smallest and
// largest displacement in cells, floor/ceil, for x, y and heading (ints
// lo0, hi0, lo1, hi1, lo2, hi2; pre-set by the caller).
__global__ __launch_bounds__(256) void reach_kernel(Consts K, const float* __restrict__ bu, const float* __restrict__ raw,
                                                    int ns, int* __restrict__ out) {
    const int64_t nst = (int64_t)K.n[2] * K.n[3] * K.n[4] * K.n[5];
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float mn[3] = {0.0f, 0.0f, 0.0f}, mx[3] = {0.0f, 0.0f, 0.0f};
    if (idx < nst) {
        int64_t r = idx;
        int i[6];
#pragma unroll
        for (int k = 5; k >= 3; --k) {
            i[k] = (int)(r % K.n[k]);
            r /= K.n[k];
        }
        i[2] = (int)r;
        const float h = K.lo[2] + (float)i[2] * K.step[2];
        const float vfx = K.lo[3] + (float)i[3] * K.step[3];
        const float vfy = K.lo[4] + (float)i[4] * K.step[4];
        const float w = K.lo[5] + (float)i[5] * K.step[5];
        float sh0, ch0;
        sincos_ref(h, sh0, ch0);
        const float vx = ch0 * vfx + sh0 * vfy;
        const float vy = -sh0 * vfx + ch0 * vfy;
        float tau = K.tau;
        if (K.cfl) tau = cfl_tau_dev(K, vfx, vfy, w);
        const int nrung = K.ntau > 1 ? K.ntau : 0;
        for (int c = 0; c < K.nctl + ns; ++c) {
            float bx, by, bw;
            if (c < K.nctl) {
                bx = bu[3 * c]; by = bu[3 * c + 1]; bw = bu[3 * c + 2];
            } else {
                const float* u = raw + 3 * (c - K.nctl);
                bu_of(K, u[0], u[1], u[2], bx, by, bw);
            }
            float tl = tau * 0.5f;
            for (int q = -2; q < nrung; ++q) {
                // q = -2: the cell's horizon; -1: the dt rung; 0..: ladder rungs
                float t;
                if (q == -2) {
                    t = tau;
                } else if (q == -1) {
                    if (!K.cfl) continue;
                    t = K.tau;
                } else {
                    t = tl;
                    tl = tl * K.tau_ratio;
                }
                if (q >= -1) {
                    t = t > K.tau_min ? t : K.tau_min;
                    t = t < K.tau_max ? t : K.tau_max;
                }
                // the horizons and substeps the sweep takes for them
                int tn = K.nsub;
                float ths = K.hs;
                if (q >= -1 || K.cfl) {
                    tn = substeps_dev(K, t);
                    ths = t / (float)tn;
                }
                float nx = 0.0f, ny = 0.0f, nh = h, nvx = vx, nvy = vy, nw = w;
                const bool ok = K.rk2 ? rk2_state(K, ths, tn, bx, by, bw, nx, ny, nh, nvx, nvy, nw)
                                      : euler(K, ths, tn, bx, by, bw, nx, ny, nh, nvx, nvy, nw);
                const float d[3] = {nx, ny, nh - h};
                if (!(ok & finite_f(d[0]) & finite_f(d[1]) & finite_f(d[2]) & finite_f(nvx) & finite_f(nvy) &
                      finite_f(nw)))
                    continue;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    mn[a] = d[a] < mn[a] ? d[a] : mn[a];
                    mx[a] = d[a] > mx[a] ? d[a] : mx[a];
                }
            }
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        // in cells, far values held at +-1e6 (they reach the whole axis)
        const float lo = __builtin_floorf(mn[a] / K.step[a]), hi = __builtin_ceilf(mx[a] / K.step[a]);
        const int il = lo < -1.0e6f ? -1000000 : (int)lo;
        const int ih = hi > 1.0e6f ? 1000000 : (int)hi;
        atomicMin(&out[2 * a], il);
        atomicMax(&out[2 * a + 1], ih);
    }
}

// Deterministic sum of the per-block partials (a fixed order for each n):
// block b sums the b-th of gridDim.x contiguous slices into out[b].
__global__ __launch_bounds__(256) void reduce_partials_kernel(const double* __restrict__ part,
                                                              int64_t n, double* __restrict__ out) {
    __shared__ double s[256];
    const int64_t per = (n + gridDim.x - 1) / gridDim.x;
    const int64_t i0 = (int64_t)blockIdx.x * per;
    const int64_t i1 = i0 + per < n ? i0 + per : n;
    double acc = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) acc += part[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off >= 1; off >>= 1) {
        if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

__global__ void fill_kernel(float* __restrict__ V, int64_t n, float v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) V[i] = v;
}

__global__ void seed_kernel(float* __restrict__ V, const int64_t* __restrict__ seeds, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) V[seeds[i]] = 0.0f;
}

// The seed reset of an FP32 table under an active set: a seed whose value
// it changes marks its box changed for the next sweep.
__global__ void seed_mark_kernel(float* __restrict__ V, const int64_t* __restrict__ seeds, int64_t n, Consts K,
                                 BoxGeom G, uint32_t* __restrict__ mark) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int64_t c = seeds[j];
    const float v = V[c];
    V[c] = 0.0f;
    if (__float_as_uint(v) != 0u) {
        int64_t r = c;
        int i[6];
        for (int a = 5; a >= 0; --a) {
            i[a] = (int)(r % K.n[a]);
            r /= K.n[a];
        }
        box_mark(mark, G, box_of_cell(G, i));
    }
}

}  // namespace pgk

// ---------------------------------------------------------------------------
// launch wrappers (host side)
// ---------------------------------------------------------------------------

// Whether a sweep needs the production candidates (warm start, pattern
// search, horizon ladder).
static bool prod_scheme(const pgk::Consts& K, const int8_t* wpol) {
    return wpol != nullptr || K.rounds > 0 || K.ntau > 1;
}

static pgk::TileCells whole_grid(const pgk::Consts& K) {
    pgk::TileCells T;
    for (int a = 0; a < 3; ++a) {
        T.t0[a] = 0;
        T.tn[a] = K.n[a];
    }
    T.ncell = K.ncell;
    return T;
}

template <bool WV, bool WP, bool ESC>
static void launch_global(const pgk::Consts& K, int64_t nb, hipStream_t st, const float* Vin, float* Vout,
                          const uint8_t* occ, const float* bu, const float* uraw, uint8_t* pol, int8_t* wpol,
                          double* part, unsigned* linf_bits) {
    const pgk::VGlobal vin{Vin};
    const pgk::TileCells T = whole_grid(K);
    if (prod_scheme(K, wpol))
        hipLaunchKernelGGL((pgk::sweep_global_kernel<WV, WP, ESC, true, pgk::VGlobal>), dim3((unsigned)nb), dim3(256),
                           0, st, vin, Vout, T, occ, bu, uraw, K, pol, wpol, part, linf_bits);
    else
        hipLaunchKernelGGL((pgk::sweep_global_kernel<WV, WP, ESC, false, pgk::VGlobal>), dim3((unsigned)nb), dim3(256),
                           0, st, vin, Vout, T, occ, bu, uraw, K, pol, wpol, part, linf_bits);
}

hipError_t pgk_launch_sweep(const pgk::Consts& K, const float* Vin, float* Vout,
                            const uint8_t* occ, const float* bu, const float* uraw, uint8_t* pol, int8_t* wpol,
                            double* part, unsigned* linf_bits, int write_v, hipStream_t st, int64_t* nblocks) {
    const int64_t nb = (K.ncell + 255) / 256;
    *nblocks = nb;
    if (write_v && pol)
        launch_global<true, true, false>(K, nb, st, Vin, Vout, occ, bu, uraw, pol, wpol, part, linf_bits);
    else if (write_v)
        launch_global<true, false, false>(K, nb, st, Vin, Vout, occ, bu, uraw, pol, wpol, part, linf_bits);
    else
        launch_global<false, true, false>(K, nb, st, Vin, Vout, occ, bu, uraw, pol, wpol, part, linf_bits);
    return hipGetLastError();
}

hipError_t pgk_launch_escape(const pgk::Consts& K, const float* Vin, float* Vout, const uint8_t* occ,
                             const float* bu, const float* uraw, uint8_t* pol, double* part, unsigned* linf_bits,
                             hipStream_t st, int64_t* nblocks) {
    const int64_t nb = (K.ncell + 255) / 256;
    *nblocks = nb;
    if (pol)
        launch_global<true, true, true>(K, nb, st, Vin, Vout, occ, bu, uraw, pol, nullptr, part, linf_bits);
    else
        launch_global<true, false, true>(K, nb, st, Vin, Vout, occ, bu, uraw, pol, nullptr, part, linf_bits);
    return hipGetLastError();
}

// A tile's window reader and cells; false if either needs 2^31 or more cells
// (a dispatch counts its work-items in 32 bits).
static bool tile_reader(const pgk::Consts& K, const int32_t t0[3], const int32_t tn[3], const float* win,
                        const int32_t wo[3], const int32_t wl[3], unsigned* miss, pgk::VWindow& Vw,
                        pgk::TileCells& T) {
    const int64_t colv = (int64_t)K.n[3] * K.n[4] * K.n[5];
    Vw.W = win;
    Vw.miss = miss;
    for (int a = 0; a < 3; ++a) {
        T.t0[a] = t0[a];
        T.tn[a] = tn[a];
        Vw.o[a] = wo[a];
        Vw.L[a] = wl[a];
    }
    T.ncell = (int64_t)tn[0] * tn[1] * tn[2] * colv;
    Vw.ws[5] = 1;
    Vw.ws[4] = K.n[5];
    Vw.ws[3] = (int64_t)K.n[4] * K.n[5];
    Vw.ws[2] = colv;
    Vw.ws[1] = colv * wl[2];
    Vw.ws[0] = Vw.ws[1] * wl[1];
    Vw.size = Vw.ws[0] * wl[0];
    return T.ncell < ((int64_t)1 << 31);
}

hipError_t pgk_launch_escape_tile(const pgk::Consts& K, const int32_t t0[3], const int32_t tn[3],
                                  const float* win, const int32_t wo[3], const int32_t wl[3], float* out,
                                  const uint8_t* occ, const float* bu, const float* uraw, double* part,
                                  unsigned* linf, unsigned* miss, hipStream_t st, int64_t* nblocks) {
    pgk::VWindow Vw;
    pgk::TileCells T;
    if (!tile_reader(K, t0, tn, win, wo, wl, miss, Vw, T)) return hipErrorInvalidValue;
    const int64_t nb = (T.ncell + 255) / 256;
    *nblocks = nb;
    if (prod_scheme(K, nullptr))
        hipLaunchKernelGGL((pgk::sweep_global_kernel<true, false, true, true, pgk::VWindow>), dim3((unsigned)nb),
                           dim3(256), 0, st, Vw, out, T, occ, bu, uraw, K, nullptr, nullptr, part, linf);
    else
        hipLaunchKernelGGL((pgk::sweep_global_kernel<true, false, true, false, pgk::VWindow>), dim3((unsigned)nb),
                           dim3(256), 0, st, Vw, out, T, occ, bu, uraw, K, nullptr, nullptr, part, linf);
    return hipGetLastError();
}

hipError_t pgk_launch_sweep_global_tile(const pgk::Consts& K, const int32_t t0[3], const int32_t tn[3],
                                        const float* win, const int32_t wo[3], const int32_t wl[3], float* out,
                                        const uint8_t* occ, const float* bu, const float* uraw, int8_t* wpol,
                                        double* part, unsigned* linf, unsigned* miss, hipStream_t st,
                                        int64_t* nblocks) {
    pgk::VWindow Vw;
    pgk::TileCells T;
    if (!tile_reader(K, t0, tn, win, wo, wl, miss, Vw, T)) return hipErrorInvalidValue;
    const int64_t nb = (T.ncell + 255) / 256;
    *nblocks = nb;
    if (prod_scheme(K, wpol))
        hipLaunchKernelGGL((pgk::sweep_global_kernel<true, false, false, true, pgk::VWindow>), dim3((unsigned)nb),
                           dim3(256), 0, st, Vw, out, T, occ, bu, uraw, K, nullptr, wpol, part, linf);
    else
        hipLaunchKernelGGL((pgk::sweep_global_kernel<true, false, false, false, pgk::VWindow>), dim3((unsigned)nb),
                           dim3(256), 0, st, Vw, out, T, occ, bu, uraw, K, nullptr, wpol, part, linf);
    return hipGetLastError();
}

hipError_t pgk_launch_reach(const pgk::Consts& K, const float* bu, const float* raw, int ns, int* out,
                            hipStream_t st) {
    const int64_t nst = (int64_t)K.n[2] * K.n[3] * K.n[4] * K.n[5];
    const int64_t nb = (nst + 255) / 256;
    hipLaunchKernelGGL(pgk::reach_kernel, dim3((unsigned)nb), dim3(256), 0, st, K, bu, raw, ns, out);
    return hipGetLastError();
}

hipError_t pgk_launch_reduce(const double* part, int64_t n, double* out, double* mid, hipStream_t st) {
    if (mid && n > 64 * 256) {
        hipLaunchKernelGGL(pgk::reduce_partials_kernel, dim3(kReduceBlocks), dim3(256), 0, st, part, n, mid);
        hipLaunchKernelGGL(pgk::reduce_partials_kernel, dim3(1), dim3(256), 0, st, mid, (int64_t)kReduceBlocks, out);
    } else {
        hipLaunchKernelGGL(pgk::reduce_partials_kernel, dim3(1), dim3(256), 0, st, part, n, out);
    }
    return hipGetLastError();
}

hipError_t pgk_launch_fill(float* V, int64_t n, float v, hipStream_t st) {
    const int64_t nb = (n + 255) / 256;
    hipLaunchKernelGGL(pgk::fill_kernel, dim3((unsigned)nb), dim3(256), 0, st, V, n, v);
    return hipGetLastError();
}

hipError_t pgk_launch_seed(float* V, const int64_t* seeds, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int64_t nb = (n + 255) / 256;
    hipLaunchKernelGGL(pgk::seed_kernel, dim3((unsigned)nb), dim3(256), 0, st, V, seeds, n);
    return hipGetLastError();
}

hipError_t pgk_launch_seed_mark(float* V, const int64_t* seeds, int64_t n, const pgk::Consts& K,
                                const pgk::BoxGeom& G, uint32_t* mark, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int64_t nb = (n + 255) / 256;
    hipLaunchKernelGGL(pgk::seed_mark_kernel, dim3((unsigned)nb), dim3(256), 0, st, V, seeds, n, K, G, mark);
    return hipGetLastError();
}
