// LDS-staged Bellman sweep for gfx950.
//
// One workgroup owns a box of states (pg_box.hpp): it stages the window of V
// that the box's successors can read into LDS, precomputes each state's starting terms, then spreads
// the box's (state, control) pairs over its 256 threads. Every 64-corner
// gather inside the window is 64 ds_read_b32 with immediate offsets from one
// base address; a successor outside the window (rare: the plan sizes the
// window per box) reads global memory with the reference's clamped indices.
// The arithmetic is pg_device.hpp's, so the results equal the global-path
// kernel's, bit for bit.
#include "pg_box.hpp"
#include "pg_kernels.h"

namespace pgk {

constexpr int kMaxCtlLds = 128;   // control table held in LDS (nctl <= 128)

// Per-box shared state: starting terms of each state, the running
// (value, control) minimum per state, and the control table.
template <int MAXS>
struct BoxShared {
    float x[MAXS], y[MAXS], h[MAXS], vx[MAXS], vy[MAXS], w[MAXS];
    float sh[MAXS], ch[MAXS], sx[MAXS], sy[MAXS], sw[MAXS], dx[MAXS], dy[MAXS], dw[MAXS], w2[MAXS];
    unsigned long long best[MAXS];
    uint8_t blocked[MAXS];
    float bu[3 * kMaxCtlLds];
};

template <int MAXS>
__device__ __forceinline__ void load_cell_start(const BoxShared<MAXS>& sm, int s, CellStart& c) {
    c.x = sm.x[s]; c.y = sm.y[s]; c.h = sm.h[s]; c.vx = sm.vx[s]; c.vy = sm.vy[s]; c.w = sm.w[s];
    c.sh = sm.sh[s]; c.ch = sm.ch[s]; c.sx = sm.sx[s]; c.sy = sm.sy[s]; c.sw = sm.sw[s];
    c.dx = sm.dx[s]; c.dy = sm.dy[s]; c.dw = sm.dw[s]; c.w2 = sm.w2[s];
}

template <int MAXS>
__device__ __forceinline__ void setup_box(const Consts& K, const BoxInfo& B,
                                          const uint8_t* __restrict__ occ,
                                          const float* __restrict__ bu_g, BoxShared<MAXS>& sm) {
    for (int s = threadIdx.x; s < B.ns; s += kThreads) {
        int i[6];
        state_coords(B, s, i);
        CellStart c;
        cell_start(K, i[0], i[1], i[2], i[3], i[4], i[5], c);
        sm.x[s] = c.x; sm.y[s] = c.y; sm.h[s] = c.h; sm.vx[s] = c.vx; sm.vy[s] = c.vy; sm.w[s] = c.w;
        sm.sh[s] = c.sh; sm.ch[s] = c.ch; sm.sx[s] = c.sx; sm.sy[s] = c.sy; sm.sw[s] = c.sw;
        sm.dx[s] = c.dx; sm.dy[s] = c.dy; sm.dw[s] = c.dw; sm.w2[s] = c.w2;
        const bool blk = occ[((int64_t)i[0] * K.n[1] + i[1]) * K.n[2] + i[2]] != 0;
        sm.blocked[s] = blk ? 1 : 0;
        // blocked cells keep (cap, coast) (PeregrineSolver.jl:1226-1228)
        sm.best[s] = blk ? ((unsigned long long)__float_as_uint(K.cap) << 32)
                         : ((unsigned long long)0x7f800000u << 32);
    }
    for (int j = threadIdx.x; j < 3 * K.nctl; j += kThreads) sm.bu[j] = bu_g[j];
}

// One lookahead up to the successor's fractional coordinates. Returns 0 if
// rejected (value cap), 1 if the successor must be interpolated.
template <bool DIAG>
__device__ __forceinline__ int successor(const Consts& K, const OccG& og, const CellStart& c,
                                         float bux, float buy, float buw, float f[6]) {
    float nx, ny, nh, nvx, nvy, nw;
    const bool ok = euler_from<DIAG>(K, c, bux, buy, buw, nx, ny, nh, nvx, nvy, nw);
    const float dx = nx - c.x, dy = ny - c.y, dh = nh - c.h;
    if (!(ok & finite_f(dx) & finite_f(dy) & finite_f(dh) & finite_f(nvx) & finite_f(nvy) &
          finite_f(nw)))
        return 0;
    if (!probes_free(K, og, c.x, c.y, c.h, dx, dy, dh)) return 0;
    float snh, cnh;
    sincos_ref(nh, snh, cnh);
    const float nvfx = cnh * nvx - snh * nvy;
    const float nvfy = snh * nvx + cnh * nvy;
    return interp_coords(K, nx, ny, nh, nvfx, nvfy, nw, f) ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Plan: per box, the smallest and largest base corner index per axis over
// every (state, control) pair whose successor is interpolated. Heading is
// measured relative to the box's first heading, in (-n3/2, n3/2].
// ---------------------------------------------------------------------------
template <bool DIAG, int MAXS>
__global__ __launch_bounds__(256) void plan_kernel(const uint8_t* __restrict__ occ,
                                                   const float* __restrict__ bu_g, Consts K,
                                                   BoxGeom G, int16_t* __restrict__ worig,
                                                   uint8_t* __restrict__ wext) {
    const int64_t box = box_of_block(G);
    if (box >= G.nbox) return;
    __shared__ BoxShared<MAXS> sm;
    __shared__ int mn[6], mx[6];
    BoxInfo B;
    box_info(K.n, G, box, B);
    if (threadIdx.x < 6) { mn[threadIdx.x] = 1 << 30; mx[threadIdx.x] = -(1 << 30); }
    setup_box<MAXS>(K, B, occ, bu_g, sm);
    __syncthreads();
    const OccG og{occ, K.n[1], K.n[2]};
    const int ns = B.ns;
    const int total = ns * K.nctl;
    int s = threadIdx.x % ns, k = threadIdx.x / ns;
    const int ds = kThreads % ns, dk = kThreads / ns;
    int lmn[6], lmx[6];
#pragma unroll
    for (int a = 0; a < 6; ++a) { lmn[a] = 1 << 30; lmx[a] = -(1 << 30); }
    for (int p = threadIdx.x; p < total; p += kThreads) {
        if (!sm.blocked[s]) {
            CellStart c;
            load_cell_start(sm, s, c);
            float f[6];
            if (successor<DIAG>(K, og, c, sm.bu[3 * k], sm.bu[3 * k + 1], sm.bu[3 * k + 2], f)) {
#pragma unroll
                for (int a = 0; a < 6; ++a) {
                    int i = (int)__builtin_floorf(f[a]);
                    if (a == 2) {
                        int r = i - B.s0[2];
                        r = r < 0 ? r + K.n[2] : r;
                        r = r >= K.n[2] ? r - K.n[2] : r;
                        i = (2 * r > K.n[2]) ? r - K.n[2] : r;
                    }
                    lmn[a] = i < lmn[a] ? i : lmn[a];
                    lmx[a] = i > lmx[a] ? i : lmx[a];
                }
            }
        }
        s += ds; k += dk;
        if (s >= ns) { s -= ns; ++k; }
    }
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        atomicMin(&mn[a], lmn[a]);
        atomicMax(&mx[a], lmx[a]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int a = 0; a < 6; ++a) {
            int lo = mn[a], hi = mx[a];
            if (lo > hi) { lo = (a == 2) ? 0 : B.s0[a]; hi = lo; }   // nothing interpolated
            int o = lo;
            if (a == 2) {
                o = B.s0[2] + lo;
                o %= K.n[2];
                if (o < 0) o += K.n[2];
            }
            worig[box * 6 + a] = (int16_t)o;
            const int ext = hi - lo + 2;
            wext[box * 6 + a] = (uint8_t)(ext > 255 ? 255 : ext);
        }
    }
}

// ---------------------------------------------------------------------------
// Sweep
// ---------------------------------------------------------------------------
template <class W, bool DIAG, int MAXS, bool WRITE_V, bool WRITE_POL>
__global__ __launch_bounds__(256) void sweep_lds_kernel(
    const float* __restrict__ Vin, float* __restrict__ Vout, const uint8_t* __restrict__ occ,
    const float* __restrict__ bu_g, Consts K, BoxGeom G, const int16_t* __restrict__ worig,
    uint8_t* __restrict__ pol, double* __restrict__ part, unsigned* __restrict__ linf_bits) {
    __shared__ float win[W::SIZE];
    __shared__ BoxShared<MAXS> sm;
    const int64_t box = box_of_block(G);
    const bool live = box < G.nbox;
    BoxInfo B;
    int o[6];
    if (live) {
        box_info(K.n, G, box, B);
#pragma unroll
        for (int a = 0; a < 6; ++a) o[a] = worig[box * 6 + a];
        // stage the window; cells off the grid hold cap (read only with zero weight)
        for (int j = threadIdx.x; j < W::CELLS; j += kThreads) {
            int l[6], gi[6];
            const bool on = window_cell<W>(K.n, o, j, l, gi);
            float v = K.cap;
            if (on) {
                int64_t fl = 0;
#pragma unroll
                for (int a = 0; a < 6; ++a) fl += (int64_t)gi[a] * K.stride[a];
                v = Vin[fl];
            }
            int lo = 0;
#pragma unroll
            for (int a = 0; a < 6; ++a) lo += l[a] * W::S[a];
            win[lo] = v;
        }
        setup_box<MAXS>(K, B, occ, bu_g, sm);
    }
    __syncthreads();
    if (live) {
        const OccG og{occ, K.n[1], K.n[2]};
        const int ns = B.ns;
        const int total = ns * K.nctl;
        int s = threadIdx.x % ns, k = threadIdx.x / ns;
        const int ds = kThreads % ns, dk = kThreads / ns;
        for (int p = threadIdx.x; p < total; p += kThreads) {
            if (!sm.blocked[s]) {
                CellStart c;
                load_cell_start(sm, s, c);
                float f[6];
                float val = K.cap;
                if (successor<DIAG>(K, og, c, sm.bu[3 * k], sm.bu[3 * k + 1], sm.bu[3 * k + 2], f)) {
                    float t[6], u[6];
                    int i[6];
#pragma unroll
                    for (int a = 0; a < 6; ++a) {
                        const float b = __builtin_floorf(f[a]);
                        t[a] = f[a] - b;
                        u[a] = 1.0f - t[a];
                        i[a] = (int)b;
                    }
                    Weights Wt;
                    corner_weights(t, u, Wt);
                    int l[6];
                    bool inwin = true;
#pragma unroll
                    for (int a = 0; a < 6; ++a) {
                        int d = i[a] - o[a];
                        if (a == 2) {
                            d = d < 0 ? d + K.n[2] : d;
                            d = d >= K.n[2] ? d - K.n[2] : d;
                        }
                        l[a] = d;
                        inwin = inwin & ((unsigned)d <= (unsigned)(W::E[a] - 2));
                    }
                    float vals[64];
                    if (inwin) {
                        int base = 0;
#pragma unroll
                        for (int a = 0; a < 6; ++a) base += l[a] * W::S[a];
                        const float* wp = win + base;
#pragma unroll
                        for (int d = 0; d < 64; ++d) {
                            const int off = ((d >> 0) & 1) * W::S1 + ((d >> 1) & 1) * W::S2 +
                                            ((d >> 2) & 1) * W::S3 + ((d >> 3) & 1) * W::S4 +
                                            ((d >> 4) & 1) * W::S5 + ((d >> 5) & 1) * W::S6;
                            vals[d] = wp[off];
                        }
                    } else {
                        fetch_global(K, Vin, i, vals);
                    }
                    const float acc = corner_sum(Wt, vals);
                    const float nv = (acc < K.cap) ? acc : K.cap;
                    const float cc = K.tau + nv;
                    val = (cc < K.cap) ? cc : K.cap;
                }
                const unsigned long long key =
                    ((unsigned long long)__float_as_uint(val) << 32) | (unsigned)k;
                atomicMin(&sm.best[s], key);
            }
            s += ds; k += dk;
            if (s >= ns) { s -= ns; ++k; }
        }
    }
    __syncthreads();
    double dsum = 0.0;
    float dmax = 0.0f;
    if (live) {
        for (int s = threadIdx.x; s < B.ns; s += kThreads) {
            int i[6];
            state_coords(B, s, i);
            int64_t idx = 0;
#pragma unroll
            for (int a = 0; a < 6; ++a) idx += (int64_t)i[a] * K.stride[a];
            const unsigned long long key = sm.best[s];
            const float best = __uint_as_float((unsigned)(key >> 32));
            const int bk = (int)(key & 0xffffffffu);
            if (WRITE_V) {
                const float old = Vin[idx];
                float out = old;
                if (best < old) {
                    out = best;
                    const float d = old - best;
                    dsum += (double)d;
                    dmax = fmaxf(dmax, d);
                }
                Vout[idx] = out;
            }
            if (WRITE_POL) pol[idx] = (uint8_t)bk;
        }
    }
    if (WRITE_V) block_reduce_sum_max(dsum, dmax, part, linf_bits);
}

}  // namespace pgk

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------

namespace {

template <class W, bool DIAG, int MAXS>
hipError_t launch_w(const pgk::Consts& K, const pgk::BoxGeom& G, const float* Vin, float* Vout,
                    const uint8_t* occ, const float* bu, const int16_t* worig, uint8_t* pol,
                    double* part, unsigned* linf, int write_v, hipStream_t st) {
    const unsigned nb = (unsigned)(((G.nbox + 7) / 8) * 8);
    if (write_v && pol)
        hipLaunchKernelGGL((pgk::sweep_lds_kernel<W, DIAG, MAXS, true, true>), dim3(nb), dim3(256), 0,
                           st, Vin, Vout, occ, bu, K, G, worig, pol, part, linf);
    else if (write_v)
        hipLaunchKernelGGL((pgk::sweep_lds_kernel<W, DIAG, MAXS, true, false>), dim3(nb), dim3(256),
                           0, st, Vin, Vout, occ, bu, K, G, worig, pol, part, linf);
    else
        hipLaunchKernelGGL((pgk::sweep_lds_kernel<W, DIAG, MAXS, false, true>), dim3(nb), dim3(256),
                           0, st, Vin, Vout, occ, bu, K, G, worig, pol, part, linf);
    return hipGetLastError();
}

}  // namespace

// Window shapes built into the library; `pgk_window_shapes` lists them so the
// host can pick the smallest that holds a box's successors.
static const int kShapes[][6] = {
    {4, 4, 4, 4, 4, 6},
    {4, 4, 4, 4, 4, 4},
    {4, 4, 4, 5, 5, 5},
};

int pgk_window_shapes(const int (**shapes)[6]) {
    *shapes = kShapes;
    return (int)(sizeof kShapes / sizeof kShapes[0]);
}

hipError_t pgk_launch_plan(const pgk::Consts& K, const pgk::BoxGeom& G, const uint8_t* occ,
                           const float* bu, int16_t* worig, uint8_t* wext, int diag,
                           hipStream_t st) {
    const unsigned nb = (unsigned)(((G.nbox + 7) / 8) * 8);
    if (diag)
        hipLaunchKernelGGL((pgk::plan_kernel<true, 128>), dim3(nb), dim3(256), 0, st, occ, bu, K, G,
                           worig, wext);
    else
        hipLaunchKernelGGL((pgk::plan_kernel<false, 128>), dim3(nb), dim3(256), 0, st, occ, bu, K, G,
                           worig, wext);
    return hipGetLastError();
}

hipError_t pgk_launch_sweep_lds(int shape, int diag, const pgk::Consts& K, const pgk::BoxGeom& G,
                                const float* Vin, float* Vout, const uint8_t* occ, const float* bu,
                                const int16_t* worig, uint8_t* pol, double* part, unsigned* linf,
                                int write_v, hipStream_t st) {
    using W0 = pgk::Win0;
    using W1 = pgk::Win1;
    using W2 = pgk::Win2;
    switch (shape * 2 + (diag ? 1 : 0)) {
        case 0: return launch_w<W0, false, 128>(K, G, Vin, Vout, occ, bu, worig, pol, part, linf, write_v, st);
        case 1: return launch_w<W0, true, 128>(K, G, Vin, Vout, occ, bu, worig, pol, part, linf, write_v, st);
        case 2: return launch_w<W1, false, 128>(K, G, Vin, Vout, occ, bu, worig, pol, part, linf, write_v, st);
        case 3: return launch_w<W1, true, 128>(K, G, Vin, Vout, occ, bu, worig, pol, part, linf, write_v, st);
        case 4: return launch_w<W2, false, 128>(K, G, Vin, Vout, occ, bu, worig, pol, part, linf, write_v, st);
        case 5: return launch_w<W2, true, 128>(K, G, Vin, Vout, occ, bu, worig, pol, part, linf, write_v, st);
        default: return hipErrorInvalidValue;
    }
}
