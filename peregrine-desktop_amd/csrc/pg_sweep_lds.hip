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
#include <type_traits>

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
    float tau[MAXS], hs[MAXS];   // per-cell horizon (cfl > 0)
    float sh1[MAXS], ch1[MAXS];  // sincos of the heading after the first substep
    unsigned long long best[MAXS];
    uint8_t blocked[MAXS];
    uint8_t nsub[MAXS];
    float bu[3 * kMaxCtlLds];
};

template <int MAXS, bool GEN = true>
__device__ __forceinline__ void load_cell_start(const Consts& K, const BoxShared<MAXS>& sm, int s, CellStart& c) {
    c.x = sm.x[s]; c.y = sm.y[s]; c.h = sm.h[s]; c.vx = sm.vx[s]; c.vy = sm.vy[s]; c.w = sm.w[s];
    c.sh = sm.sh[s]; c.ch = sm.ch[s]; c.sx = sm.sx[s]; c.sy = sm.sy[s]; c.sw = sm.sw[s];
    c.dx = sm.dx[s]; c.dy = sm.dy[s]; c.dw = sm.dw[s]; c.w2 = sm.w2[s];
    if (GEN && K.cfl) {
        c.tau = sm.tau[s]; c.hs = sm.hs[s]; c.nsub = sm.nsub[s];
    } else {
        c.tau = K.tau; c.hs = K.hs; c.nsub = K.nsub;
    }
    cell_sub1(c);
    c.sh1 = sm.sh1[s]; c.ch1 = sm.ch1[s];
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
        sm.tau[s] = c.tau; sm.hs[s] = c.hs; sm.nsub[s] = (uint8_t)c.nsub;
        sm.sh1[s] = c.sh1; sm.ch1[s] = c.ch1;
        const bool blk = occ[((int64_t)i[0] * K.n[1] + i[1]) * K.n[2] + i[2]] != 0;
        sm.blocked[s] = blk ? 1 : 0;
        // blocked cells keep (cap, coast) (PeregrineSolver.jl:1226-1228)
        sm.best[s] = blk ? ((unsigned long long)__float_as_uint(K.cap) << 32)
                         : ((unsigned long long)0x7f800000u << 32);
    }
    for (int j = threadIdx.x; j < 3 * K.nctl; j += kThreads) sm.bu[j] = bu_g[j];
}

// One lookahead up to the successor's fractional coordinates. GEN: every
// scheme option as K sets it (explicit midpoint substeps, adaptive probes,
// velocity clamp); otherwise the default scheme, compiled without them.
// Returns 0 if rejected (value cap), 1 if the successor must be interpolated.
// noprobe: the caller knows the swept probes pass (a box whose every pair
// passed them in the plan; a successor never depends on V), so they are not
// run. REACH: when the probes are reached, reach[] receives the grid x, y
// and unwrapped heading coordinates of the probe chord's far end (c + d, as
// the probes form it; every probe lies between the start and that end on
// each axis) and *probed whether they passed. gc: where the successor's
// grid coordinates take their constants (K, or an LDS copy).
template <bool DIAG, bool GEN, class OCC = OccG, bool REACH = false, class GC = Consts>
__device__ __forceinline__ int successor_t(const Consts& K, const OCC& og, const CellStart& c,
                                           float bux, float buy, float buw, float f[6], bool noprobe = false,
                                           float* reach = nullptr, int* probed = nullptr, const GC* gc = nullptr) {
    float nx, ny, nh, nvx, nvy, nw;
    const bool ok = (GEN && K.rk2) ? rk2_from<DIAG>(K, c, bux, buy, buw, nx, ny, nh, nvx, nvy, nw)
                                   : euler_from<DIAG>(K, c, bux, buy, buw, nx, ny, nh, nvx, nvy, nw);
    const float dx = nx - c.x, dy = ny - c.y, dh = nh - c.h;
    if (!(ok & finite_f(dx) & finite_f(dy) & finite_f(dh) & finite_f(nvx) & finite_f(nvy) &
          finite_f(nw)))
        return 0;
    if constexpr (REACH) {
        reach[0] = qdiv((c.x + dx) - K.lo[0], K.rstep[0]);
        reach[1] = qdiv((c.y + dy) - K.lo[1], K.rstep[1]);
        reach[2] = qdiv((c.h + dh) - K.lo[2], K.rstep[2]);
    }
    if (!noprobe) {
        const bool pass = probes_free<OCC, GEN>(K, og, c.x, c.y, c.h, dx, dy, dh);
        if constexpr (REACH) *probed = pass ? 1 : 0;
        if (!pass) return 0;
    }
    float snh, cnh;
    sincos_ref(nh, snh, cnh);
    const float nvfx = cnh * nvx - snh * nvy;
    const float nvfy = snh * nvx + cnh * nvy;
    if constexpr (std::is_same<GC, Consts>::value) return interp_coords<GEN>(K, nx, ny, nh, nvfx, nvfy, nw, f) ? 1 : 0;
    else return interp_coords<GEN>(K, *gc, nx, ny, nh, nvfx, nvfy, nw, f) ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Plan: per box, the smallest and largest base corner index per axis over
// every (state, control) pair whose successor is interpolated. Heading is
// measured relative to the box's first heading, in (-n3/2, n3/2]. FAST
// plans with the precision form's successor (its read ranges; the window
// origins always come from the exact successor). With pfree, the plan also
// flags the boxes whose swept probes can only meet free cells: no obstacle
// lies in the (x, y, heading) cells between the box's states and the far end
// of any probe chord. The precision form's sweeps, and its read ranges here,
// skip such a box's probes. The exact plan writes two bits: 1, the chords
// meet free cells only (the exact sweeps skip the occupancy lookups, which
// would all answer "free"); 2, every (state, control) pair of the box that
// reaches its probes passes them (the exact sweeps skip the probes: a pair's
// successor, and so its probes' outcome, is the same in every sweep, and the
// plan computes it with the sweeps' own arithmetic). Neither changes a bit.
// ---------------------------------------------------------------------------

// Whether an obstacle lies in x [x0, x1] and y [y0, y1] (clamped to the
// grid) at the headings hb + [h0, h1] (modulo n3; every heading when the
// span covers the circle). Every thread of the block calls it.
__device__ __forceinline__ bool obstacle_in(const Consts& K, const OccG& og, int x0, int x1, int y0, int y1, int hb,
                                            int h0, int h1) {
    x0 = x0 < 0 ? 0 : x0;
    x1 = x1 > K.n[0] - 1 ? K.n[0] - 1 : x1;
    y0 = y0 < 0 ? 0 : y0;
    y1 = y1 > K.n[1] - 1 ? K.n[1] - 1 : y1;
    if (h1 - h0 + 1 >= K.n[2]) { h0 = 0; h1 = K.n[2] - 1; }
    const int nx = x1 >= x0 ? x1 - x0 + 1 : 0, ny = y1 >= y0 ? y1 - y0 + 1 : 0, nh = h1 - h0 + 1;
    int hit = 0;
    for (int j = threadIdx.x; j < nx * ny * nh; j += kThreads) {
        const int ih = j % nh, iy = (j / nh) % ny, ix = j / (nh * ny);
        int ch = hb + h0 + ih;
        ch %= K.n[2];
        if (ch < 0) ch += K.n[2];
        hit |= og(x0 + ix, y0 + iy, ch) ? 1 : 0;
    }
    return __syncthreads_or(hit) != 0;
}

// Widens the (x, y, heading relative to hb) cell range [mn, mx] by the cells
// rint can take between floor(g) - margin and ceil(g) + margin; coordinates
// far outside the grid are clamped (the bounds tests reject them).
__device__ __forceinline__ void widen_reach(const float g[3], int hb, int margin, int mn[3], int mx[3]) {
    const float lo[3] = {g[0], g[1], g[2] - (float)hb};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float fl = __builtin_floorf(lo[a]), ce = __builtin_ceilf(lo[a]);
        const int il = (fl < -1.0e6f ? -1000000 : (int)fl) - margin;
        const int ic = (ce > 1.0e6f ? 1000000 : (int)ce) + margin;
        mn[a] = il < mn[a] ? il : mn[a];
        mx[a] = ic > mx[a] ? ic : mx[a];
    }
}

template <bool DIAG, bool FAST, int MAXS>
__global__ __launch_bounds__(256) void plan_kernel(const uint8_t* __restrict__ occ,
                                                   const float* __restrict__ bu_g, Consts K, FastConsts F,
                                                   BoxGeom G, int16_t* __restrict__ worig,
                                                   uint8_t* __restrict__ wext, uint8_t* __restrict__ pfree) {
    const int64_t box = box_of_block(G);
    if (box >= G.nbox) return;
    __shared__ BoxShared<MAXS> sm;
    __shared__ int mn[6], mx[6];
    __shared__ int rn[3], rx[3];
    BoxInfo B;
    box_info(G, box, B);
    if (threadIdx.x < 6) { mn[threadIdx.x] = 1 << 30; mx[threadIdx.x] = -(1 << 30); }
    if (threadIdx.x < 3) { rn[threadIdx.x] = 1 << 30; rx[threadIdx.x] = -(1 << 30); }
    setup_box<MAXS>(K, B, occ, bu_g, sm);
    __syncthreads();
    const OccG og{occ, K.n[1], K.n[2]};
    const int ns = B.ns;
    const int total = ns * K.nctl;
    int s = threadIdx.x % ns, k = threadIdx.x / ns;
    const int ds = kThreads % ns, dk = kThreads / ns;
    // cells (x, y, heading relative to the box) the probes can reach: the
    // box's own, widened by every chord's far end
    int rmn[3] = {B.s0[0], B.s0[1], 0}, rmx[3] = {B.s0[0] + B.vb[0] - 1, B.s0[1] + B.vb[1] - 1, B.vb[2] - 1};
    bool probe = true;
    if constexpr (FAST) {
        for (int p = threadIdx.x; p < total; p += kThreads) {
            if (!sm.blocked[s]) {
                CellStart c;
                load_cell_start(K, sm, s, c);
                FastStart S;
                fast_start<DIAG>(K, F, c, S);
                float g[3];
                if (successor_reach_fast<DIAG>(K, F, S, sm.bu[3 * k], sm.bu[3 * k + 1], sm.bu[3 * k + 2], g))
                    widen_reach(g, B.s0[2], 0, rmn, rmx);
            }
            s += ds; k += dk;
            if (s >= ns) { s -= ns; ++k; }
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            atomicMin(&rn[a], rmn[a]);
            atomicMax(&rx[a], rmx[a]);
        }
        __syncthreads();
        probe = obstacle_in(K, og, rn[0], rx[0], rn[1], rx[1], B.s0[2], rn[2], rx[2]);
        if (threadIdx.x == 0) pfree[B.gbox] = probe ? 0 : 1;
        s = threadIdx.x % ns;
        k = threadIdx.x / ns;
    }
    int lmn[6], lmx[6];
#pragma unroll
    for (int a = 0; a < 6; ++a) { lmn[a] = 1 << 30; lmx[a] = -(1 << 30); }
    bool allpass = true;
    for (int p = threadIdx.x; p < total; p += kThreads) {
        if (!sm.blocked[s]) {
            CellStart c;
            load_cell_start(K, sm, s, c);
            float f[6];
            bool interp;
            if constexpr (FAST) {
                FastStart S;
                fast_start<DIAG>(K, F, c, S);
                interp = successor_fast<DIAG>(K, F, og, S, sm.bu[3 * k], sm.bu[3 * k + 1], sm.bu[3 * k + 2], probe,
                                              f);
            } else {
                // the chord's far end as the probes form it; one cell of
                // margin each side on top of rint's floor/ceil bracket
                float g[3] = {__builtin_nanf(""), 0.0f, 0.0f};
                int probed = -1;
                interp = successor_t<DIAG, true, OccG, true>(K, og, c, sm.bu[3 * k], sm.bu[3 * k + 1],
                                                             sm.bu[3 * k + 2], f, false, g, &probed) != 0;
                if (finite_f(g[0]) & finite_f(g[1]) & finite_f(g[2])) widen_reach(g, B.s0[2], 1, rmn, rmx);
                allpass = allpass && probed != 0;
            }
            if (interp) {
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
    if constexpr (!FAST) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            atomicMin(&rn[a], rmn[a]);
            atomicMax(&rx[a], rmx[a]);
        }
    }
    __syncthreads();
    if (!FAST && pfree) {
        const bool hit = obstacle_in(K, og, rn[0], rx[0], rn[1], rx[1], B.s0[2], rn[2], rx[2]);
        const bool pass = __syncthreads_and(allpass ? 1 : 0) != 0;
        if (threadIdx.x == 0) pfree[B.gbox] = (uint8_t)((hit ? 0 : 1) | (pass ? 2 : 0));
    }
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
            worig[B.gbox * 6 + a] = (int16_t)o;
            const int ext = hi - lo + 2;
            wext[B.gbox * 6 + a] = (uint8_t)(ext > 255 ? 255 : ext);
        }
    }
}

// ---------------------------------------------------------------------------
// Where a sweep reads V and writes its results: the whole table in HBM, or
// (out-of-core) the window of one tile and that tile's interior.
// ---------------------------------------------------------------------------
// (no __restrict__ on the table: an in-place sweep reads and writes one
// buffer, other workgroups' results arriving while a window is staged)
struct DevSrc {
    const float* V;
    // a window cell to stage (may lie anywhere on the grid)
    __device__ __forceinline__ float stage(const Consts& K, const int gi[6]) const { return read(K, gi); }
    // a value the backup uses
    __device__ __forceinline__ float read(const Consts& K, const int gi[6]) const {
        int64_t fl = 0;
#pragma unroll
        for (int a = 0; a < 6; ++a) fl += (int64_t)gi[a] * K.stride[a];
        return V[fl];
    }
    // one corner of a gather (read_acc), and the gather's end (report): the
    // whole table holds every cell
    __device__ __forceinline__ float read_acc(const Consts& K, const int gi[6], bool&) const { return read(K, gi); }
    __device__ __forceinline__ void report(bool) const {}
};

struct DevDst {
    float* V;
    uint8_t* __restrict__ pol;
    __device__ __forceinline__ void put(const Consts& K, const int i[6], float v, int k) const {
        int64_t fl = 0;
#pragma unroll
        for (int a = 0; a < 6; ++a) fl += (int64_t)i[a] * K.stride[a];
        if (V) V[fl] = v;
        if (pol) pol[fl] = (uint8_t)k;
    }
};

// One tile's window: cells x in [o0, o0 + L0), y in [o1, o1 + L1), heading in
// [o2, o2 + L2) modulo n3, every velocity cell; flat, omega fastest. The
// window holds every corner any backup of the tile reads (it is the union
// of the plan's per-box ranges), so a read outside it is a defect: it
// returns cap and raises *miss. Staged cells outside it are never read as
// corners and stage cap quietly.
struct TileSrc {
    const float* __restrict__ W;
    int32_t o[3], L[3];
    int32_t ws[6];
    unsigned* miss;
    __device__ __forceinline__ bool local(const Consts& K, const int gi[6], int32_t& fl) const {
        int l2 = gi[2] - o[2];
        l2 = l2 < 0 ? l2 + K.n[2] : l2;
        const int l0 = gi[0] - o[0], l1 = gi[1] - o[1];
        // unsigned arithmetic: an index outside the window may wrap, and is
        // never used
        fl = (int32_t)((unsigned)l0 * (unsigned)ws[0] + (unsigned)l1 * (unsigned)ws[1] +
                       (unsigned)l2 * (unsigned)ws[2] + (unsigned)gi[3] * (unsigned)ws[3] +
                       (unsigned)gi[4] * (unsigned)ws[4] + (unsigned)gi[5]);
        return ((unsigned)l0 < (unsigned)L[0]) & ((unsigned)l1 < (unsigned)L[1]) & ((unsigned)l2 < (unsigned)L[2]);
    }
    __device__ __forceinline__ float stage(const Consts& K, const int gi[6]) const {
        int32_t fl;
        return local(K, gi, fl) ? W[fl] : K.cap;
    }
    __device__ __forceinline__ float read(const Consts& K, const int gi[6]) const {
        int32_t fl;
        if (local(K, gi, fl)) return W[fl];
        atomicOr(miss, 1u);
        return K.cap;
    }
    // One corner of a gather without the defect check's branch: cap when
    // outside the window, noted in bad; report raises *miss once for the
    // gather. (A branch and an atomic per corner, over 64 corners, ran the
    // tile kernel out of scalar registers.)
    __device__ __forceinline__ float read_acc(const Consts& K, const int gi[6], bool& bad) const {
        int32_t fl;
        const bool in = local(K, gi, fl);
        bad |= !in;
        const float v = W[in ? fl : 0];
        return in ? v : K.cap;
    }
    __device__ __forceinline__ void report(bool bad) const {
        if (bad) atomicOr(miss, 1u);
    }
};

// The tile's interior: cells from t0 (x, y, heading) over r cells, every
// velocity cell; flat, omega fastest.
struct TileDst {
    float* __restrict__ V;
    uint8_t* __restrict__ pol;
    int32_t t0[3];
    int32_t os[6];
    __device__ __forceinline__ void put(const Consts& K, const int i[6], float v, int k) const {
        const int32_t fl = (i[0] - t0[0]) * os[0] + (i[1] - t0[1]) * os[1] + (i[2] - t0[2]) * os[2] +
                           i[3] * os[3] + i[4] * os[4] + i[5];
        if (V) V[fl] = v;
        if (pol) pol[fl] = (uint8_t)k;
    }
};

// The 64 corners of the successor cell i[] with the reference's clamped /
// periodic indices (PeregrineSolver.jl:455-463), read through src.
template <class SRC>
__device__ __forceinline__ void fetch_src(const Consts& K, const SRC& src, const int i[6], float vals[64]) {
    int j0[6], j1[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        if (k == 2) {
            j0[k] = (i[k] >= K.n[2]) ? i[k] - K.n[2] : i[k];
            const int jj = i[k] + 1;
            j1[k] = (jj >= K.n[2]) ? jj - K.n[2] : jj;
        } else {
            j0[k] = i[k];
            j1[k] = (i[k] + 1 < K.n[k] - 1) ? i[k] + 1 : K.n[k] - 1;
        }
    }
    bool bad = false;
#pragma unroll
    for (int d = 0; d < 64; ++d) {
        int c[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) c[k] = ((d >> k) & 1) ? j1[k] : j0[k];
        vals[d] = src.read_acc(K, c, bad);
    }
    src.report(bad);
}

// The value at the grid node nearest the successor (nearest_cell lookups,
// PeregrineSolver.jl:424-428): per axis the lower or the upper corner of the
// successor's cell as rint(f) falls, from the window or through src.
template <class SRC, class W>
__device__ __forceinline__ float nearest_value(const Consts& K, const SRC& src, const float* win, W,
                                               const int l[6], const int i[6], const float f[6], bool inwin) {
    int c[6];
    int off = 0;
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        const int up = __builtin_rintf(f[a]) != __builtin_floorf(f[a]) ? 1 : 0;
        off += (l[a] + up) * W::S[a];
        int j = i[a] + up;
        if (a == 2) {
            j = j >= K.n[2] ? j - K.n[2] : j;
            j = j >= K.n[2] ? j - K.n[2] : j;
        } else {
            j = j < K.n[a] - 1 ? j : K.n[a] - 1;
        }
        c[a] = j;
    }
    return inwin ? win[off] : src.read(K, c);
}

// The Kuhn simplex interpolant of the successor (simplex lookups): its 7
// vertices from the window or through src.
template <class SRC, class W>
__device__ __forceinline__ float kuhn_value(const Consts& K, const SRC& src, const float* win, W, const int l[6],
                                            const int i[6], const float f[6], bool inwin) {
    int base = 0;
#pragma unroll
    for (int a = 0; a < 6; ++a) base += l[a] * W::S[a];
    return kuhn_interp(f, [&](int mask) {
        if (inwin) {
            int off = base;
#pragma unroll
            for (int a = 0; a < 6; ++a) off += ((mask >> a) & 1) * W::S[a];
            return win[off];
        }
        int c[6];
#pragma unroll
        for (int a = 0; a < 6; ++a) {
            int j = i[a] + ((mask >> a) & 1);
            if (a == 2) {
                j = j >= K.n[2] ? j - K.n[2] : j;
                j = j >= K.n[2] ? j - K.n[2] : j;
            } else {
                j = j < K.n[a] - 1 ? j : K.n[a] - 1;
            }
            c[a] = j;
        }
        return src.read(K, c);
    });
}

// ---------------------------------------------------------------------------
// Sweep
// ---------------------------------------------------------------------------
template <class W, bool DIAG, bool GEN, int MAXS, bool WRITE_V, bool WRITE_POL, class SRC, class DST>
__global__ __launch_bounds__(256, 4) void sweep_lds_kernel(
    SRC src, DST dst, const uint8_t* __restrict__ occ,
    const float* __restrict__ bu_g, Consts K, BoxGeom G, const int16_t* __restrict__ worig,
    const uint8_t* __restrict__ pfree, double* __restrict__ part, unsigned* __restrict__ linf_bits, ActiveG A) {
    __shared__ float win[W::SIZE];
    __shared__ BoxShared<MAXS> sm;
    __shared__ GridC gcs;   // read once per lookahead: 27 fewer scalar registers held across the loop
    const int64_t box = box_of_block(G);
    BoxInfo B;
    int o[6];
    if (box < G.nbox) box_info(G, box, B);
    // live: the box is computed this sweep
    const bool live = box < G.nbox && box_active(A, G, K.n, B);
    if (live) {
        if (A.nwork && threadIdx.x == 0) atomicAdd(A.nwork, 1u);
        if (threadIdx.x == 0) gridc_fill(K, gcs);
#pragma unroll
        for (int a = 0; a < 6; ++a) o[a] = worig[B.gbox * 6 + a];
        // stage the window; cells off the grid hold cap (read only with zero weight)
        for (int j = threadIdx.x; j < W::CELLS; j += kThreads) {
            int l[6], gi[6];
            const bool on = window_cell<W>(K.n, o, j, l, gi);
            const float v = on ? src.stage(K, gi) : K.cap;
            int lo = 0;
#pragma unroll
            for (int a = 0; a < 6; ++a) lo += l[a] * W::S[a];
            win[lo] = v;
        }
        setup_box<MAXS>(K, B, occ, bu_g, sm);
    }
    __syncthreads();
    if (live) {
        // the plan's flags: bit 1, no obstacle along any probe chord of the
        // box (no occupancy lookups); bit 2, every pair's probes pass (none run)
        const unsigned pf = pfree != nullptr ? pfree[B.gbox] : 0u;
        const OccG og{occ, K.n[1], K.n[2], (pf & 1u) != 0};
        const bool noprobe = (pf & 2u) != 0;
        const int ns = B.ns;
        // tps = 256 / ns threads share a state (ns <= 128, so tps >= 2): each
        // holds the state's start in registers and walks scan positions
        // r0, r0 + tps, ... (coast, then the lattice or this sweep's slice of
        // it in scan order), keeping its running minimum in registers with a
        // strict `<` (the earliest position wins a tie); the state's minimum
        // key over its threads then keeps the earliest position among equal
        // values, as the reference's strict `<` scan does
        const int tps = kThreads / ns;
        const int s = threadIdx.x % ns, r0 = threadIdx.x / ns;
        if (r0 < tps && !sm.blocked[s]) {
            float bv = __uint_as_float(0x7f800000u);
            int br = 0;
            for (int r = r0; r <= K.span; r += tps) {
                // the state's start is read from LDS every iteration: held
                // in registers across the loop it would spill
                asm volatile("" ::: "memory");
                CellStart c;
                load_cell_start<MAXS, GEN>(K, sm, s, c);
                float f[6];
                float val = K.cap;
                const int k = slice_control(K, r);
                if (successor_t<DIAG, GEN, OccG, false, GridC>(K, og, c, sm.bu[3 * k], sm.bu[3 * k + 1],
                                                               sm.bu[3 * k + 2], f, noprobe, nullptr, nullptr, &gcs)) {
                    float t[6], u[6];
                    int i[6];
#pragma unroll
                    for (int a = 0; a < 6; ++a) {
                        const float b = __builtin_floorf(f[a]);
                        t[a] = f[a] - b;
                        u[a] = 1.0f - t[a];
                        i[a] = (int)b;
                    }
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
                    float acc;
                    if (GEN && K.nearest) {
                        acc = nearest_value(K, src, win, W{}, l, i, f, inwin);
                    } else if (GEN && K.simplex) {
                        acc = kuhn_value(K, src, win, W{}, l, i, f, inwin);
                    } else {
                    Weights Wt;
                    corner_weights(t, u, Wt);
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
                        fetch_src(K, src, i, vals);
                    }
                    acc = corner_sum(Wt, vals);
                    }
                    const float nv = (acc < K.cap) ? acc : K.cap;
                    const float cc = c.tau + nv;
                    val = (cc < K.cap) ? cc : K.cap;
                }
                // the order of the (value bits, position) keys: for the
                // non-negative values of a value sweep, the order of values
                if (__float_as_uint(val) < __float_as_uint(bv)) {
                    bv = val;
                    br = r;
                }
            }
            atomicMin(&sm.best[s], ((unsigned long long)__float_as_uint(bv) << 32) | (unsigned)br);
        }
    }
    __syncthreads();
    double dsum = 0.0;
    float dmax = 0.0f;
    if (live) {
        for (int s = threadIdx.x; s < B.ns; s += kThreads) {
            int i[6];
            state_coords(B, s, i);
            const unsigned long long key = sm.best[s];
            const float best = __uint_as_float((unsigned)(key >> 32));
            const int bk = slice_control(K, (int)(key & 0xffffffffu));
            float out = 0.0f;
            if (WRITE_V) {
                const float old = src.read(K, i);
                out = old;
                if (best < old) {
                    out = best;
                    const float d = old - best;
                    dsum += (double)d;
                    dmax = fmaxf(dmax, d);
                }
            }
            dst.put(K, i, out, bk);
        }
    }
    if (WRITE_V) {
        const float m = block_reduce_sum_max(dsum, dmax, part + G.blk0, linf_bits);
        if (A.next && m > 0.0f) box_mark(A.next, G, B.gbox);
    }
}

// ---------------------------------------------------------------------------
// Escape sweep (gpu_escape_kernel!, Gpu.jl:51-67; EscapeView / OccView,
// PeregrineSolver.jl:807-874), LDS-staged, Jacobi. A cell with a real route
// (0 <= v < cap) keeps its value; every other cell, blocked ones included,
// takes -c when its backup c beats the escape time it holds (-v if v < 0,
// else cap). Corners read as the escape pass reads them: a stored escape as
// its time, a real route as 0, the rest as cap, and -- for a free updating
// cell -- an obstacle (x, y, h) cell as cap; a blocked updating cell's probes
// and corners ignore the obstacles. The window is staged already read that
// way, walls applied when the box holds a free state; a blocked state of
// such a box reads its corners through global memory. Results equal the
// global-gather escape sweep's bit for bit.
// ---------------------------------------------------------------------------
struct OccE {
    const uint8_t* __restrict__ occ;
    int32_t n2, n3;
    bool ignore;
    __device__ __forceinline__ bool operator()(int ci, int cj, int ck) const {
        return !ignore && occ[((int64_t)ci * n2 + cj) * n3 + ck] != 0;
    }
};

// A stored value as the escape pass reads it: a stored escape as its time,
// a real route as 0, the rest as cap.
__device__ __forceinline__ float escape_view(const Consts& K, float v) {
    return v < 0.0f ? -v : (v < K.cap ? 0.0f : K.cap);
}

__device__ __forceinline__ bool wall_at(const Consts& K, const uint8_t* __restrict__ occ, const int gi[6]) {
    return occ[((int64_t)gi[0] * K.n[1] + gi[1]) * K.n[2] + gi[2]] != 0;
}

// Where an escape sweep reads V. raw: a state's own stored value; read: a
// corner as the escape pass reads it, an obstacle (x, y, h) cell as cap
// when walls apply; stage: the same for a window cell to stage. The whole
// table in HBM:
struct EscDev {
    const float* V;
    const uint8_t* __restrict__ occ;
    __device__ __forceinline__ float raw(const Consts& K, const int gi[6]) const {
        int64_t fl = 0;
#pragma unroll
        for (int a = 0; a < 6; ++a) fl += (int64_t)gi[a] * K.stride[a];
        return V[fl];
    }
    __device__ __forceinline__ float read(const Consts& K, const int gi[6], bool walls) const {
        if (walls && wall_at(K, occ, gi)) return K.cap;
        return escape_view(K, raw(K, gi));
    }
    __device__ __forceinline__ float stage(const Consts& K, const int gi[6], bool walls) const {
        return read(K, gi, walls);
    }
    // one corner of a gather, without branches: both reads are made and
    // the wall selects (a branch per corner over 64 corners runs the kernel
    // out of scalar registers)
    __device__ __forceinline__ float read_acc(const Consts& K, const int gi[6], bool walls, bool&) const {
        const bool wall = walls & wall_at(K, occ, gi);
        const float v = escape_view(K, raw(K, gi));
        return wall ? K.cap : v;
    }
    __device__ __forceinline__ void report(bool) const {}
};

// ... or (out of core) one tile's window, which holds the tile's interior
// and every corner its escape backups read: a read outside it is a defect
// (cap, and *miss raised); a staged cell outside it is never read as a
// corner and stages cap quietly.
struct EscTile {
    TileSrc t;
    const uint8_t* __restrict__ occ;
    __device__ __forceinline__ float raw(const Consts& K, const int gi[6]) const { return t.read(K, gi); }
    __device__ __forceinline__ float read(const Consts& K, const int gi[6], bool walls) const {
        if (walls && wall_at(K, occ, gi)) return K.cap;
        return escape_view(K, t.read(K, gi));
    }
    __device__ __forceinline__ float stage(const Consts& K, const int gi[6], bool walls) const {
        if (walls && wall_at(K, occ, gi)) return K.cap;
        int32_t fl;
        return t.local(K, gi, fl) ? escape_view(K, t.W[fl]) : K.cap;
    }
    __device__ __forceinline__ float read_acc(const Consts& K, const int gi[6], bool walls, bool& bad) const {
        const bool wall = walls & wall_at(K, occ, gi);
        bool out = false;
        const float v = t.read_acc(K, gi, out);
        bad |= out && !wall;   // as read: a wall reads as cap without touching the window
        return wall ? K.cap : escape_view(K, v);
    }
    __device__ __forceinline__ void report(bool bad) const { t.report(bad); }
};

// A source with its walls fixed, as the corner readers take it.
template <class ESRC>
struct EscView {
    const ESRC& src;
    bool walls;
    __device__ __forceinline__ float read(const Consts& K, const int gi[6]) const { return src.read(K, gi, walls); }
    __device__ __forceinline__ float read_acc(const Consts& K, const int gi[6], bool& bad) const {
        return src.read_acc(K, gi, walls, bad);
    }
    __device__ __forceinline__ void report(bool bad) const { src.report(bad); }
};

template <class W, bool DIAG, bool GEN, int MAXS, class ESRC, class DST>
__global__ __launch_bounds__(256, 4) void escape_lds_kernel(
    ESRC src, DST dst, const uint8_t* __restrict__ occ, const float* __restrict__ bu_g,
    Consts K, BoxGeom G, const int16_t* __restrict__ worig, const uint8_t* __restrict__ pfree,
    double* __restrict__ part, unsigned* __restrict__ linf_bits) {
    __shared__ float win[W::SIZE];
    __shared__ BoxShared<MAXS> sm;
    __shared__ uint8_t term[MAXS];
    __shared__ uint8_t blist[MAXS];   // the blocked states still to back up
    __shared__ int nblist;
    const int64_t box = box_of_block(G);
    const bool live = box < G.nbox;
    BoxInfo B;
    int o[6];
    bool free_here = false;
    if (threadIdx.x == 0) nblist = 0;
    __syncthreads();
    if (live) {
        box_info(G, box, B);
#pragma unroll
        for (int a = 0; a < 6; ++a) o[a] = worig[B.gbox * 6 + a];
        setup_box<MAXS>(K, B, occ, bu_g, sm);
        // the same states as setup_box's loop, so no barrier is needed
        for (int s = threadIdx.x; s < B.ns; s += kThreads) {
            int i[6];
            state_coords(B, s, i);
            const float old = src.raw(K, i);
            term[s] = (old >= 0.0f && old < K.cap) ? 1 : 0;
            sm.best[s] = (unsigned long long)0x7f800000u << 32;
            free_here |= sm.blocked[s] == 0;
            if (sm.blocked[s] && !term[s]) blist[atomicAdd(&nblist, 1)] = (uint8_t)s;
        }
    }
    // walls in the staged window when any state of the box is free
    const bool walls = __syncthreads_or(free_here ? 1 : 0) != 0;
    auto stage = [&](bool with_walls) {
        for (int j = threadIdx.x; j < W::CELLS; j += kThreads) {
            int l[6], gi[6];
            const bool on = window_cell<W>(K.n, o, j, l, gi);
            const float v = on ? src.stage(K, gi, with_walls) : K.cap;
            int lo = 0;
#pragma unroll
            for (int a = 0; a < 6; ++a) lo += l[a] * W::S[a];
            win[lo] = v;
        }
    };
    if (live) stage(walls);
    __syncthreads();
    // the value plan's flags (free states only: the plan skips blocked
    // ones, whose probes here ignore the obstacles): bit 1, no obstacle
    // along any probe chord; bit 2, every pair's probes pass
    const unsigned pf = live && pfree != nullptr ? pfree[B.gbox] : 0u;
    const bool pclear = (pf & 1u) != 0, pallpass = (pf & 2u) != 0;
    // Pass 0: free states (and blocked ones when the window holds no walls).
    // Pass 1, for a box that holds walls and blocked states to back up: the
    // window is staged again without the walls and the blocked states' pairs
    // spread over the threads (a blocked state's probes and corners ignore
    // the obstacles; read through global memory instead, every corner of
    // such a pair left the window). One loop, so the backup is compiled once.
    const int npass = (walls && nblist > 0) ? 2 : 1;
#pragma unroll 1
    for (int pass = 0; pass < npass; ++pass) {
        if (pass == 1) {
            __syncthreads();
            if (live) stage(false);
            __syncthreads();
        }
        if (!live) continue;
        const bool walled = pass == 0 && walls;
        const int n = pass == 0 ? B.ns : nblist;
        const int total = n * (K.span + 1);
        int j = threadIdx.x % n, r = threadIdx.x / n;
        const int dj = kThreads % n, dr = kThreads / n;
        for (int p = threadIdx.x; p < total; p += kThreads) {
            const int s = pass == 0 ? j : blist[j];
            if (!term[s] && !(walled && sm.blocked[s])) {
                const bool blk = sm.blocked[s] != 0;
                const OccE og{occ, K.n[1], K.n[2], blk || pclear};
                const EscView<ESRC> er{src, !blk};
                CellStart c;
                load_cell_start<MAXS, GEN>(K, sm, s, c);
                float f[6];
                float val = K.cap;
                const int k = slice_control(K, r);
                if (successor_t<DIAG, GEN>(K, og, c, sm.bu[3 * k], sm.bu[3 * k + 1], sm.bu[3 * k + 2], f,
                                           pallpass && !blk)) {
                    float t[6], u[6];
                    int i[6];
#pragma unroll
                    for (int a = 0; a < 6; ++a) {
                        const float b = __builtin_floorf(f[a]);
                        t[a] = f[a] - b;
                        u[a] = 1.0f - t[a];
                        i[a] = (int)b;
                    }
                    int l[6];
                    bool inwin = !(blk && walled);
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
                    float acc;
                    if (GEN && K.nearest) {
                        acc = nearest_value(K, er, win, W{}, l, i, f, inwin);
                    } else if (GEN && K.simplex) {
                        acc = kuhn_value(K, er, win, W{}, l, i, f, inwin);
                    } else {
                        Weights Wt;
                        corner_weights(t, u, Wt);
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
                            fetch_src(K, er, i, vals);
                        }
                        acc = corner_sum(Wt, vals);
                    }
                    const float nv = (acc < K.cap) ? acc : K.cap;
                    const float cc = c.tau + nv;
                    val = (cc < K.cap) ? cc : K.cap;
                }
                const unsigned long long key = ((unsigned long long)__float_as_uint(val) << 32) | (unsigned)r;
                atomicMin(&sm.best[s], key);
            }
            j += dj; r += dr;
            if (j >= n) { j -= n; ++r; }
        }
    }
    __syncthreads();
    double dsum = 0.0;
    float dmax = 0.0f;
    if (live) {
        for (int s = threadIdx.x; s < B.ns; s += kThreads) {
            int i[6];
            state_coords(B, s, i);
            const float old = src.raw(K, i);
            float out = old;
            if (!term[s]) {
                const float best = __uint_as_float((unsigned)(sm.best[s] >> 32));
                const float e_old = old < 0.0f ? -old : K.cap;
                if (best < e_old) {
                    out = -best;
                    const float d = e_old - best;
                    dsum += (double)d;
                    dmax = fmaxf(dmax, d);
                }
            }
            dst.put(K, i, out, 0);
        }
    }
    block_reduce_sum_max(dsum, dmax, part + G.blk0, linf_bits);
}

}  // namespace pgk

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------

namespace {

template <class W, bool DIAG, bool GEN, int MAXS, class SRC, class DST>
hipError_t launch_w(const pgk::Consts& K, const pgk::BoxGeom& G, const SRC& src, const DST& dst,
                    const uint8_t* occ, const float* bu, const int16_t* worig, const uint8_t* pfree, double* part,
                    unsigned* linf, int write_v, int write_pol, const pgk::ActiveG& A, hipStream_t st) {
    return pgk::for_box_chunks(G, [&](const pgk::BoxGeom& g, unsigned nb) {
        if (write_v && write_pol)
            hipLaunchKernelGGL((pgk::sweep_lds_kernel<W, DIAG, GEN, MAXS, true, true, SRC, DST>), dim3(nb), dim3(256), 0,
                               st, src, dst, occ, bu, K, g, worig, pfree, part, linf, A);
        else if (write_v)
            hipLaunchKernelGGL((pgk::sweep_lds_kernel<W, DIAG, GEN, MAXS, true, false, SRC, DST>), dim3(nb), dim3(256),
                               0, st, src, dst, occ, bu, K, g, worig, pfree, part, linf, A);
        else
            hipLaunchKernelGGL((pgk::sweep_lds_kernel<W, DIAG, GEN, MAXS, false, true, SRC, DST>), dim3(nb), dim3(256),
                               0, st, src, dst, occ, bu, K, g, worig, pfree, part, linf, A);
    });
}

template <class SRC, class DST>
hipError_t launch_shape(int shape, int diag, const pgk::Consts& K, const pgk::BoxGeom& G, const SRC& src,
                        const DST& dst, const uint8_t* occ, const float* bu, const int16_t* worig,
                        const uint8_t* pfree, double* part, unsigned* linf, int write_v, int write_pol,
                        const pgk::ActiveG& A, hipStream_t st) {
    // the general instantiation when any scheme option is set
    const bool gen = K.rk2 || K.nearest || K.simplex || K.cfl || K.adaptive || K.vclamp;
    switch ((shape * 2 + (diag ? 1 : 0)) * 2 + (gen ? 1 : 0)) {
        case 0: return launch_w<pgk::Win0, false, false, 128>(K, G, src, dst, occ, bu, worig, pfree, part, linf, write_v, write_pol, A, st);
        case 1: return launch_w<pgk::Win0, false, true, 128>(K, G, src, dst, occ, bu, worig, pfree, part, linf, write_v, write_pol, A, st);
        case 2: return launch_w<pgk::Win0, true, false, 128>(K, G, src, dst, occ, bu, worig, pfree, part, linf, write_v, write_pol, A, st);
        case 3: return launch_w<pgk::Win0, true, true, 128>(K, G, src, dst, occ, bu, worig, pfree, part, linf, write_v, write_pol, A, st);
        case 4: return launch_w<pgk::Win1, false, false, 128>(K, G, src, dst, occ, bu, worig, pfree, part, linf, write_v, write_pol, A, st);
        case 5: return launch_w<pgk::Win1, false, true, 128>(K, G, src, dst, occ, bu, worig, pfree, part, linf, write_v, write_pol, A, st);
        case 6: return launch_w<pgk::Win1, true, false, 128>(K, G, src, dst, occ, bu, worig, pfree, part, linf, write_v, write_pol, A, st);
        case 7: return launch_w<pgk::Win1, true, true, 128>(K, G, src, dst, occ, bu, worig, pfree, part, linf, write_v, write_pol, A, st);
        case 8: return launch_w<pgk::Win2, false, false, 128>(K, G, src, dst, occ, bu, worig, pfree, part, linf, write_v, write_pol, A, st);
        case 9: return launch_w<pgk::Win2, false, true, 128>(K, G, src, dst, occ, bu, worig, pfree, part, linf, write_v, write_pol, A, st);
        case 10: return launch_w<pgk::Win2, true, false, 128>(K, G, src, dst, occ, bu, worig, pfree, part, linf, write_v, write_pol, A, st);
        case 11: return launch_w<pgk::Win2, true, true, 128>(K, G, src, dst, occ, bu, worig, pfree, part, linf, write_v, write_pol, A, st);
        default: return hipErrorInvalidValue;
    }
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

hipError_t pgk_launch_plan(const pgk::Consts& K, const pgk::FastConsts* F, const pgk::BoxGeom& G,
                           const uint8_t* occ, const float* bu, int16_t* worig, uint8_t* wext, uint8_t* pfree,
                           int diag, hipStream_t st) {
    const pgk::FastConsts f = F ? *F : pgk::FastConsts{};
    return pgk::for_box_chunks(G, [&](const pgk::BoxGeom& g, unsigned nb) {
        const dim3 grid(nb), blk(256);
        switch ((F ? 2 : 0) + (diag ? 1 : 0)) {
            case 0: hipLaunchKernelGGL((pgk::plan_kernel<false, false, 128>), grid, blk, 0, st, occ, bu, K, f, g, worig, wext, pfree); break;
            case 1: hipLaunchKernelGGL((pgk::plan_kernel<true, false, 128>), grid, blk, 0, st, occ, bu, K, f, g, worig, wext, pfree); break;
            case 2: hipLaunchKernelGGL((pgk::plan_kernel<false, true, 128>), grid, blk, 0, st, occ, bu, K, f, g, worig, wext, pfree); break;
            default: hipLaunchKernelGGL((pgk::plan_kernel<true, true, 128>), grid, blk, 0, st, occ, bu, K, f, g, worig, wext, pfree); break;
        }
    });
}

hipError_t pgk_launch_sweep_lds(int shape, int diag, const pgk::Consts& K, const pgk::BoxGeom& G,
                                const float* Vin, float* Vout, const uint8_t* occ, const float* bu,
                                const int16_t* worig, const uint8_t* pfree, uint8_t* pol, double* part,
                                unsigned* linf, int write_v, const pgk::ActiveG& A, hipStream_t st) {
    const pgk::DevSrc src{Vin};
    const pgk::DevDst dst{write_v ? Vout : nullptr, pol};
    return launch_shape(shape, diag, K, G, src, dst, occ, bu, worig, pfree, part, linf, write_v, pol != nullptr, A,
                        st);
}

namespace {

template <class ESRC, class DST>
hipError_t launch_escape_shape(int shape, int diag, const pgk::Consts& K, const pgk::BoxGeom& G, const ESRC& src,
                               const DST& dst, const uint8_t* occ, const float* bu, const int16_t* worig,
                               const uint8_t* pfree, double* part, unsigned* linf, hipStream_t st) {
    const bool gen = K.rk2 || K.nearest || K.simplex || K.cfl || K.adaptive || K.vclamp;
    return pgk::for_box_chunks(G, [&](const pgk::BoxGeom& g, unsigned nb) {
        const dim3 grid(nb), blk(256);
#define PGK_ESC(WS, D, GN) hipLaunchKernelGGL((pgk::escape_lds_kernel<WS, D, GN, 128, ESRC, DST>), grid, blk, 0, st, \
                                             src, dst, occ, bu, K, g, worig, pfree, part, linf)
        switch ((shape * 2 + (diag ? 1 : 0)) * 2 + (gen ? 1 : 0)) {
            case 0: PGK_ESC(pgk::Win0, false, false); break;
            case 1: PGK_ESC(pgk::Win0, false, true); break;
            case 2: PGK_ESC(pgk::Win0, true, false); break;
            case 3: PGK_ESC(pgk::Win0, true, true); break;
            case 4: PGK_ESC(pgk::Win1, false, false); break;
            case 5: PGK_ESC(pgk::Win1, false, true); break;
            case 6: PGK_ESC(pgk::Win1, true, false); break;
            case 7: PGK_ESC(pgk::Win1, true, true); break;
            case 8: PGK_ESC(pgk::Win2, false, false); break;
            case 9: PGK_ESC(pgk::Win2, false, true); break;
            case 10: PGK_ESC(pgk::Win2, true, false); break;
            default: PGK_ESC(pgk::Win2, true, true); break;
        }
#undef PGK_ESC
    });
}

// A tile's window reader and interior writer (x, y, heading cut; every
// velocity cell), both indexed in 32 bits: false when either holds 2^31
// cells or more.
bool tile_io(const pgk::Consts& K, const pgk::BoxGeom& G, const float* win, const int32_t wo[3],
             const int32_t wl[3], unsigned* miss, float* out, uint8_t* pol, pgk::TileSrc& src, pgk::TileDst& dst) {
    src.W = win;
    src.miss = miss;
    dst.V = out;
    dst.pol = pol;
    for (int a = 0; a < 3; ++a) {
        src.o[a] = wo[a];
        src.L[a] = wl[a];
        dst.t0[a] = G.org[a];
    }
    // window strides: L0 x L1 x L2 x n4 x n5 x n6; interior: rn0 x rn1 x rn2 x ...
    const int64_t colv = (int64_t)K.n[3] * K.n[4] * K.n[5];
    if ((int64_t)wl[0] * wl[1] * wl[2] * colv >= ((int64_t)1 << 31) ||
        (int64_t)G.rn[0] * G.rn[1] * G.rn[2] * colv >= ((int64_t)1 << 31))
        return false;
    src.ws[5] = 1;
    src.ws[4] = K.n[5];
    src.ws[3] = K.n[4] * K.n[5];
    src.ws[2] = src.ws[3] * K.n[3];
    src.ws[1] = src.ws[2] * wl[2];
    src.ws[0] = src.ws[1] * wl[1];
    dst.os[5] = 1;
    dst.os[4] = K.n[5];
    dst.os[3] = K.n[4] * K.n[5];
    dst.os[2] = dst.os[3] * K.n[3];
    dst.os[1] = dst.os[2] * G.rn[2];
    dst.os[0] = dst.os[1] * G.rn[1];
    return true;
}

}  // namespace

hipError_t pgk_launch_escape_lds(int shape, int diag, const pgk::Consts& K, const pgk::BoxGeom& G, const float* Vin,
                                 float* Vout, const uint8_t* occ, const float* bu, const int16_t* worig,
                                 const uint8_t* pfree, double* part, unsigned* linf, hipStream_t st) {
    const pgk::EscDev src{Vin, occ};
    const pgk::DevDst dst{Vout, nullptr};
    return launch_escape_shape(shape, diag, K, G, src, dst, occ, bu, worig, pfree, part, linf, st);
}

hipError_t pgk_launch_escape_lds_tile(int shape, int diag, const pgk::Consts& K, const pgk::BoxGeom& G,
                                      const float* win, const int32_t wo[3], const int32_t wl[3], float* out,
                                      const uint8_t* occ, const float* bu, const int16_t* worig,
                                      const uint8_t* pfree, double* part, unsigned* linf, unsigned* miss,
                                      hipStream_t st) {
    pgk::EscTile src;
    pgk::TileDst dst;
    if (!tile_io(K, G, win, wo, wl, miss, out, nullptr, src.t, dst)) return hipErrorInvalidValue;
    src.occ = occ;
    return launch_escape_shape(shape, diag, K, G, src, dst, occ, bu, worig, pfree, part, linf, st);
}

hipError_t pgk_launch_sweep_tile(int shape, int diag, const pgk::Consts& K, const pgk::BoxGeom& G,
                                 const float* win, const int32_t wo[3], const int32_t wl[3], float* out,
                                 uint8_t* pol, const uint8_t* occ, const float* bu, const int16_t* worig,
                                 const uint8_t* pfree, double* part, unsigned* linf, unsigned* miss, hipStream_t st) {
    pgk::TileSrc src;
    pgk::TileDst dst;
    if (!tile_io(K, G, win, wo, wl, miss, out, pol, src, dst)) return hipErrorInvalidValue;
    return launch_shape(shape, diag, K, G, src, dst, occ, bu, worig, pfree, part, linf, out != nullptr,
                        pol != nullptr, pgk::ActiveG{}, st);
}
