// Precision-form sweep: contracted FP32 backups over an LDS window, with V
// held in FP32 or as 16-bit codes with a per-tile FP32 offset/scale.
//
// 16-bit stores. The grid is cut into quantization tiles of 4^6 index cells
// (edge tiles partial). A cell's value is off_t + scale_t * q with q a 16-bit
// code -- IEEE binary16 (11 significant bits of (v - off)) or fixed point
// (65,535 uniform levels over the tile's range) -- and unreached cells (value
// cap) hold an out-of-band code, so the cap never widens a tile's range.
// A sweep decodes into the FP32 LDS window, writes its FP32 results to a
// scratch buffer and each box's min/max of reached values into its tile's
// statistics. A tile's offset/scale are refit only when its values leave the
// representable range (with headroom below, since values decrease);
// otherwise only the cells whose value changed are re-encoded, so a settled
// cell keeps its code bit for bit.
#include "pg_box.hpp"
#include "pg_fast.hpp"
#include "pg_kernels.h"

#include <type_traits>

namespace pgk {

namespace fastk {

constexpr int kMaxCtl = 128;
constexpr uint16_t kCapF16 = 0x7c00;   // binary16 +inf
constexpr uint16_t kCapU16 = 0xffff;
constexpr float kNcodeF16 = 1000.0f;   // fitted range maps to codes [0, 1000] (< 1024: one binade spacing)
constexpr float kNcodeU16 = 65534.0f;  // fitted range maps to codes [0, 65534]

template <int STORE>
struct Codec;

template <>
struct Codec<PG_STORE_F16> {
    static constexpr float ncode = kNcodeF16;
    static __device__ __forceinline__ float decode(uint16_t q, float2 p, float cap) {
        if (q == kCapF16) return cap;
        const float v = p.x + p.y * (float)__builtin_bit_cast(_Float16, q);
        return v < cap ? v : cap;
    }
    static __device__ __forceinline__ uint16_t encode(float v, float2 p, float cap) {
        if (!(v < cap)) return kCapF16;
        return __builtin_bit_cast(uint16_t, (_Float16)((v - p.x) / p.y));
    }
};

template <>
struct Codec<PG_STORE_U16> {
    static constexpr float ncode = kNcodeU16;
    static __device__ __forceinline__ float decode(uint16_t q, float2 p, float cap) {
        if (q == kCapU16) return cap;
        const float v = p.x + p.y * (float)q;
        return v < cap ? v : cap;
    }
    static __device__ __forceinline__ uint16_t encode(float v, float2 p, float cap) {
        if (!(v < cap)) return kCapU16;
        float u = __builtin_rintf((v - p.x) / p.y);
        u = u < 0.0f ? 0.0f : (u > kNcodeU16 ? kNcodeU16 : u);
        return (uint16_t)u;
    }
};

__device__ __forceinline__ int64_t tile_of_coords(const QuantGeom& Q, const int* gi) {
    int64_t t = 0;
#pragma unroll
    for (int a = 0; a < 6; ++a) t = t * Q.nt[a] + (gi[a] >> kQuantShift);
    return t;
}

__device__ __forceinline__ int64_t flat_of(const Consts& K, const int* gi) {
    int64_t fl = 0;
#pragma unroll
    for (int a = 0; a < 6; ++a) fl += (int64_t)gi[a] * K.stride[a];
    return fl;
}

// Reads of V in either store: stage() for a window cell, read() for a value
// a backup uses. (no __restrict__ on the table: in place, a sweep writes
// what it reads)
// An in-HBM table holds fewer than 2^32 cells (the solver refuses larger
// grids), so its cell and tile indices are 32-bit.
__device__ __forceinline__ uint32_t flat32(const Consts& K, const int* gi) {
    uint32_t fl = 0;
#pragma unroll
    for (int a = 0; a < 6; ++a) fl += (uint32_t)gi[a] * (uint32_t)K.stride[a];
    return fl;
}

__device__ __forceinline__ uint32_t tile32(const QuantGeom& Q, const int* gi) {
    uint32_t t = 0;
#pragma unroll
    for (int a = 0; a < 6; ++a) t = t * (uint32_t)Q.nt[a] + (uint32_t)(gi[a] >> kQuantShift);
    return t;
}

template <int STORE>
struct VRead {
    const uint16_t* V;
    const float2* __restrict__ tp;
    __device__ __forceinline__ float read(const Consts& K, const QuantGeom& Q, const int* gi) const {
        return Codec<STORE>::decode(V[flat32(K, gi)], tp[tile32(Q, gi)], K.cap);
    }
    __device__ __forceinline__ float stage(const Consts& K, const QuantGeom& Q, const int* gi) const {
        return read(K, Q, gi);
    }
};

template <>
struct VRead<PG_STORE_F32> {
    const float* V;
    __device__ __forceinline__ float read(const Consts& K, const QuantGeom&, const int* gi) const {
        return V[flat32(K, gi)];
    }
    __device__ __forceinline__ float stage(const Consts& K, const QuantGeom& Q, const int* gi) const {
        return read(K, Q, gi);
    }
};

// An out-of-core tile's window of 16-bit codes: cells x in [o0, o0 + L0),
// y in [o1, o1 + L1), heading in [o2, o2 + L2) modulo n3, every velocity
// cell; omega fastest. Offsets/scales stay those of the whole grid. The
// window holds every cell the tile's backups read, so a read outside it is a
// defect: it returns cap and raises *miss. Staged cells outside it are never
// read as corners and stage cap quietly. A window holds fewer than 2^31
// cells (the host checks), so its indices are 32-bit.
template <int STORE>
struct VReadTile {
    const uint16_t* __restrict__ W;
    const float2* __restrict__ tp;
    int32_t o[3], L[3];
    int32_t ws[6];
    unsigned* miss;
    __device__ __forceinline__ bool local(const Consts& K, const int* gi, int32_t& fl) const {
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
    __device__ __forceinline__ float stage(const Consts& K, const QuantGeom& Q, const int* gi) const {
        int32_t fl;
        return local(K, gi, fl) ? Codec<STORE>::decode(W[fl], tp[tile_of_coords(Q, gi)], K.cap) : K.cap;
    }
    __device__ __forceinline__ float read(const Consts& K, const QuantGeom& Q, const int* gi) const {
        int32_t fl;
        if (local(K, gi, fl)) return Codec<STORE>::decode(W[fl], tp[tile_of_coords(Q, gi)], K.cap);
        atomicOr(miss, 1u);
        return K.cap;
    }
};

// The states' own FP32 values and the sweep's FP32 results: the whole grid,
// or an out-of-core tile's interior (x, y, heading from t0; omega fastest;
// fewer than 2^31 cells).
struct SelfGlobal {
    const float* self;
    float* out;
    __device__ __forceinline__ int64_t index(const Consts& K, const int* i) const { return flat_of(K, i); }
};

struct SelfTile {
    const float* __restrict__ self;
    float* __restrict__ out;
    int32_t t0[3];
    int32_t os[6];
    __device__ __forceinline__ int32_t index(const Consts&, const int* i) const {
        return (i[0] - t0[0]) * os[0] + (i[1] - t0[1]) * os[1] + (i[2] - t0[2]) * os[2] + i[3] * os[3] +
               i[4] * os[4] + i[5];
    }
};

template <int MAXS>
struct BoxShared {
    unsigned long long best[MAXS];   // per state: (value bits, control) minimum
    float bu[3 * kMaxCtl];
    unsigned bmin, bmax;
};

}  // namespace fastk

// One workgroup per box. The tps = 256 / ns threads of a state share its
// controls (k = k0, k0 + tps, ...): each thread holds the state's
// control-independent start in registers, keeps its running minimum of
// (value, control) in registers (strict <, so the lowest control wins a
// tie) and folds it into the state's LDS minimum once. Threads 0..ns-1
// then write the states.
template <class W, bool DIAG, int STORE, int MAXS, bool WRITE_V, bool WRITE_POL, class RD, class IO>
__global__ __launch_bounds__(256, 4) void sweep_fast_kernel(
    RD vin, IO io, uint16_t* codes_out, const uint8_t* __restrict__ occ,
    const float* __restrict__ bu_g, Consts K, FastConsts F, BoxGeom G, QuantGeom Q,
    const int16_t* __restrict__ worig, const uint8_t* __restrict__ pfree, uint8_t* __restrict__ pol,
    double* __restrict__ part, unsigned* __restrict__ linf_bits, unsigned* __restrict__ tstat, ActiveG A) {
    using namespace fastk;
    // the 16-bit stores solve each backup's diagonal (see below); with the
    // FP32 store the kernel is the contracted twin of the exact backup
    constexpr bool SELF = STORE != PG_STORE_F32;
    __shared__ float win[W::SIZE];
    __shared__ BoxShared<MAXS> sm;
    const int64_t box = box_of_block(G);
    BoxInfo B;
    int o[6];
    if (threadIdx.x == 0) { sm.bmin = 0x7f800000u; sm.bmax = 0u; }
    if (box < G.nbox) box_info(G, box, B);
    // live: the box is computed this sweep; a skipped box keeps its values
    const bool live = box < G.nbox && box_active(A, G, K.n, B);
    if (live) {
        if (A.nwork && threadIdx.x == 0) atomicAdd(A.nwork, 1u);
#pragma unroll
        for (int a = 0; a < 6; ++a) o[a] = worig[B.gbox * 6 + a];
        for (int j = threadIdx.x; j < W::CELLS; j += kThreads) {
            int l[6], gi[6];
            const bool on = window_cell<W>(K.n, o, j, l, gi);
            float v = K.cap;
            if (on) v = vin.stage(K, Q, gi);
            int lo = 0;
#pragma unroll
            for (int a = 0; a < 6; ++a) lo += l[a] * W::S[a];
            win[lo] = v;
        }
        for (int j = threadIdx.x; j < B.ns; j += kThreads) {
            int i[6];
            state_coords(B, j, i);
            const bool blk = occ[((int64_t)i[0] * K.n[1] + i[1]) * K.n[2] + i[2]] != 0;
            // blocked cells keep (cap, coast) (PeregrineSolver.jl:1226-1228)
            sm.best[j] = blk ? ((unsigned long long)__float_as_uint(K.cap) << 32)
                             : ((unsigned long long)0x7f800000u << 32);
        }
        for (int j = threadIdx.x; j < 3 * K.nctl; j += kThreads) sm.bu[j] = bu_g[j];
    }
    __syncthreads();
    if (live) {
        const int ns = B.ns;
        const int tps = kThreads / ns;
        const int s = threadIdx.x % ns, k0 = threadIdx.x / ns;
        int i[6];
        state_coords(B, s, i);
        const bool blk = occ[((int64_t)i[0] * K.n[1] + i[1]) * K.n[2] + i[2]] != 0;
        if (k0 < tps && !blk) {
            const OccG og{occ, K.n[1], K.n[2]};
            const bool probe = pfree[B.gbox] == 0;
            FastStart S;
            {
                CellStart c;
                cell_start(K, i[0], i[1], i[2], i[3], i[4], i[5], c);
                fast_start<DIAG>(K, F, c, S);
            }
            // the state's own cell in window coordinates, 4 bits per axis
            // (15: outside the window, so never a corner of an in-window
            // successor)
            unsigned pl = 0;
            float vmir = 0.0f;
            if (SELF) {
#pragma unroll
                for (int a = 0; a < 6; ++a) {
                    int d = i[a] - o[a];
                    if (a == 2) {
                        d = d < 0 ? d + K.n[2] : d;
                        d = d >= K.n[2] ? d - K.n[2] : d;
                    }
                    pl |= (unsigned)((unsigned)d < (unsigned)W::E[a] ? d : 15) << (4 * a);
                }
                vmir = vin.read(K, Q, i);
            }
            // One backup: the value of control k, or false when a corner of
            // its successor lies outside the window and OUTSIDE is not set
            // (the main loop leaves those pairs to a second pass, so that
            // their addressed reads do not raise the loop's register
            // pressure).
            auto backup = [&](int k, auto outside, float& val) -> bool {
                constexpr bool OUTSIDE = decltype(outside)::value;
                float f[6];
                val = K.cap;
                if (!successor_fast<DIAG>(K, F, og, S, sm.bu[3 * k], sm.bu[3 * k + 1], sm.bu[3 * k + 2], probe, f))
                    return true;
                float t[6];
                int si[6];
                int l[6];
                bool inwin = true;
#pragma unroll
                for (int a = 0; a < 6; ++a) {
                    const float b = __builtin_floorf(f[a]);
                    t[a] = f[a] - b;
                    si[a] = (int)b;
                    int d = si[a] - o[a];
                    if (a == 2) {
                        d = d < 0 ? d + K.n[2] : d;
                        d = d >= K.n[2] ? d - K.n[2] : d;
                    }
                    l[a] = d;
                    inwin = inwin & ((unsigned)d <= (unsigned)(W::E[a] - 2));
                }
                // Weight of the state's own cell in the successor's stencil:
                // the product over the axes of 1 - t where the cell is the
                // lower corner and t where it is the upper one, 0 when it is
                // neither on some axis.
                float wself = 0.0f;
                float acc;
                if constexpr (OUTSIDE) {
                    if (SELF) {
                        int own[6];
                        state_coords(B, s, own);
                        wself = 1.0f;
#pragma unroll
                        for (int a = 0; a < 6; ++a) {
                            int d = own[a] - si[a];
                            if (a == 2) d = d < 0 ? d + K.n[2] : d;
                            wself *= d == 0 ? 1.0f - t[a] : (d == 1 ? t[a] : 0.0f);
                        }
                    }
                    // corners outside the window: global memory, the
                    // reference's clamped / periodic indices
                    acc = tensor_interp_rolled(
                        [&](int d) {
                            int gi[6];
#pragma unroll
                            for (int a = 0; a < 6; ++a) {
                                int j = si[a] + ((d >> a) & 1);
                                if (a == 2) {
                                    j = j >= K.n[2] ? j - K.n[2] : j;
                                    j = j >= K.n[2] ? j - K.n[2] : j;
                                } else {
                                    j = j < K.n[a] - 1 ? j : K.n[a] - 1;
                                }
                                gi[a] = j;
                            }
                            return vin.read(K, Q, gi);
                        },
                        t);
                } else {
                    if (!inwin) return false;
                    if (SELF) {
                        // window coordinates: the cell is a corner on axis a
                        // when its offset from the base corner is 0 or 1
                        unsigned bad = 0;
                        float ws = 1.0f;
#pragma unroll
                        for (int a = 0; a < 6; ++a) {
                            const int d = (int)((pl >> (4 * a)) & 15u) - l[a];
                            bad |= (unsigned)d;
                            ws *= d == 0 ? 1.0f - t[a] : t[a];
                        }
                        wself = bad > 1u ? 0.0f : ws;
                    }
                    int base = 0;
#pragma unroll
                    for (int a = 0; a < 6; ++a) base += l[a] * W::S[a];
                    const float* wp = win + base;
                    acc = tensor_interp_grouped(
                        [&](int d) {
                            return wp[((d >> 0) & 1) * W::S1 + ((d >> 1) & 1) * W::S2 + ((d >> 2) & 1) * W::S3 +
                                      ((d >> 3) & 1) * W::S4 + ((d >> 4) & 1) * W::S5 + ((d >> 5) & 1) * W::S6];
                        },
                        t);
                }
                if (SELF) {
                    // V = tau + w V + rest solved for V (Jacobi on the
                    // policy system with its diagonal): same fixed point
                    // as tau + min(acc, cap), without the 1/(1-w)-times
                    // slower approach through the cell's own value. With
                    // w = 0 this is tau + acc clamped to cap, the other
                    // branch's value (acc and vmir are finite), without a
                    // branch.
                    const float rest = acc - wself * vmir;
                    const float r = wself > 0.0f ? __builtin_amdgcn_rcpf(1.0f - wself) : 1.0f;
                    const float cc = (K.tau + rest) * r;
                    val = cc < K.cap ? cc : K.cap;
                } else {
                    const float nv = acc < K.cap ? acc : K.cap;
                    const float cc = K.tau + nv;
                    val = cc < K.cap ? cc : K.cap;
                }
                return true;
            };
            float bv = __uint_as_float(0x7f800000u);
            int bk = 0;
            // bit j: the pair of the thread's j-th control was left for the
            // second pass (a thread holds at most 128 / 2 = 64 controls)
            unsigned long long deferred = 0;
            int j = 0;
            for (int k = k0; k < K.nctl; k += tps, ++j) {
                float val;
                if (!backup(k, std::false_type{}, val)) {
                    deferred |= 1ull << j;
                    continue;
                }
                if (val < bv) {
                    bv = val;
                    bk = k;
                }
            }
            // The deferred pairs, merged as the scan would have taken them:
            // the smallest value, the lowest control among equal values.
            while (deferred) {
                const int jj = __builtin_ctzll(deferred);
                deferred &= deferred - 1;
                const int k = k0 + jj * tps;
                float val;
                backup(k, std::true_type{}, val);
                if (val < bv || (val == bv && k < bk)) {
                    bv = val;
                    bk = k;
                }
            }
            atomicMin(&sm.best[s], ((unsigned long long)__float_as_uint(bv) << 32) | (unsigned)bk);
        }
    }
    __syncthreads();
    double dsum = 0.0;
    float dmax = 0.0f;
    unsigned lmin = 0x7f800000u, lmax = 0u;
    if (live) {
        for (int s = threadIdx.x; s < B.ns; s += kThreads) {
            int i[6];
            state_coords(B, s, i);
            int64_t idx = 0;
#pragma unroll
            for (int a = 0; a < 6; ++a) idx += (int64_t)i[a] * K.stride[a];
            const unsigned long long key = sm.best[s];
            const float best = __uint_as_float((unsigned)(key >> 32));
            if (WRITE_V) {
                const int64_t io_i = io.index(K, i);
                const float old = io.self[io_i];
                float out = old;
                if (best < old) {
                    out = best;
                    const float d = old - best;
                    dsum += (double)d;
                    dmax = fmaxf(dmax, d);
                }
                io.out[io_i] = out;
                // in place: this state's code now, for the boxes swept after it
                if constexpr (STORE != PG_STORE_F32) {
                    if (codes_out) codes_out[idx] = Codec<STORE>::encode(out, vin.tp[tile_of_coords(Q, i)], K.cap);
                }
                if (out < K.cap) {
                    lmin = min(lmin, __float_as_uint(out));
                    lmax = max(lmax, __float_as_uint(out));
                }
            }
            if (WRITE_POL) pol[idx] = (uint8_t)(key & 0xffffffffu);
        }
    } else if (STORE != PG_STORE_F32 && WRITE_V && box < G.nbox) {
        // skipped: its unchanged values still count in its tile's statistics
        for (int s = threadIdx.x; s < B.ns; s += kThreads) {
            int i[6];
            state_coords(B, s, i);
            int64_t idx = 0;
#pragma unroll
            for (int a = 0; a < 6; ++a) idx += (int64_t)i[a] * K.stride[a];
            const float out = io.self[io.index(K, i)];
            if (out < K.cap) {
                lmin = min(lmin, __float_as_uint(out));
                lmax = max(lmax, __float_as_uint(out));
            }
        }
    }
    if (WRITE_V) {
        if (STORE != PG_STORE_F32) {
            atomicMin(&sm.bmin, lmin);
            atomicMax(&sm.bmax, lmax);
        }
        const float m = block_reduce_sum_max(dsum, dmax, part + G.blk0, linf_bits);
        if (A.next && m > 0.0f) box_mark(A.next, G, B.gbox);
        if (STORE != PG_STORE_F32 && box < G.nbox && threadIdx.x == 0 && sm.bmin <= sm.bmax) {
            int64_t t = 0;
#pragma unroll
            for (int a = 0; a < 6; ++a) t = t * Q.nt[a] + (B.s0[a] >> kQuantShift);
            atomicMin(&tstat[2 * t], sm.bmin);
            atomicMax(&tstat[2 * t + 1], sm.bmax);
        }
    }
}

// The seeds are reset to 0 after the sweep kernel took its statistics; fold
// that 0 into their tiles so that every seed stays exactly on the code grid.
__global__ __launch_bounds__(256) void seed_stats_kernel(const int64_t* __restrict__ seeds, int64_t nseeds,
                                                         Consts K, QuantGeom Q, unsigned* __restrict__ tstat) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nseeds) return;
    int64_t r = seeds[j], t = 0, mul = 1;
    for (int a = 5; a >= 0; --a) {
        const int i = (int)(r % K.n[a]);
        r /= K.n[a];
        t += (int64_t)(i >> kQuantShift) * mul;
        mul *= Q.nt[a];
    }
    atomicMin(&tstat[2 * t], 0u);
}

// Per-tile statistics of an FP32 table (used when a table is loaded whole).
__global__ __launch_bounds__(256) void tile_stats_kernel(const float* __restrict__ V, Consts K,
                                                         QuantGeom Q, unsigned* __restrict__ tstat) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= K.ncell) return;
    const float v = V[idx];
    if (!(v < K.cap)) return;
    int64_t r = idx, t = 0, mul = 1;
    for (int a = 5; a >= 0; --a) {
        const int i = (int)(r % K.n[a]);
        r /= K.n[a];
        t += (int64_t)(i >> kQuantShift) * mul;
        mul *= Q.nt[a];
    }
    atomicMin(&tstat[2 * t], __float_as_uint(v));
    atomicMax(&tstat[2 * t + 1], __float_as_uint(v));
}

// Refit each tile's offset/scale when its reached values leave the current
// representable range [off, off + ncode*scale] (or always, with force), and
// flag the refit tiles; clear the statistics. A refit leaves headroom of a
// quarter of the span below the minimum, since values only decrease.
// Refit of tile t from its statistics (which it resets): the offset/scale
// read from tpo, the result written to tpn (the same array in HBM; the
// previous and the next sweep's arrays out of core).
__device__ __forceinline__ void refit_tile(int64_t t, unsigned* __restrict__ tstat, const float2* tpo, float2* tpn,
                                           uint8_t* __restrict__ trefit, float ncode, int force) {
    const unsigned mnb = tstat[2 * t], mxb = tstat[2 * t + 1];
    tstat[2 * t] = 0x7f800000u;
    tstat[2 * t + 1] = 0u;
    uint8_t flag = 0;
    if (mnb <= mxb) {
        const float mn = __uint_as_float(mnb), mx = __uint_as_float(mxb);
        const float2 p = tpo[t];
        const bool unset = !(p.y > 0.0f);
        const float top = p.x + p.y * ncode;
        if (force || unset || mn < p.x || mx > top) {
            const float span = mx - mn;
            // values are >= 0, so the headroom stops at 0 (a seed's 0 stays
            // exactly on the code grid)
            float off, scale;
            if (span > 0.0f) {
                off = fmaxf(mn - 0.25f * span, 0.0f);
                scale = (mx - off) / ncode;
            } else {
                const float mag = fmaxf(fabsf(mn), 1.0f);
                off = fmaxf(mn - mag * (1.0f / 1024.0f), 0.0f);
                scale = (mn + mag * (1.0f / 1024.0f) - off) / ncode;
            }
            tpn[t] = make_float2(off, scale);
            flag = 1;
        } else {
            tpn[t] = p;
        }
    } else if (force) {
        tpn[t] = make_float2(0.0f, 0.0f);   // nothing reached: unset
        flag = 1;
    } else {
        tpn[t] = tpo[t];
    }
    trefit[t] = flag;
}

__global__ __launch_bounds__(256) void tile_refit_kernel(unsigned* __restrict__ tstat, float2* tp,
                                                         uint8_t* __restrict__ trefit, int64_t ntile,
                                                         float ncode, int force) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntile) return;
    refit_tile(t, tstat, tp, tp, trefit, ncode, force);
}

// An out-of-core tile's interior: x, y, heading cells t0 .. t0 + tn, every
// velocity cell, omega fastest; its quantization tiles lie inside it (the
// interior is cut on multiples of the tile edge).
struct Interior {
    int32_t t0[3], tn[3];
    __device__ __forceinline__ void coords(const Consts& K, int64_t j, int* gi) const {
        for (int a = 5; a >= 3; --a) {
            gi[a] = (int)(j % K.n[a]);
            j /= K.n[a];
        }
        gi[2] = t0[2] + (int)(j % tn[2]);
        j /= tn[2];
        gi[1] = t0[1] + (int)(j % tn[1]);
        gi[0] = t0[0] + (int)(j / tn[1]);
    }
};

// Statistics of an interior buffer's reached values into the grid's tiles.
__global__ __launch_bounds__(256) void interior_stats_kernel(const float* __restrict__ V, int64_t n, Interior I,
                                                             Consts K, QuantGeom Q, unsigned* __restrict__ tstat) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const float v = V[j];
    if (!(v < K.cap)) return;
    int gi[6];
    I.coords(K, j, gi);
    const int64_t t = fastk::tile_of_coords(Q, gi);
    atomicMin(&tstat[2 * t], __float_as_uint(v));
    atomicMax(&tstat[2 * t + 1], __float_as_uint(v));
}

// Refit of the tiles under an interior.
__global__ __launch_bounds__(256) void interior_refit_kernel(Interior I, QuantGeom Q, unsigned* __restrict__ tstat,
                                                             const float2* __restrict__ tpo, float2* __restrict__ tpn,
                                                             uint8_t* __restrict__ trefit, float ncode, int force) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int32_t q0[3], qn[3];
    int64_t nv = 1, n = 1;
    for (int a = 0; a < 3; ++a) {
        q0[a] = I.t0[a] >> kQuantShift;
        qn[a] = ((I.t0[a] + I.tn[a] - 1) >> kQuantShift) - q0[a] + 1;
        n *= qn[a];
    }
    for (int a = 3; a < 6; ++a) nv *= Q.nt[a];
    if (j >= n * nv) return;
    const int64_t v = j % nv;
    j /= nv;
    const int32_t c2 = (int32_t)(j % qn[2]);
    j /= qn[2];
    const int32_t c1 = (int32_t)(j % qn[1]), c0 = (int32_t)(j / qn[1]);
    const int64_t t = ((((int64_t)(q0[0] + c0) * Q.nt[1] + (q0[1] + c1)) * Q.nt[2] + (q0[2] + c2)) * nv) + v;
    refit_tile(t, tstat, tpo, tpn, trefit, ncode, force);
}

// Codes of an interior buffer.
template <int STORE>
__global__ __launch_bounds__(256) void interior_encode_kernel(const float* __restrict__ V, int64_t n, Interior I,
                                                              Consts K, QuantGeom Q, const float2* __restrict__ tp,
                                                              uint16_t* __restrict__ codes) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int gi[6];
    I.coords(K, j, gi);
    codes[j] = fastk::Codec<STORE>::encode(V[j], tp[fastk::tile_of_coords(Q, gi)], K.cap);
}

// FP32 table -> codes. In a refit tile every cell is encoded afresh; in
// any other tile only a cell whose code changed is written.
// With `mark`, the box of every cell whose decoded value changes (a new code,
// or a refit tile) is marked changed for the next sweep's active set; the
// lanes of a wave that hit one word of bits fold them into one atomic.
template <int STORE>
__global__ __launch_bounds__(256) void encode_kernel(const float* __restrict__ V32, Consts K,
                                                     QuantGeom Q, const float2* __restrict__ tp,
                                                     const uint8_t* __restrict__ trefit,
                                                     uint16_t* __restrict__ codes, uint32_t* __restrict__ mark,
                                                     BoxGeom G) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool changed = false;
    int64_t word = 0;
    uint32_t bit = 0;
    if (idx < K.ncell) {
        int64_t r = idx;
        int i[6];
        for (int a = 5; a >= 0; --a) {
            i[a] = (int)(r % K.n[a]);
            r /= K.n[a];
        }
        const int64_t t = fastk::tile_of_coords(Q, i);
        const uint16_t qn = fastk::Codec<STORE>::encode(V32[idx], tp[t], K.cap);
        if (trefit[t] || qn != codes[idx]) {
            codes[idx] = qn;
            changed = true;
            const int64_t gb = box_of_cell(G, i);
            word = gb / G.nbg[5];
            bit = 1u << (int)(gb % G.nbg[5]);
        }
    }
    if (!mark) return;
    unsigned long long pending = __ballot(changed);
    const int lane = threadIdx.x & 63;
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const int64_t lw = __shfl(word, leader, 64);
        const bool same = changed && word == lw;
        uint32_t b = same ? bit : 0u;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) b |= __shfl_xor(b, off, 64);
        if (lane == leader) atomicOr(&mark[lw], b);
        pending &= ~__ballot(same);
    }
}

// Codes -> FP32 table.
template <int STORE>
__global__ __launch_bounds__(256) void decode_kernel(const uint16_t* __restrict__ codes, Consts K,
                                                     QuantGeom Q, const float2* __restrict__ tp,
                                                     int64_t off, int64_t n, float* __restrict__ V32) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int64_t idx = off + j;
    int64_t r = idx;
    int i[6];
    for (int a = 5; a >= 0; --a) {
        i[a] = (int)(r % K.n[a]);
        r /= K.n[a];
    }
    V32[j] = fastk::Codec<STORE>::decode(codes[idx], tp[fastk::tile_of_coords(Q, i)], K.cap);
}

}  // namespace pgk

namespace {

template <class W, bool DIAG, int STORE>
hipError_t launch_fast_w(const pgk::Consts& K, const pgk::FastConsts& F, const pgk::BoxGeom& G,
                         const pgk::QuantGeom& Q, const void* Vin, const float2* tp, const float* vself, float* Vout,
                         uint16_t* codes_out,
                         const uint8_t* occ, const float* bu, const int16_t* worig, const uint8_t* pfree, uint8_t* pol,
                         double* part, unsigned* linf, unsigned* tstat, int write_v, const pgk::ActiveG& A, hipStream_t st) {
    using RD = pgk::fastk::VRead<STORE>;
    using IO = pgk::fastk::SelfGlobal;
    RD rd;
    if constexpr (STORE == PG_STORE_F32) {
        rd.V = (const float*)Vin;
    } else {
        rd.V = (const uint16_t*)Vin;
        rd.tp = tp;
    }
    const IO io{vself, Vout};
    return pgk::for_box_chunks(G, [&](const pgk::BoxGeom& g, unsigned nb) {
        if (write_v && pol)
            hipLaunchKernelGGL((pgk::sweep_fast_kernel<W, DIAG, STORE, 128, true, true, RD, IO>), dim3(nb), dim3(256),
                               0, st, rd, io, codes_out, occ, bu, K, F, g, Q, worig, pfree, pol, part, linf, tstat, A);
        else if (write_v)
            hipLaunchKernelGGL((pgk::sweep_fast_kernel<W, DIAG, STORE, 128, true, false, RD, IO>), dim3(nb), dim3(256),
                               0, st, rd, io, codes_out, occ, bu, K, F, g, Q, worig, pfree, pol, part, linf, tstat, A);
        else
            hipLaunchKernelGGL((pgk::sweep_fast_kernel<W, DIAG, STORE, 128, false, true, RD, IO>), dim3(nb), dim3(256),
                               0, st, rd, io, codes_out, occ, bu, K, F, g, Q, worig, pfree, pol, part, linf, tstat, A);
    });
}

template <bool DIAG, int STORE>
hipError_t launch_fast_shape(int shape, const pgk::Consts& K, const pgk::FastConsts& F,
                             const pgk::BoxGeom& G, const pgk::QuantGeom& Q, const void* Vin,
                             const float2* tp, const float* vself, float* Vout, uint16_t* codes_out,
                             const uint8_t* occ, const float* bu,
                             const int16_t* worig, const uint8_t* pfree, uint8_t* pol, double* part, unsigned* linf,
                             unsigned* tstat, int write_v, const pgk::ActiveG& A, hipStream_t st) {
    using W0 = pgk::Win0;
    using W1 = pgk::Win1;
    using W2 = pgk::Win2;
    switch (shape) {
        case 0: return launch_fast_w<W0, DIAG, STORE>(K, F, G, Q, Vin, tp, vself, Vout, codes_out, occ, bu, worig, pfree, pol, part, linf, tstat, write_v, A, st);
        case 1: return launch_fast_w<W1, DIAG, STORE>(K, F, G, Q, Vin, tp, vself, Vout, codes_out, occ, bu, worig, pfree, pol, part, linf, tstat, write_v, A, st);
        case 2: return launch_fast_w<W2, DIAG, STORE>(K, F, G, Q, Vin, tp, vself, Vout, codes_out, occ, bu, worig, pfree, pol, part, linf, tstat, write_v, A, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t pgk_launch_sweep_fast(int shape, int diag, int store, const pgk::Consts& K,
                                 const pgk::FastConsts& F, const pgk::BoxGeom& G,
                                 const pgk::QuantGeom& Q, const void* Vin, const float2* tp,
                                 const float* vself, float* Vout, uint16_t* codes_out, const uint8_t* occ,
                                 const float* bu, const int16_t* worig, const uint8_t* pfree, uint8_t* pol, double* part, unsigned* linf,
                                 unsigned* tstat, int write_v, const pgk::ActiveG& A, hipStream_t st) {
    switch (store * 2 + (diag ? 1 : 0)) {
        case PG_STORE_F32 * 2 + 0: return launch_fast_shape<false, PG_STORE_F32>(shape, K, F, G, Q, Vin, tp, vself, Vout, codes_out, occ, bu, worig, pfree, pol, part, linf, tstat, write_v, A, st);
        case PG_STORE_F32 * 2 + 1: return launch_fast_shape<true, PG_STORE_F32>(shape, K, F, G, Q, Vin, tp, vself, Vout, codes_out, occ, bu, worig, pfree, pol, part, linf, tstat, write_v, A, st);
        case PG_STORE_F16 * 2 + 0: return launch_fast_shape<false, PG_STORE_F16>(shape, K, F, G, Q, Vin, tp, vself, Vout, codes_out, occ, bu, worig, pfree, pol, part, linf, tstat, write_v, A, st);
        case PG_STORE_F16 * 2 + 1: return launch_fast_shape<true, PG_STORE_F16>(shape, K, F, G, Q, Vin, tp, vself, Vout, codes_out, occ, bu, worig, pfree, pol, part, linf, tstat, write_v, A, st);
        case PG_STORE_U16 * 2 + 0: return launch_fast_shape<false, PG_STORE_U16>(shape, K, F, G, Q, Vin, tp, vself, Vout, codes_out, occ, bu, worig, pfree, pol, part, linf, tstat, write_v, A, st);
        case PG_STORE_U16 * 2 + 1: return launch_fast_shape<true, PG_STORE_U16>(shape, K, F, G, Q, Vin, tp, vself, Vout, codes_out, occ, bu, worig, pfree, pol, part, linf, tstat, write_v, A, st);
        default: return hipErrorInvalidValue;
    }
}

namespace {

template <class W, bool DIAG, int STORE>
hipError_t launch_fast_tile_w(const pgk::Consts& K, const pgk::FastConsts& F, const pgk::BoxGeom& G,
                              const pgk::QuantGeom& Q, const pgk::fastk::VReadTile<STORE>& rd,
                              const pgk::fastk::SelfTile& io, const uint8_t* occ, const float* bu,
                              const int16_t* worig, const uint8_t* pfree, double* part, unsigned* linf,
                              unsigned* tstat, hipStream_t st) {
    using RD = pgk::fastk::VReadTile<STORE>;
    using IO = pgk::fastk::SelfTile;
    return pgk::for_box_chunks(G, [&](const pgk::BoxGeom& g, unsigned nb) {
        hipLaunchKernelGGL((pgk::sweep_fast_kernel<W, DIAG, STORE, 128, true, false, RD, IO>), dim3(nb), dim3(256), 0,
                           st, rd, io, nullptr, occ, bu, K, F, g, Q, worig, pfree, nullptr, part, linf, tstat,
                           pgk::ActiveG{});
    });
}

template <bool DIAG, int STORE>
hipError_t launch_fast_tile(int shape, const pgk::Consts& K, const pgk::FastConsts& F, const pgk::BoxGeom& G,
                            const pgk::QuantGeom& Q, const uint16_t* win, const float2* tp, const int32_t wo[3],
                            const int32_t wl[3], unsigned* miss, const float* self, float* out, const uint8_t* occ,
                            const float* bu, const int16_t* worig, const uint8_t* pfree, double* part, unsigned* linf,
                            unsigned* tstat, hipStream_t st) {
    pgk::fastk::VReadTile<STORE> rd{};
    rd.W = win;
    rd.tp = tp;
    rd.miss = miss;
    const int32_t colv = K.n[3] * K.n[4] * K.n[5];
    for (int a = 0; a < 3; ++a) { rd.o[a] = wo[a]; rd.L[a] = wl[a]; }
    rd.ws[5] = 1;
    rd.ws[4] = K.n[5];
    rd.ws[3] = K.n[4] * K.n[5];
    rd.ws[2] = colv;
    rd.ws[1] = colv * wl[2];
    rd.ws[0] = colv * wl[2] * wl[1];
    pgk::fastk::SelfTile io{};
    io.self = self;
    io.out = out;
    for (int a = 0; a < 3; ++a) io.t0[a] = G.org[a];
    io.os[5] = 1;
    io.os[4] = K.n[5];
    io.os[3] = K.n[4] * K.n[5];
    io.os[2] = colv;
    io.os[1] = colv * G.rn[2];
    io.os[0] = colv * G.rn[2] * G.rn[1];
    switch (shape) {
        case 0: return launch_fast_tile_w<pgk::Win0, DIAG, STORE>(K, F, G, Q, rd, io, occ, bu, worig, pfree, part, linf, tstat, st);
        case 1: return launch_fast_tile_w<pgk::Win1, DIAG, STORE>(K, F, G, Q, rd, io, occ, bu, worig, pfree, part, linf, tstat, st);
        case 2: return launch_fast_tile_w<pgk::Win2, DIAG, STORE>(K, F, G, Q, rd, io, occ, bu, worig, pfree, part, linf, tstat, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t pgk_launch_sweep_fast_tile(int shape, int diag, int store, const pgk::Consts& K, const pgk::FastConsts& F,
                                      const pgk::BoxGeom& G, const pgk::QuantGeom& Q, const uint16_t* win,
                                      const float2* tp, const int32_t wo[3], const int32_t wl[3], unsigned* miss,
                                      const float* self, float* out, const uint8_t* occ, const float* bu,
                                      const int16_t* worig, const uint8_t* pfree, double* part, unsigned* linf,
                                      unsigned* tstat, hipStream_t st) {
    switch (store * 2 + (diag ? 1 : 0)) {
        case PG_STORE_F16 * 2 + 0: return launch_fast_tile<false, PG_STORE_F16>(shape, K, F, G, Q, win, tp, wo, wl, miss, self, out, occ, bu, worig, pfree, part, linf, tstat, st);
        case PG_STORE_F16 * 2 + 1: return launch_fast_tile<true, PG_STORE_F16>(shape, K, F, G, Q, win, tp, wo, wl, miss, self, out, occ, bu, worig, pfree, part, linf, tstat, st);
        case PG_STORE_U16 * 2 + 0: return launch_fast_tile<false, PG_STORE_U16>(shape, K, F, G, Q, win, tp, wo, wl, miss, self, out, occ, bu, worig, pfree, part, linf, tstat, st);
        case PG_STORE_U16 * 2 + 1: return launch_fast_tile<true, PG_STORE_U16>(shape, K, F, G, Q, win, tp, wo, wl, miss, self, out, occ, bu, worig, pfree, part, linf, tstat, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t pgk_launch_interior_stats(const float* V, const int32_t t0[3], const int32_t tn[3], const pgk::Consts& K,
                                     const pgk::QuantGeom& Q, unsigned* tstat, hipStream_t st) {
    pgk::Interior I{};
    for (int a = 0; a < 3; ++a) { I.t0[a] = t0[a]; I.tn[a] = tn[a]; }
    const int64_t n = (int64_t)tn[0] * tn[1] * tn[2] * K.n[3] * K.n[4] * K.n[5];
    hipLaunchKernelGGL(pgk::interior_stats_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, V, n, I, K, Q,
                       tstat);
    return hipGetLastError();
}

hipError_t pgk_launch_interior_refit(int store, const int32_t t0[3], const int32_t tn[3], const pgk::QuantGeom& Q,
                                     unsigned* tstat, const float2* tpo, float2* tpn, uint8_t* trefit, int force,
                                     hipStream_t st) {
    pgk::Interior I{};
    int64_t n = 1;
    for (int a = 0; a < 3; ++a) {
        I.t0[a] = t0[a];
        I.tn[a] = tn[a];
        n *= ((t0[a] + tn[a] - 1) >> pgk::kQuantShift) - (t0[a] >> pgk::kQuantShift) + 1;
    }
    for (int a = 3; a < 6; ++a) n *= Q.nt[a];
    const float ncode = store == PG_STORE_U16 ? pgk::fastk::kNcodeU16 : pgk::fastk::kNcodeF16;
    hipLaunchKernelGGL(pgk::interior_refit_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, I, Q, tstat,
                       tpo, tpn, trefit, ncode, force);
    return hipGetLastError();
}

hipError_t pgk_launch_interior_encode(int store, const float* V, const int32_t t0[3], const int32_t tn[3],
                                      const pgk::Consts& K, const pgk::QuantGeom& Q, const float2* tp,
                                      uint16_t* codes, hipStream_t st) {
    pgk::Interior I{};
    for (int a = 0; a < 3; ++a) { I.t0[a] = t0[a]; I.tn[a] = tn[a]; }
    const int64_t n = (int64_t)tn[0] * tn[1] * tn[2] * K.n[3] * K.n[4] * K.n[5];
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (store == PG_STORE_U16)
        hipLaunchKernelGGL(pgk::interior_encode_kernel<PG_STORE_U16>, dim3(nb), dim3(256), 0, st, V, n, I, K, Q, tp,
                           codes);
    else
        hipLaunchKernelGGL(pgk::interior_encode_kernel<PG_STORE_F16>, dim3(nb), dim3(256), 0, st, V, n, I, K, Q, tp,
                           codes);
    return hipGetLastError();
}

hipError_t pgk_launch_tile_stats(const float* V, const pgk::Consts& K, const pgk::QuantGeom& Q,
                                 unsigned* tstat, hipStream_t st) {
    const int64_t nb = (K.ncell + 255) / 256;
    hipLaunchKernelGGL(pgk::tile_stats_kernel, dim3((unsigned)nb), dim3(256), 0, st, V, K, Q, tstat);
    return hipGetLastError();
}

hipError_t pgk_launch_seed_stats(const int64_t* seeds, int64_t nseeds, const pgk::Consts& K,
                                 const pgk::QuantGeom& Q, unsigned* tstat, hipStream_t st) {
    if (nseeds <= 0) return hipSuccess;
    const int64_t nb = (nseeds + 255) / 256;
    hipLaunchKernelGGL(pgk::seed_stats_kernel, dim3((unsigned)nb), dim3(256), 0, st, seeds, nseeds, K, Q, tstat);
    return hipGetLastError();
}

hipError_t pgk_launch_tile_refit(int store, unsigned* tstat, float2* tp, uint8_t* trefit, int64_t ntile,
                                 int force, hipStream_t st) {
    const int64_t nb = (ntile + 255) / 256;
    const float ncode = store == PG_STORE_U16 ? pgk::fastk::kNcodeU16 : pgk::fastk::kNcodeF16;
    hipLaunchKernelGGL(pgk::tile_refit_kernel, dim3((unsigned)nb), dim3(256), 0, st, tstat, tp, trefit,
                       ntile, ncode, force);
    return hipGetLastError();
}

hipError_t pgk_launch_encode(int store, const float* V32, const pgk::Consts& K, const pgk::QuantGeom& Q,
                             const float2* tp, const uint8_t* trefit, uint16_t* codes, uint32_t* mark,
                             const pgk::BoxGeom& G, hipStream_t st) {
    const int64_t nb = (K.ncell + 255) / 256;
    if (store == PG_STORE_U16)
        hipLaunchKernelGGL(pgk::encode_kernel<PG_STORE_U16>, dim3((unsigned)nb), dim3(256), 0, st, V32, K, Q, tp,
                           trefit, codes, mark, G);
    else
        hipLaunchKernelGGL(pgk::encode_kernel<PG_STORE_F16>, dim3((unsigned)nb), dim3(256), 0, st, V32, K, Q, tp,
                           trefit, codes, mark, G);
    return hipGetLastError();
}

hipError_t pgk_launch_decode(int store, const uint16_t* codes, const pgk::Consts& K, const pgk::QuantGeom& Q,
                             const float2* tp, int64_t off, int64_t n, float* V32, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int64_t nb = (n + 255) / 256;
    if (store == PG_STORE_U16)
        hipLaunchKernelGGL(pgk::decode_kernel<PG_STORE_U16>, dim3((unsigned)nb), dim3(256), 0, st, codes, K, Q, tp,
                           off, n, V32);
    else
        hipLaunchKernelGGL(pgk::decode_kernel<PG_STORE_F16>, dim3((unsigned)nb), dim3(256), 0, st, codes, K, Q, tp,
                           off, n, V32);
    return hipGetLastError();
}
