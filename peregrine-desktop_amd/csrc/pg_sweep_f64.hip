// FP64 Bellman sweeps over LDS windows: the FP64 tables the precision form
// is measured against, and FP64 residual checks of any table.
//
// Same box / window / (state, control)-pair organisation as the FP32 LDS
// sweep (the window origins are the FP32 plan's; a successor that FP64
// rounding moves out of its window reads global memory), with double values
// in LDS and pg_f64.hpp's arithmetic. Three modes:
//   kPlain    V' = min(V, T V)                                (Jacobi)
//   kDiagonal V' = min(V, C V), C solving each control's own-cell term:
//             tau + w V_s + rest = V_s  =>  V_s = (tau + rest) / (1 - w);
//             the fixed point is T's, reached in fewer sweeps
//   kResidual no write; |T V - V| over states where either is below cap,
//             target states (held at 0, bit set in `seedmask`) excepted
// The per-state minimum over controls is an atomicMin on the value's bits
// (non-negative doubles order as their bit patterns); the argmin is an
// atomicMin on (value bits with the low 8 bits cleared, control index), so
// controls within 2^-44 relative of the minimum tie to the lowest index.
#include "pg_box.hpp"
#include "pg_f64.hpp"
#include "pg_kernels.h"

namespace pgk {

namespace f64k {

constexpr int kMaxCtl = 128;

template <int MAXS>
struct BoxShared {
    double x[MAXS], y[MAXS], h[MAXS], vx[MAXS], vy[MAXS], w[MAXS];
    double vold[MAXS];
    unsigned long long best[MAXS];   // value bits
    unsigned long long key[MAXS];    // (value bits >> 8) << 8 | control
    uint16_t loc[MAXS];              // coordinates inside the box, 3 bits per axis
    uint8_t blocked[MAXS];
    double bu[3 * kMaxCtl];
};

__device__ __forceinline__ unsigned long long dbits(double v) {
    return (unsigned long long)__double_as_longlong(v);
}

// Corner values of the successor cell i[] from global memory, with the
// reference's clamped / periodic indices.
__device__ __forceinline__ double corner_global64(const Consts64& K, const double* __restrict__ V,
                                                  const int i[6], int d) {
    int64_t a = 0;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        int j = i[k] + ((d >> k) & 1);
        if (k == 2) {
            j = j >= K.n[2] ? j - K.n[2] : j;
            j = j >= K.n[2] ? j - K.n[2] : j;
        } else {
            j = j < K.n[k] - 1 ? j : K.n[k] - 1;
        }
        a += (int64_t)j * K.stride[k];
    }
    return V[a];
}

}  // namespace f64k

template <class W, int MODE, int MAXS, bool WRITE_POL>
__global__ __launch_bounds__(256) void sweep_f64_kernel(
    const double* __restrict__ Vin, double* __restrict__ Vout, const uint8_t* __restrict__ occ,
    const double* __restrict__ bu_g, Consts64 K, BoxGeom G, const int16_t* __restrict__ worig,
    const uint32_t* __restrict__ seedmask, uint8_t* __restrict__ pol, double* __restrict__ part,
    unsigned* __restrict__ linf_bits) {
    using namespace f64k;
    __shared__ double win[W::SIZE];
    __shared__ BoxShared<MAXS> sm;
    const int64_t box = box_of_block(G);
    const bool live = box < G.nbox;
    BoxInfo B;
    int o[6];
    if (live) {
        box_info(G, box, B);
#pragma unroll
        for (int a = 0; a < 6; ++a) o[a] = worig[B.gbox * 6 + a];
        for (int j = threadIdx.x; j < W::CELLS; j += kThreads) {
            int l[6], gi[6];
            const bool on = window_cell<W>(K.n, o, j, l, gi);
            double v = K.cap;
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
        for (int s = threadIdx.x; s < B.ns; s += kThreads) {
            int i[6];
            state_coords(B, s, i);
            Start64 c;
            cell_start64(K, i, c);
            sm.x[s] = c.x; sm.y[s] = c.y; sm.h[s] = c.h; sm.vx[s] = c.vx; sm.vy[s] = c.vy; sm.w[s] = c.w;
            int64_t fl = 0;
            unsigned pk = 0;
#pragma unroll
            for (int a = 0; a < 6; ++a) {
                fl += (int64_t)i[a] * K.stride[a];
                pk |= (unsigned)(i[a] - B.s0[a]) << (3 * a);
            }
            sm.vold[s] = Vin[fl];
            sm.loc[s] = (uint16_t)pk;
            const bool blk = occ[((int64_t)i[0] * K.n[1] + i[1]) * K.n[2] + i[2]] != 0;
            sm.blocked[s] = blk ? 1 : 0;
            // blocked cells keep (cap, coast) (PeregrineSolver.jl:1226-1228)
            sm.best[s] = blk ? dbits(K.cap) : dbits(__builtin_inf());
            sm.key[s] = blk ? ((dbits(K.cap) >> 8) << 8) : ((dbits(__builtin_inf()) >> 8) << 8);
        }
        for (int j = threadIdx.x; j < 3 * K.nctl; j += kThreads) sm.bu[j] = bu_g[j];
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
                Start64 c;
                c.x = sm.x[s]; c.y = sm.y[s]; c.h = sm.h[s]; c.vx = sm.vx[s]; c.vy = sm.vy[s]; c.w = sm.w[s];
                double f[6];
                double val = K.cap;
                if (successor64(K, og, c, sm.bu[3 * k], sm.bu[3 * k + 1], sm.bu[3 * k + 2], f)) {
                    double t[6];
                    int i[6], l[6];
                    bool inwin = true;
#pragma unroll
                    for (int a = 0; a < 6; ++a) {
                        const double b = floor(f[a]);
                        t[a] = f[a] - b;
                        i[a] = (int)b;
                        int d = i[a] - o[a];
                        if (a == 2) {
                            d = d < 0 ? d + K.n[2] : d;
                            d = d >= K.n[2] ? d - K.n[2] : d;
                        }
                        l[a] = d;
                        inwin = inwin & ((unsigned)d <= (unsigned)(W::E[a] - 2));
                    }
                    double wl[8], wh[8];
                    corner_weights64(t, wl, wh);
                    // every corner value is finite (<= cap), so a zero weight
                    // adds +0 and the reference's skip needs no branch
                    double acc = 0.0;
                    if (inwin) {
                        int base = 0;
#pragma unroll
                        for (int a = 0; a < 6; ++a) base += l[a] * W::S[a];
#pragma unroll
                        for (int d = 0; d < 64; ++d) {
                            const int off = ((d >> 0) & 1) * W::S1 + ((d >> 1) & 1) * W::S2 +
                                            ((d >> 2) & 1) * W::S3 + ((d >> 3) & 1) * W::S4 +
                                            ((d >> 4) & 1) * W::S5 + ((d >> 5) & 1) * W::S6;
                            acc = acc + wl[d & 7] * wh[d >> 3] * win[base + off];
                        }
                    } else {
                        for (int d = 0; d < 64; ++d)
                            acc = acc + wl[d & 7] * wh[d >> 3] * corner_global64(K, Vin, i, d);
                    }
                    double wself = 0.0;
                    if (MODE == PG_F64_DIAGONAL) {
                        const unsigned pk = sm.loc[s];
                        wself = 1.0;
#pragma unroll
                        for (int a = 0; a < 6; ++a) {
                            int d = B.s0[a] + (int)((pk >> (3 * a)) & 7u) - i[a];
                            if (a == 2) d = d < 0 ? d + K.n[2] : d;
                            wself *= d == 0 ? 1.0 - t[a] : (d == 1 ? t[a] : 0.0);
                        }
                    }
                    if (MODE == PG_F64_DIAGONAL && wself > 0.0) {
                        const double cc = (K.tau + (acc - wself * sm.vold[s])) / (1.0 - wself);
                        val = cc < K.cap ? cc : K.cap;
                    } else {
                        const double nv = acc < K.cap ? acc : K.cap;
                        const double cc = K.tau + nv;
                        val = cc < K.cap ? cc : K.cap;
                    }
                }
                atomicMin(&sm.best[s], dbits(val));
                if (WRITE_POL) atomicMin(&sm.key[s], ((dbits(val) >> 8) << 8) | (unsigned)k);
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
            const double best = __longlong_as_double((long long)sm.best[s]);
            const double old = sm.vold[s];
            if (MODE == PG_F64_RESIDUAL) {
                const bool seed = (seedmask[idx >> 5] >> (idx & 31)) & 1u;
                if (!seed && (best < K.cap || old < K.cap)) {
                    const double d = fabs(best - old);
                    dsum += d;
                    dmax = fmaxf(dmax, (float)d);
                }
            } else if (Vout) {
                double out = old;
                if (best < old) {
                    out = best;
                    dsum += old - best;
                    dmax = fmaxf(dmax, (float)(old - best));
                }
                Vout[idx] = out;
            }
            if (WRITE_POL) pol[idx] = (uint8_t)(sm.key[s] & 0xffu);
        }
    }
    if (part) block_reduce_sum_max(dsum, dmax, part + G.blk0, linf_bits);
}

__global__ void fill_f64_kernel(double* __restrict__ V, int64_t n, double v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) V[i] = v;
}

__global__ void seed_f64_kernel(double* __restrict__ V, const int64_t* __restrict__ seeds, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) V[seeds[i]] = 0.0;
}

__global__ void seed_mask_kernel(uint32_t* __restrict__ mask, const int64_t* __restrict__ seeds, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicOr(&mask[seeds[i] >> 5], 1u << (seeds[i] & 31));
}

__global__ void widen_kernel(const float* __restrict__ V32, double* __restrict__ V64, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) V64[i] = (double)V32[i];
}

}  // namespace pgk

namespace {

template <class W, int MODE>
hipError_t launch_f64_w(const pgk::Consts64& K, const pgk::BoxGeom& G, const double* Vin, double* Vout,
                        const uint8_t* occ, const double* bu, const int16_t* worig, const uint32_t* seedmask,
                        uint8_t* pol, double* part, unsigned* linf, hipStream_t st) {
    return pgk::for_box_chunks(G, [&](const pgk::BoxGeom& g, unsigned nb) {
        if (pol)
            hipLaunchKernelGGL((pgk::sweep_f64_kernel<W, MODE, 128, true>), dim3(nb), dim3(256), 0, st, Vin, Vout,
                               occ, bu, K, g, worig, seedmask, pol, part, linf);
        else
            hipLaunchKernelGGL((pgk::sweep_f64_kernel<W, MODE, 128, false>), dim3(nb), dim3(256), 0, st, Vin, Vout,
                               occ, bu, K, g, worig, seedmask, pol, part, linf);
    });
}

template <int MODE>
hipError_t launch_f64_mode(int shape, const pgk::Consts64& K, const pgk::BoxGeom& G, const double* Vin,
                           double* Vout, const uint8_t* occ, const double* bu, const int16_t* worig,
                           const uint32_t* seedmask, uint8_t* pol, double* part, unsigned* linf,
                           hipStream_t st) {
    switch (shape) {
        case 0: return launch_f64_w<pgk::Win0, MODE>(K, G, Vin, Vout, occ, bu, worig, seedmask, pol, part, linf, st);
        case 1: return launch_f64_w<pgk::Win1, MODE>(K, G, Vin, Vout, occ, bu, worig, seedmask, pol, part, linf, st);
        case 2: return launch_f64_w<pgk::Win2, MODE>(K, G, Vin, Vout, occ, bu, worig, seedmask, pol, part, linf, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t pgk_launch_sweep_f64(int mode, int shape, const pgk::Consts64& K, const pgk::BoxGeom& G,
                                const double* Vin, double* Vout, const uint8_t* occ, const double* bu,
                                const int16_t* worig, const uint32_t* seedmask, uint8_t* pol, double* part,
                                unsigned* linf, hipStream_t st) {
    switch (mode) {
        case PG_F64_PLAIN:
            return launch_f64_mode<PG_F64_PLAIN>(shape, K, G, Vin, Vout, occ, bu, worig, seedmask, pol, part, linf, st);
        case PG_F64_DIAGONAL:
            return launch_f64_mode<PG_F64_DIAGONAL>(shape, K, G, Vin, Vout, occ, bu, worig, seedmask, pol, part, linf, st);
        case PG_F64_RESIDUAL:
            return launch_f64_mode<PG_F64_RESIDUAL>(shape, K, G, Vin, nullptr, occ, bu, worig, seedmask, pol, part, linf, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t pgk_launch_fill_f64(double* V, int64_t n, double v, hipStream_t st) {
    const int64_t nb = (n + 255) / 256;
    hipLaunchKernelGGL(pgk::fill_f64_kernel, dim3((unsigned)nb), dim3(256), 0, st, V, n, v);
    return hipGetLastError();
}

hipError_t pgk_launch_seed_f64(double* V, const int64_t* seeds, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int64_t nb = (n + 255) / 256;
    hipLaunchKernelGGL(pgk::seed_f64_kernel, dim3((unsigned)nb), dim3(256), 0, st, V, seeds, n);
    return hipGetLastError();
}

hipError_t pgk_launch_widen(const float* V32, double* V64, int64_t n, hipStream_t st) {
    const int64_t nb = (n + 255) / 256;
    hipLaunchKernelGGL(pgk::widen_kernel, dim3((unsigned)nb), dim3(256), 0, st, V32, V64, n);
    return hipGetLastError();
}

hipError_t pgk_launch_seed_mask(uint32_t* mask, const int64_t* seeds, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int64_t nb = (n + 255) / 256;
    hipLaunchKernelGGL(pgk::seed_mask_kernel, dim3((unsigned)nb), dim3(256), 0, st, mask, seeds, n);
    return hipGetLastError();
}
