// Adjoint of L = mean over start states of the multilinear V at the start,
// with respect to the 30 drivetrain parameters theta = B (9, row-major),
// A (9), q (3), diag S (3), diag D (3), c (3), at an FP64 fixed point V of
// the plain backup (no counterpart in the reference, whose only derivative
// is the Kuhn interpolant's gradient in the state, PeregrineSolver.jl:533-570).
//
// With the argmin control held fixed every reached state satisfies
// V_i = tau + P_i . V (P_i: the 64 corner weights of its successor); targets,
// blocked and capped states have no row. Differentiating,
//   dL/dtheta = sum_i lambda_i d(P_i . V)/dtheta,   (I - P)^T lambda = g,
// with g the starts' mean corner weights. d(P_i . V)/dtheta is the gradient
// of the interpolant at the successor times the successor's tangent, carried
// by forward-mode differentiation through the Euler substeps. lambda is
// iterated as lambda_j <- (g_j + sum_{i != j} P_ij lambda_i) / (1 - P_jj):
// a scatter of P^T lambda with FP64 atomics, then the diagonal-solved update.
// Everything is FP64.
#include "pg_f64.hpp"
#include "pg_kernels.h"

namespace pgk {

namespace adjk {

constexpr int kThreads = 256;

__device__ __forceinline__ void unflat(const Consts64& K, int64_t idx, int i[6]) {
    int64_t r = idx;
#pragma unroll
    for (int k = 5; k >= 1; --k) {
        i[k] = (int)(r % K.n[k]);
        r /= K.n[k];
    }
    i[0] = (int)r;
}

// The 64 corners of fractional coordinates f: flat index and weight.
template <class Fn>
__device__ __forceinline__ void corners(const Consts64& K, const double t[6], const int b[6], Fn fn) {
    int64_t o0[6], o1[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        int j0, j1;
        if (k == 2) {
            j0 = b[k] >= K.n[2] ? b[k] - K.n[2] : b[k];
            j1 = b[k] + 1;
            j1 = j1 >= K.n[2] ? j1 - K.n[2] : j1;
        } else {
            j0 = b[k];
            j1 = b[k] + 1 < K.n[k] - 1 ? b[k] + 1 : K.n[k] - 1;
        }
        o0[k] = (int64_t)j0 * K.stride[k];
        o1[k] = (int64_t)j1 * K.stride[k];
    }
    for (int d = 0; d < 64; ++d) {
        int64_t a = 0;
        double w = 1.0;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int c = (d >> k) & 1;
            a += c ? o1[k] : o0[k];
            w *= c ? t[k] : 1.0 - t[k];
        }
        fn(a, w);
    }
}

__device__ __forceinline__ void split(const double f[6], double t[6], int b[6]) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const double fl = floor(f[k]);
        t[k] = f[k] - fl;
        b[k] = (int)fl;
    }
}

struct Occ64 {
    const uint8_t* __restrict__ occ;
    int32_t n2, n3;
    __device__ __forceinline__ bool operator()(int ci, int cj, int ck) const {
        return occ[((int64_t)ci * n2 + cj) * n3 + ck] != 0;
    }
};

// d a_r / d theta_j at a fixed state.
__device__ __forceinline__ double accel_param_partial(const Consts64& K, int j, int r, const double u[3],
                                                      const double v[3]) {
    if (j < 9) return (j / 3 == r) ? u[j % 3] : 0.0;
    if (j < 18) return ((j - 9) / 3 == r) ? v[(j - 9) % 3] : 0.0;
    if (j < 21) return (j - 18 == r) ? v[2] * v[2] : 0.0;
    if (j < 24) return (j - 21 == r) ? clamp1_64(v[r] / K.eps[r]) : 0.0;
    if (j < 27) return (j - 24 == r) ? fabs(v[r]) * v[r] : 0.0;
    return (j - 27 == r) ? 1.0 : 0.0;
}

// Successor state (body frame) and its tangent with respect to theta_j.
__device__ __forceinline__ void euler_tangent(const Consts64& K, const Start64& c, const double bu[3],
                                              const double u[3], int j, double s[6], double ds[6]) {
    s[0] = c.x; s[1] = c.y; s[2] = c.h; s[3] = c.vx; s[4] = c.vy; s[5] = c.w;
#pragma unroll
    for (int q = 0; q < 6; ++q) ds[q] = 0.0;
    for (int k = 0; k < K.nsub; ++k) {
        const double h = s[2], vx = s[3], vy = s[4], w = s[5];
        const double dh = ds[2], dvx = ds[3], dvy = ds[4], dw = ds[5];
        double ax, ay, al;
        accel64(K, bu[0], bu[1], bu[2], vx, vy, w, ax, ay, al);
        const double v[3] = {vx, vy, w}, dv[3] = {dvx, dvy, dw};
        double dcs[3], dqd[3];
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
            dcs[cc] = fabs(v[cc] / K.eps[cc]) < 1.0 ? dv[cc] / K.eps[cc] : 0.0;
            dqd[cc] = 2.0 * fabs(v[cc]) * dv[cc];
        }
        double da[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            double t = 0.0;
#pragma unroll
            for (int cc = 0; cc < 3; ++cc)
                t += K.A[3 * r + cc] * dv[cc] + K.S[3 * r + cc] * dcs[cc] + K.D[3 * r + cc] * dqd[cc];
            t += 2.0 * K.q[r] * w * dw;
            da[r] = t + accel_param_partial(K, j, r, u, v);
        }
        double sh, ch;
        sincos(h, &sh, &ch);
        const double a1 = dvx - vy * dh, a2 = dvy + vx * dh;
        const double F[6] = {ch * vx - sh * vy, sh * vx + ch * vy, w, ax + w * vy, ay - w * vx, al};
        const double dF[6] = {ch * a1 - sh * a2, sh * a1 + ch * a2, dw,
                              da[0] + dw * vy + w * dvy, da[1] - dw * vx - w * dvx, da[2]};
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            s[q] = s[q] + F[q] * K.hs;
            ds[q] = ds[q] + dF[q] * K.hs;
        }
    }
}

}  // namespace adjk

// One row per state: the successor's split coordinates under the FP64
// argmin control, and the weight of the state's own cell; dead rows have
// tcoord[0] = -1.
__global__ __launch_bounds__(256) void adj_linearize_kernel(const double* __restrict__ V, const uint8_t* __restrict__ pol,
                                                            const uint32_t* __restrict__ seedmask,
                                                            const uint8_t* __restrict__ occ, const double* __restrict__ bu,
                                                            Consts64 K, double* __restrict__ tcoord,
                                                            int16_t* __restrict__ bcoord, double* __restrict__ self) {
    using namespace adjk;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= K.ncell) return;
    tcoord[idx * 6] = -1.0;
    self[idx] = 0.0;
    int i[6];
    unflat(K, idx, i);
    const bool seed = (seedmask[idx >> 5] >> (idx & 31)) & 1u;
    if (seed || occ[((int64_t)i[0] * K.n[1] + i[1]) * K.n[2] + i[2]] || !(V[idx] < K.cap)) return;
    Start64 c;
    cell_start64(K, i, c);
    const int k = pol[idx];
    double f[6];
    const Occ64 og{occ, K.n[1], K.n[2]};
    if (!successor64(K, og, c, bu[3 * k], bu[3 * k + 1], bu[3 * k + 2], f)) return;
    double t[6];
    int b[6];
    split(f, t, b);
    double acc = 0.0;
    corners(K, t, b, [&](int64_t a, double w) { acc = acc + w * V[a]; });
    if (!(acc < K.cap) || !(K.tau + acc < K.cap)) return;
    double w = 1.0;
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        int d = i[a] - b[a];
        if (a == 2) d = d < 0 ? d + K.n[2] : d;
        w *= d == 0 ? 1.0 - t[a] : (d == 1 ? t[a] : 0.0);
        tcoord[idx * 6 + a] = t[a];
        bcoord[idx * 6 + a] = (int16_t)b[a];
    }
    self[idx] = w;
}

// g: each start's corner weights / nstarts; vstart: V at each start.
__global__ void adj_starts_kernel(const double* __restrict__ V, const double* __restrict__ starts, int64_t n,
                                  Consts64 K, double* __restrict__ g, double* __restrict__ vstart) {
    using namespace adjk;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double f[6];
#pragma unroll
    for (int a = 0; a < 6; ++a) f[a] = (starts[6 * k + a] - K.lo[a]) / K.step[a];
    f[2] = f[2] - floor(f[2] / K.n3) * K.n3;
    bool in = true;
#pragma unroll
    for (int a = 0; a < 6; ++a) in = in & (f[a] >= 0.0) & (f[a] <= K.lim[a]);
    if (!in) { vstart[k] = K.cap; return; }
    double t[6];
    int b[6];
    split(f, t, b);
    double acc = 0.0;
    corners(K, t, b, [&](int64_t a, double w) { acc = acc + w * V[a]; });
    vstart[k] = acc < K.cap ? acc : K.cap;
    if (!(acc < K.cap)) return;
    const double inv = 1.0 / (double)n;
    corners(K, t, b, [&](int64_t a, double w) { if (w != 0.0) unsafeAtomicAdd(&g[a], w * inv); });
}

// nxt += P^T lam (scatter over live rows).
__global__ __launch_bounds__(256) void adj_scatter_kernel(const double* __restrict__ lam, const double* __restrict__ tcoord,
                                                          const int16_t* __restrict__ bcoord, Consts64 K,
                                                          double* __restrict__ nxt) {
    using namespace adjk;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= K.ncell) return;
    const double li = lam[idx];
    if (li == 0.0 || tcoord[idx * 6] < 0.0) return;
    double t[6];
    int b[6];
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        t[a] = tcoord[idx * 6 + a];
        b[a] = bcoord[idx * 6 + a];
    }
    corners(K, t, b, [&](int64_t a, double w) { if (w != 0.0) unsafeAtomicAdd(&nxt[a], w * li); });
}

// lam <- (g + nxt - self lam) / (1 - self); nxt <- 0; max |change| and max
// |lam| into stat[0], stat[1] (as ordered bits of non-negative doubles).
__global__ __launch_bounds__(256) void adj_update_kernel(double* __restrict__ lam, double* __restrict__ nxt,
                                                         const double* __restrict__ g, const double* __restrict__ self,
                                                         int64_t n, unsigned long long* __restrict__ stat) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double dch = 0.0, mag = 0.0;
    if (j < n) {
        const double sj = self[j];
        const double v = (g[j] + nxt[j] - sj * lam[j]) / (1.0 - sj);
        dch = fabs(v - lam[j]);
        mag = fabs(v);
        lam[j] = v;
        nxt[j] = 0.0;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        dch = fmax(dch, __shfl_xor(dch, off, 64));
        mag = fmax(mag, __shfl_xor(mag, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&stat[0], (unsigned long long)__double_as_longlong(dch));
        atomicMax(&stat[1], (unsigned long long)__double_as_longlong(mag));
    }
}

// Per block: sum over its live states of lam_i * d(P_i . V)/dtheta_j, j < 30.
__global__ __launch_bounds__(256) void adj_grad_kernel(const double* __restrict__ V, const double* __restrict__ lam,
                                                       const uint8_t* __restrict__ pol, const double* __restrict__ tcoord,
                                                       const uint8_t* __restrict__ occ, const double* __restrict__ bu,
                                                       const double* __restrict__ usat, Consts64 K,
                                                       double* __restrict__ part) {
    using namespace adjk;
    __shared__ double red[30][kThreads / 64];
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = idx < K.ncell && lam[idx] != 0.0 && tcoord[idx * 6] >= 0.0;
    Start64 c{};
    double bk[3] = {0, 0, 0}, u[3] = {0, 0, 0}, gi[6] = {0, 0, 0, 0, 0, 0};
    double li = 0.0;
    if (live) {
        int i[6];
        unflat(K, idx, i);
        cell_start64(K, i, c);
        const int k = pol[idx];
        bk[0] = bu[3 * k]; bk[1] = bu[3 * k + 1]; bk[2] = bu[3 * k + 2];
        u[0] = usat[3 * k]; u[1] = usat[3 * k + 1]; u[2] = usat[3 * k + 2];
        // the interpolant's gradient at the successor
        double s[6], ds[6];
        euler_tangent(K, c, bk, u, 0, s, ds);
        double snh, cnh;
        sincos(s[2], &snh, &cnh);
        double f[6] = {s[0], s[1], s[2], cnh * s[3] - snh * s[4], snh * s[3] + cnh * s[4], s[5]};
#pragma unroll
        for (int a = 0; a < 6; ++a) f[a] = (f[a] - K.lo[a]) / K.step[a];
        f[2] = f[2] - floor(f[2] / K.n3) * K.n3;
        double t[6];
        int b[6];
        split(f, t, b);
        int d = 0;
        corners(K, t, b, [&](int64_t a, double) {
            const double v = V[a];
#pragma unroll
            for (int ax = 0; ax < 6; ++ax) {
                double wa = ((d >> ax) & 1) ? 1.0 : -1.0;
#pragma unroll
                for (int q = 0; q < 6; ++q)
                    if (q != ax) wa *= ((d >> q) & 1) ? t[q] : 1.0 - t[q];
                gi[ax] += wa * v;
            }
            ++d;
        });
        li = lam[idx];
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int j = 0; j < 30; ++j) {
        double r = 0.0;
        if (live) {
            double s[6], ds[6];
            euler_tangent(K, c, bk, u, j, s, ds);
            double snh, cnh;
            sincos(s[2], &snh, &cnh);
            const double a1 = ds[3] - s[4] * ds[2], a2 = ds[4] + s[3] * ds[2];
            const double dsv[6] = {ds[0], ds[1], ds[2], cnh * a1 - snh * a2, snh * a1 + cnh * a2, ds[5]};
#pragma unroll
            for (int a = 0; a < 6; ++a) r += gi[a] * dsv[a] / K.step[a];
            r *= li;
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) r += __shfl_xor(r, off, 64);
        if (lane == 0) red[j][wv] = r;
    }
    __syncthreads();
    if (threadIdx.x < 30) {
        double v = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) v += red[threadIdx.x][w];
        part[(int64_t)blockIdx.x * 30 + threadIdx.x] = v;
    }
}

}  // namespace pgk

hipError_t pgk_launch_adj_linearize(const double* V, const uint8_t* pol, const uint32_t* seedmask, const uint8_t* occ,
                                    const double* bu, const pgk::Consts64& K, double* tcoord, int16_t* bcoord,
                                    double* self, hipStream_t st) {
    const int64_t nb = (K.ncell + 255) / 256;
    hipLaunchKernelGGL(pgk::adj_linearize_kernel, dim3((unsigned)nb), dim3(256), 0, st, V, pol, seedmask, occ, bu, K,
                       tcoord, bcoord, self);
    return hipGetLastError();
}

hipError_t pgk_launch_adj_starts(const double* V, const double* starts, int64_t n, const pgk::Consts64& K, double* g,
                                 double* vstart, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int64_t nb = (n + 255) / 256;
    hipLaunchKernelGGL(pgk::adj_starts_kernel, dim3((unsigned)nb), dim3(256), 0, st, V, starts, n, K, g, vstart);
    return hipGetLastError();
}

hipError_t pgk_launch_adj_iteration(double* lam, double* nxt, const double* g, const double* self, const double* tcoord,
                                    const int16_t* bcoord, const pgk::Consts64& K, unsigned long long* stat,
                                    hipStream_t st) {
    const int64_t nb = (K.ncell + 255) / 256;
    hipLaunchKernelGGL(pgk::adj_scatter_kernel, dim3((unsigned)nb), dim3(256), 0, st, lam, tcoord, bcoord, K, nxt);
    hipLaunchKernelGGL(pgk::adj_update_kernel, dim3((unsigned)nb), dim3(256), 0, st, lam, nxt, g, self, K.ncell, stat);
    return hipGetLastError();
}

int64_t pgk_adj_grad_blocks(const pgk::Consts64& K) { return (K.ncell + 255) / 256; }

hipError_t pgk_launch_adj_grad(const double* V, const double* lam, const uint8_t* pol, const double* tcoord,
                               const uint8_t* occ, const double* bu, const double* usat, const pgk::Consts64& K,
                               double* part, hipStream_t st) {
    const int64_t nb = pgk_adj_grad_blocks(K);
    hipLaunchKernelGGL(pgk::adj_grad_kernel, dim3((unsigned)nb), dim3(256), 0, st, V, lam, pol, tcoord, occ, bu, usat,
                       K, part);
    return hipGetLastError();
}
