// Boxes and windows of the LDS-staged sweeps.
//
// The grid is cut into boxes of states (b1 x ... x b6). One workgroup owns a
// box and stages into LDS the window of V that the box's successors read: a
// 6-D block of E1 x ... x E6 cells (omega fastest) whose origin comes from a
// one-time plan pass.
#pragma once
#include "pg_device.hpp"

namespace pgk {

constexpr int kThreads = 256;

// Padded LDS strides of an E1..E6 window, omega fastest. Each stride above
// omega's rows is rounded up to an odd number of elements so that
// neighbouring rows start in different banks; P2 and P1 add to the y and x
// strides. The corners a wave gathers at once belong to successors a cell
// or two apart on each axis, so two lanes collide in a bank when a small
// combination of strides is a multiple of the bank count: with x and y
// strides 1621 and 405 (both 21 modulo 64), a step of +1 in x against -1
// in y is one. 1631 and 407 avoid it: a model of the gathers of a
// 564M-state, 125-control grid counts a fifth fewer bank-conflict cycles,
// and the contracted sweeps run 2-3% faster (the exact sweep, bound by its
// arithmetic, does not move).
template <int E1, int E2, int E3, int E4, int E5, int E6, int P2 = 0, int P1 = 0>
struct Win {
    static constexpr int E[6] = {E1, E2, E3, E4, E5, E6};
    static constexpr int S6 = 1;
    static constexpr int S5 = E6 * S6;
    static constexpr int S4 = (E5 * S5) | 1;
    static constexpr int S3 = (E4 * S4) | 1;
    static constexpr int S2 = ((E3 * S3) | 1) + P2;
    static constexpr int S1 = ((E2 * S2) | 1) + P1;
    static constexpr int SIZE = E1 * S1;
    static constexpr int CELLS = E1 * E2 * E3 * E4 * E5 * E6;
    static constexpr int S[6] = {S1, S2, S3, S4, S5, S6};
};

using Win0 = Win<4, 4, 4, 4, 4, 6, 2, 2>;
using Win1 = Win<4, 4, 4, 4, 4, 4>;
using Win2 = Win<4, 4, 4, 5, 5, 5>;

// The box a workgroup owns. Workgroups are dealt to the 8 XCDs in turn
// (workgroup b runs on XCD b % 8), and work per box is uneven (obstacles,
// rejected successors, boxes an active set skips), so each XCD walks chunks
// of kBoxChunk consecutive boxes (neighbours, whose windows overlap, share
// its L2) taken from all over the grid: chunk q of XCD x is chunk 8q + x.
__device__ __forceinline__ int64_t box_of_block(const BoxGeom& G) {
    int64_t b = G.blk0 + (int64_t)blockIdx.x;
    // reversed order (in-place sweeps alternate): the last boxes first
    if (G.rev) b = box_blocks(G.nbox) - 1 - b;
    const int64_t j = b / 8;
    return ((j / kBoxChunk) * 8 + b % 8) * kBoxChunk + j % kBoxChunk;
}

// A dispatch counts its work-items in 32 bits, so a grid of boxes is
// launched in chunks of at most 2^23 workgroups (2^31 work-items); each
// chunk's geometry carries its first workgroup in blk0.
constexpr int64_t kMaxBoxBlocks = (int64_t)1 << 23;

template <class Launch>
inline hipError_t for_box_chunks(const BoxGeom& G, Launch launch) {
    const int64_t nb = box_blocks(G.nbox);
    for (int64_t b0 = 0; b0 < nb; b0 += kMaxBoxBlocks) {
        BoxGeom g = G;
        g.blk0 = b0;
        const int64_t n = nb - b0 < kMaxBoxBlocks ? nb - b0 : kMaxBoxBlocks;
        launch(g, (unsigned)n);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

struct BoxInfo {
    int32_t s0[6];   // first state of the box, per axis (grid coordinates)
    int32_t vb[6];   // states of the box per axis (edge boxes are partial)
    int32_t ns;
    int64_t gbox;    // the box's number over the whole grid
};

__device__ __forceinline__ void box_info(const BoxGeom& G, int64_t box, BoxInfo& B) {
    int64_t r = box;
    int32_t bi[6];
#pragma unroll
    for (int k = 5; k >= 1; --k) {
        bi[k] = (int32_t)(r % G.nb[k]);
        r /= G.nb[k];
    }
    bi[0] = (int32_t)r;
    B.ns = 1;
    B.gbox = 0;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        B.s0[k] = G.org[k] + bi[k] * G.b[k];
        const int32_t rem = G.org[k] + G.rn[k] - B.s0[k];
        B.vb[k] = rem < G.b[k] ? rem : G.b[k];
        B.ns *= B.vb[k];
        B.gbox = B.gbox * G.nbg[k] + (G.org[k] / G.b[k] + bi[k]);
    }
}

__device__ __forceinline__ void state_coords(const BoxInfo& B, int s, int i[6]) {
    int r = s;
#pragma unroll
    for (int k = 5; k >= 1; --k) {
        i[k] = B.s0[k] + r % B.vb[k];
        r /= B.vb[k];
    }
    i[0] = B.s0[0] + r;
}

// Window cell j (row-major over E1..E6) -> its local coordinates, its grid
// coordinates (heading periodic) and whether it lies on the grid.
template <class W>
__device__ __forceinline__ bool window_cell(const int32_t n[6], const int o[6], int j, int l[6], int gi[6]) {
    int r = j;
    l[5] = r % W::E[5]; r /= W::E[5];
    l[4] = r % W::E[4]; r /= W::E[4];
    l[3] = r % W::E[3]; r /= W::E[3];
    l[2] = r % W::E[2]; r /= W::E[2];
    l[1] = r % W::E[1]; r /= W::E[1];
    l[0] = r;
    bool on = true;
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        int gv = o[a] + l[a];
        if (a == 2) {
            while (gv >= n[2]) gv -= n[2];
        } else {
            on = on & (gv >= 0) & (gv < n[a]);
        }
        gi[a] = gv;
    }
    return on;
}

// The boxes of one axis that meet the cells [c, c + ext): `cnt` boxes from
// `first`, counted modulo nb on the (periodic) heading axis. Off-grid cells
// of the other axes are clamped, as the corner indices are.
__device__ __forceinline__ void box_span(bool periodic, int c, int ext, int n, int b, int nb, int& first,
                                         int& cnt) {
    if (ext >= 255) { first = 0; cnt = nb; return; }   // a saturated extent: the whole axis
    if (periodic) {
        const int last = c + ext - 1;
        if (ext >= n) { first = 0; cnt = nb; return; }
        first = c / b;
        if (last < n) { cnt = last / b - first + 1; return; }
        cnt = (nb - first) + (last - n) / b + 1;
        if (cnt >= nb) { first = 0; cnt = nb; }
        return;
    }
    const int lo = c < 0 ? 0 : (c > n - 1 ? n - 1 : c);
    int hi = c + ext - 1;
    hi = hi > n - 1 ? n - 1 : (hi < lo ? lo : hi);
    first = lo / b;
    cnt = hi / b - first + 1;
}

// Nonzero in at least one thread of the block iff a box of the box's read
// range, or the box itself, changed in the previous sweep (every thread
// calls it and the block ORs the results).
__device__ __forceinline__ int box_probe(const ActiveG& A, const BoxGeom& G, const int32_t n[6],
                                         const BoxInfo& B) {
    const int64_t gb = B.gbox;
    int first[6], cnt[6];
#pragma unroll
    for (int a = 0; a < 6; ++a)
        box_span(a == 2, A.rorig[gb * 6 + a], A.rext[gb * 6 + a], n[a], G.b[a], G.nbg[a], first[a], cnt[a]);
    const uint32_t mask = (cnt[5] >= 32 ? ~0u : ((1u << cnt[5]) - 1u)) << first[5];
    const int total = cnt[0] * cnt[1] * cnt[2] * cnt[3] * cnt[4];
    int hit = 0;
    for (int t = threadIdx.x; t < total; t += kThreads) {
        int r = t, bi[5];
#pragma unroll
        for (int a = 4; a >= 0; --a) {
            int v = first[a] + r % cnt[a];
            r /= cnt[a];
            if (a == 2 && v >= G.nbg[2]) v -= G.nbg[2];
            bi[a] = v;
        }
        int64_t w = 0;
#pragma unroll
        for (int a = 0; a < 5; ++a) w = w * G.nbg[a] + bi[a];
        hit |= (A.prev[w] & mask) != 0u;
    }
    if (threadIdx.x == 0) hit |= (int)((A.prev[gb / G.nbg[5]] >> (int)(gb % G.nbg[5])) & 1u);
    return hit;
}

// Whether the block computes its box this sweep (block-uniform; every
// thread of a live block must call it).
__device__ __forceinline__ bool box_active(const ActiveG& A, const BoxGeom& G, const int32_t n[6],
                                           const BoxInfo& B) {
    if (!A.prev) return true;
    return __syncthreads_or(box_probe(A, G, n, B)) != 0;
}

__device__ __forceinline__ void box_mark(uint32_t* bits, const BoxGeom& G, int64_t gbox) {
    atomicOr(&bits[gbox / G.nbg[5]], 1u << (int)(gbox % G.nbg[5]));
}

// The whole-grid box number of cell i.
__device__ __forceinline__ int64_t box_of_cell(const BoxGeom& G, const int i[6]) {
    int64_t g = 0;
#pragma unroll
    for (int a = 0; a < 6; ++a) g = g * G.nbg[a] + i[a] / G.b[a];
    return g;
}

// Occupancy of an (x, y, heading) cell. skip: the caller has shown that no
// cell it can ask about holds an obstacle (a box whose probes reach only
// free cells), so the lookup is not made.
struct OccG {
    const uint8_t* __restrict__ occ;
    int32_t n2, n3;
    bool skip = false;
    __device__ __forceinline__ bool operator()(int ci, int cj, int ck) const {
        if (skip) return false;
        return occ[((int64_t)ci * n2 + cj) * n3 + ck] != 0;
    }
};

}  // namespace pgk
