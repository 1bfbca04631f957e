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
// neighbouring rows start in different banks.
template <int E1, int E2, int E3, int E4, int E5, int E6>
struct Win {
    static constexpr int E[6] = {E1, E2, E3, E4, E5, E6};
    static constexpr int S6 = 1;
    static constexpr int S5 = E6 * S6;
    static constexpr int S4 = (E5 * S5) | 1;
    static constexpr int S3 = (E4 * S4) | 1;
    static constexpr int S2 = (E3 * S3) | 1;
    static constexpr int S1 = (E2 * S2) | 1;
    static constexpr int SIZE = E1 * S1;
    static constexpr int CELLS = E1 * E2 * E3 * E4 * E5 * E6;
    static constexpr int S[6] = {S1, S2, S3, S4, S5, S6};
};

using Win0 = Win<4, 4, 4, 4, 4, 6>;
using Win1 = Win<4, 4, 4, 4, 4, 4>;
using Win2 = Win<4, 4, 4, 5, 5, 5>;

// The box a workgroup owns, with XCD-contiguous numbering: consecutive
// workgroup ids land on different XCDs, so box ranges are dealt out so that
// each XCD walks neighbouring boxes (whose windows overlap) through its L2.
__device__ __forceinline__ int64_t box_of_block(const BoxGeom& G) {
    const int64_t per = (G.nbox + 7) / 8;
    int64_t b = G.blk0 + (int64_t)blockIdx.x;
    // reversed order: the last boxes of each XCD's range first
    if (G.rev) b = (b % 8) + 8 * (per - 1 - b / 8);
    return (b % 8) * per + b / 8;
}

// A dispatch counts its work-items in 32 bits, so a grid of boxes is
// launched in chunks of at most 2^23 workgroups (2^31 work-items); each
// chunk's geometry carries its first workgroup in blk0.
constexpr int64_t kMaxBoxBlocks = (int64_t)1 << 23;

template <class Launch>
inline hipError_t for_box_chunks(const BoxGeom& G, Launch launch) {
    const int64_t nb = ((G.nbox + 7) / 8) * 8;
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

struct OccG {
    const uint8_t* __restrict__ occ;
    int32_t n2, n3;
    __device__ __forceinline__ bool operator()(int ci, int cj, int ck) const {
        return occ[((int64_t)ci * n2 + cj) * n3 + ck] != 0;
    }
};

}  // namespace pgk
