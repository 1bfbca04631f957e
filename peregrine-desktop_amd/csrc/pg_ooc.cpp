// Out-of-core residency: the Jacobi pair in pinned host memory, swept tile
// by tile through a device window.
//
// Replaces the reference's tiled driver (solver/src/OutOfCore.jl:412-584,
// tile planning solver/src/Tiles.jl:311-512) with a Jacobi design whose
// sweep equals the in-HBM sweep bit for bit:
//   * tiles cut x, y and heading (heading periodic); the velocity axes stay
//     whole, as in Tiles.jl:330-333. A tile's interior is a multiple of the
//     box edge on each cut axis, so its boxes are boxes of the whole grid and
//     reuse the whole grid's window plan.
//   * a tile's window is the union of its boxes' planned corner ranges -- the
//     exact set of cells its backups read, from the same arithmetic that does
//     the backups -- so no reach estimate or safety margin is involved, and
//     the kernel raises a flag if a read ever falls outside it.
//   * every backup reads the old table and writes the new one, so tile order
//     and overlap change nothing: tile k's window is copied in (side stream)
//     while tile k-1 is swept, and tile k-1's interior is copied out (a second
//     side stream) behind it. Two window and two interior buffers alternate.
//   * with a 16-bit store (the precision form) the host also holds the Jacobi
//     pair of codes. A tile's window streams in as codes (half the bytes),
//     its interior's own FP32 values beside them; after the tile's sweep its
//     quantization tiles (inside the interior: the cut is on multiples of the
//     tile edge) are refit into the next sweep's offset/scale array and its
//     interior is encoded, and both the FP32 values and the codes stream out.
//     Every sweep reads the previous sweep's codes and offsets/scales, as the
//     in-HBM precision form does, so the two are equal bit for bit.
#include <algorithm>
#include <cmath>
#include <cstring>
#include "pg_internal.hpp"

namespace pgh {

struct OocTile {
    int32_t t0[3], tn[3];   // interior origin and extent: x, y, heading
    int32_t wo[3], wl[3];   // window origin (heading modulo n3) and extent
    int32_t ewo[3], ewl[3]; // the escape pass's window (set on its first sweep)
    int32_t pwo[3], pwl[3]; // the production scheme's window (set on its first sweep)
    pgk::BoxGeom G;         // the interior's boxes
};

struct Ooc {
    int32_t tile[3] = {0, 0, 0};
    std::vector<OocTile> tiles;
    float* hV[2] = {nullptr, nullptr};
    int cur = 0;
    float* dwin[2] = {nullptr, nullptr};
    float* dout[2] = {nullptr, nullptr};
    // 16-bit stores: the codes' Jacobi pair on the host, window codes, the
    // interior's own values and codes, the next sweep's offsets/scales
    uint16_t* hM[2] = {nullptr, nullptr};
    uint16_t* dwin16[2] = {nullptr, nullptr};
    float* dself[2] = {nullptr, nullptr};
    uint16_t* dcode[2] = {nullptr, nullptr};
    float2* d_tpn = nullptr;
    hipStream_t h2d = nullptr, d2h = nullptr;
    hipEvent_t ev_h2d[2] = {nullptr, nullptr}, ev_k[2] = {nullptr, nullptr}, ev_d2h[2] = {nullptr, nullptr};
    unsigned* d_miss = nullptr;
    double* d_tsum = nullptr;
    int64_t device_bytes = 0;
    double amplification = 0.0;
    int64_t colv = 0;        // cells per (x, y, heading): the velocity block
    int64_t budget = 0;      // device bytes allowed
    int64_t win_floats = 0;  // capacity of each FP32 window buffer, in cells
    bool escape_ready = false;
    bool escape_lds = false;   // the escape pass runs the LDS-staged kernel
    bool prod_ready = false;
    // the Int8 warm-start store (3 per cell) on the host, and the interiors'
    // on the device (null: the warm start is off)
    int8_t* hW = nullptr;
    int8_t* dwp[2] = {nullptr, nullptr};
};

namespace {

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

struct PlanRef {
    const std::vector<int16_t>* orig;
    const std::vector<uint8_t>* ext;
};

// A tile's window: the union of its boxes' corner ranges under each plan
// given, with `interior` the interior itself, and with `align16` widened so
// that copies of 16-bit codes stay 4-byte aligned (below).
void tile_window(const pg_solver* s, const Ooc& o, const OocTile& t, const std::vector<PlanRef>& plans,
                 bool interior, bool align16, int32_t wo[3], int32_t wl[3]) {
    const pg_grid& g = s->g;
    const pgk::BoxGeom& GG = s->G;
    const pgk::BoxGeom& G = t.G;
    const int n3 = g.n[2];
    const int32_t h0 = t.t0[2];
    // union of the boxes' corner ranges; heading relative to h0
    int32_t lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
    bool full[3] = {false, false, false};
    const int32_t b0 = G.org[0] / G.b[0], b1 = G.org[1] / G.b[1], b2 = G.org[2] / G.b[2];
    const int64_t per = (int64_t)GG.nb[3] * GG.nb[4] * GG.nb[5];
    for (int32_t i0 = 0; i0 < G.nb[0]; ++i0)
        for (int32_t i1 = 0; i1 < G.nb[1]; ++i1)
            for (int32_t i2 = 0; i2 < G.nb[2]; ++i2) {
                const int64_t head = (((int64_t)(b0 + i0) * GG.nb[1] + (b1 + i1)) * GG.nb[2] + (b2 + i2));
                const int32_t sh = (b2 + i2) * G.b[2];   // the box's first heading
                for (const PlanRef& pr : plans)
                    for (int64_t v = 0; v < per; ++v) {
                        const size_t bx = (size_t)(head * per + v) * 6;
                        for (int a = 0; a < 3; ++a) {
                            const int e = (*pr.ext)[bx + a];
                            if (e >= 255) full[a] = true;
                            int32_t l = (*pr.orig)[bx + a];
                            if (a == 2) {
                                // the plan measures a box's heading range from
                                // the box's own first heading, in (-n3/2, n3/2]
                                l = (l - sh) % n3;
                                if (l < 0) l += n3;
                                if (2 * l > n3) l -= n3;
                                l += sh - h0;
                            }
                            lo[a] = std::min(lo[a], l);
                            hi[a] = std::max(hi[a], l + e);
                        }
                    }
            }
    if (interior)
        for (int a = 0; a < 3; ++a) {
            const int32_t l = a == 2 ? 0 : t.t0[a];
            lo[a] = std::min(lo[a], l);
            hi[a] = std::max(hi[a], l + t.tn[a]);
        }
    // 16-bit codes with an odd velocity block: an even first heading and an
    // even heading count keep every host offset, pitch and row of the
    // window's copy 4-byte aligned. The DMA engines move only such copies;
    // any other is done by a copy kernel that takes CUs from the sweep
    // running beside it.
    if (align16 && (o.colv & 1) && !(n3 & 1) && lo[2] <= hi[2]) {
        if ((h0 + lo[2]) & 1) --lo[2];
        if ((hi[2] - lo[2]) & 1) ++hi[2];
    }
    for (int a = 0; a < 2; ++a) {
        if (full[a] || lo[a] > hi[a]) { lo[a] = 0; hi[a] = g.n[a]; }
        lo[a] = std::max(lo[a], 0);
        hi[a] = std::min(hi[a], g.n[a]);
        wo[a] = lo[a];
        wl[a] = hi[a] - lo[a];
    }
    if (full[2] || lo[2] > hi[2] || hi[2] - lo[2] >= n3) {
        wo[2] = 0;
        wl[2] = n3;
    } else {
        wo[2] = ((h0 + lo[2]) % n3 + n3) % n3;
        wl[2] = hi[2] - lo[2];
    }
}

// Boxes and window of every tile under the plans given (with `interior` the
// 16-bit store's windows: they hold the interior too, and are aligned).
void build_tiles(const pg_solver* s, Ooc& o, const std::vector<PlanRef>& plans, bool interior) {
    const pg_grid& g = s->g;
    const pgk::BoxGeom& GG = s->G;
    const int n3 = g.n[2];
    o.tiles.clear();
    for (int32_t x0 = 0; x0 < g.n[0]; x0 += o.tile[0])
        for (int32_t y0 = 0; y0 < g.n[1]; y0 += o.tile[1])
            for (int32_t h0 = 0; h0 < n3; h0 += o.tile[2]) {
                OocTile t{};
                const int32_t start[3] = {x0, y0, h0};
                for (int a = 0; a < 3; ++a) {
                    t.t0[a] = start[a];
                    t.tn[a] = std::min(o.tile[a], g.n[a] - start[a]);
                }
                pgk::BoxGeom& G = t.G;
                G = GG;
                int64_t nbox = 1;
                for (int a = 0; a < 6; ++a) {
                    G.org[a] = a < 3 ? t.t0[a] : 0;
                    G.rn[a] = a < 3 ? t.tn[a] : g.n[a];
                    G.nb[a] = (int32_t)ceil_div(G.rn[a], G.b[a]);
                    nbox *= G.nb[a];
                }
                G.nbox = nbox;
                tile_window(s, o, t, plans, interior, interior, t.wo, t.wl);
                o.tiles.push_back(t);
            }
}

int64_t window_cells(const OocTile& t, int64_t colv) { return (int64_t)t.wl[0] * t.wl[1] * t.wl[2] * colv; }
int64_t interior_cells(const OocTile& t, int64_t colv) { return (int64_t)t.tn[0] * t.tn[1] * t.tn[2] * colv; }

// One strided block between the host table and a device buffer: depth x
// planes of height y rows of width bytes; host rows n3 * colv cells apart,
// host planes n1 rows apart.
template <class T>
hipError_t copy_block(const pg_solver* s, const Ooc& o, T* host, T* dev, size_t dev_row, size_t dev_rows,
                      size_t width, size_t height, size_t depth, bool to_host, hipStream_t st) {
    const size_t host_row = (size_t)s->g.n[2] * (size_t)o.colv * sizeof(T);
    hipPitchedPtr hp = make_hipPitchedPtr(host, host_row, width, (size_t)s->g.n[1]);
    hipPitchedPtr dp = make_hipPitchedPtr(dev, dev_row, width, dev_rows);
    hipMemcpy3DParms p{};
    p.srcPtr = to_host ? dp : hp;
    p.dstPtr = to_host ? hp : dp;
    p.extent = make_hipExtent(width, height, depth);
    p.kind = to_host ? hipMemcpyDeviceToHost : hipMemcpyHostToDevice;
    return hipMemcpy3DAsync(&p, st);
}

// Host table -> window buffer: one block per heading piece (a heading range
// that wraps is two).
template <class T>
hipError_t copy_window_in(const pg_solver* s, const Ooc& o, const OocTile& t, const T* hV, T* dw,
                          hipStream_t st) {
    const pg_grid& g = s->g;
    const int64_t colv = o.colv;
    const int n3 = g.n[2];
    const int32_t h_first = std::min(t.wl[2], n3 - t.wo[2]);
    for (int piece = 0; piece < 2; ++piece) {
        const int32_t hs = piece == 0 ? t.wo[2] : 0;
        const int32_t hl = piece == 0 ? h_first : t.wl[2] - h_first;
        const int32_t lh = piece == 0 ? 0 : h_first;
        if (hl <= 0) continue;
        T* src = const_cast<T*>(hV) + (((int64_t)t.wo[0] * g.n[1] + t.wo[1]) * n3 + hs) * colv;
        T* dst = dw + (int64_t)lh * colv;
        hipError_t e = copy_block<T>(s, o, src, dst, (size_t)t.wl[2] * colv * sizeof(T), (size_t)t.wl[1],
                                     (size_t)hl * colv * sizeof(T), (size_t)t.wl[1], (size_t)t.wl[0], false, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Interior buffer <-> host table, one block.
template <class T>
hipError_t copy_interior(const pg_solver* s, const Ooc& o, const OocTile& t, T* dbuf, T* hV, bool to_host,
                         hipStream_t st) {
    const pg_grid& g = s->g;
    const int64_t colv = o.colv;
    T* h = hV + (((int64_t)t.t0[0] * g.n[1] + t.t0[1]) * g.n[2] + t.t0[2]) * colv;
    const size_t row = (size_t)t.tn[2] * colv * sizeof(T);
    return copy_block<T>(s, o, h, dbuf, row, (size_t)t.tn[1], row, (size_t)t.tn[1], (size_t)t.tn[0], to_host, st);
}

hipError_t copy_interior_out(const pg_solver* s, const Ooc& o, const OocTile& t, float* dout, float* hV,
                             hipStream_t st) {
    return copy_interior<float>(s, o, t, dout, hV, true, st);
}

int64_t fixed_device_bytes(const pg_solver* s) {
    return s->G.nbox * 6 * (int64_t)sizeof(int16_t)                  // window plan
           + (int64_t)s->g.n[0] * s->g.n[1] * s->g.n[2]              // occupancy
           + 3 * (int64_t)s->nctl * (int64_t)sizeof(float);          // B u per control
}

}  // namespace

int ooc_create(pg_solver* s, int64_t budget_bytes, const int32_t tile_req[3]) {
    const pg_grid& g = s->g;
    if (budget_bytes <= 0) return fail(PG_ERR_ARG, "budget_bytes must be positive");
    const bool f16 = s->vstore != PG_VSTORE_F32;
    std::vector<int16_t> orig, forig;
    std::vector<uint8_t> ext, fext;
    int rc = plan(s, true, &orig, &ext);
    if (rc) return rc;
    std::vector<PlanRef> plans{{&orig, &ext}};
    if (f16) {
        // the precision form's own read ranges (its contracted successors)
        hipStream_t st = s->ctx->stream;
        HIPCHECK(hipMalloc(&s->d_forig, (size_t)s->G.nbox * 6 * sizeof(int16_t)));
        HIPCHECK(hipMalloc(&s->d_fext, (size_t)s->G.nbox * 6));
        HIPCHECK(hipMalloc(&s->d_ffree, (size_t)s->G.nbox));
        HIPCHECK(pgk_launch_plan(s->K, &s->F, s->G, s->d_occ, s->d_bu, s->d_forig, s->d_fext, s->d_ffree,
                                 s->diag ? 1 : 0, st));
        forig.resize((size_t)s->G.nbox * 6);
        fext.resize((size_t)s->G.nbox * 6);
        HIPCHECK(hipMemcpyAsync(forig.data(), s->d_forig, forig.size() * sizeof(int16_t), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipMemcpyAsync(fext.data(), s->d_fext, fext.size(), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        plans.push_back({&forig, &fext});
    }
    // the tiles need partial sums for the largest tile only
    (void)hipFree(s->d_part);
    s->d_part = nullptr;
    s->part_cap = 0;
    s->kernel = PG_KERNEL_LDS;
    Ooc* o = new Ooc;
    s->ooc = o;
    o->colv = (int64_t)g.n[3] * g.n[4] * g.n[5];
    o->budget = budget_bytes;
    const int64_t fixed = fixed_device_bytes(s);
    // partial sums: one per workgroup of the largest tile, one total per tile
    auto need = [&](Ooc& oc) {
        int64_t win = 0, in = 0, blocks = 0;
        for (const OocTile& t : oc.tiles) {
            win = std::max(win, window_cells(t, oc.colv));
            in = std::max(in, interior_cells(t, oc.colv));
            blocks = std::max(blocks, pgk::box_blocks(t.G.nbox));
        }
        int64_t b16 = 0;
        if (f16)   // codes in and out, own values in, the next offsets/scales
            b16 = 2 * (win * 2 + in * (4 + 2)) + s->Q.ntile * (int64_t)sizeof(float2);
        return fixed + 2 * (win + in) * (int64_t)sizeof(float) + b16 + blocks * (int64_t)sizeof(double) +
               (int64_t)oc.tiles.size() * (int64_t)sizeof(double);
    };
    auto even = [&](int32_t v, int a) {
        // whole boxes, and with codes whole quantization tiles
        const int32_t b = f16 ? std::max(s->G.b[a], (int32_t)1 << pgk::kQuantShift) : s->G.b[a];
        v = std::max(v, b);
        v = (v + b - 1) / b * b;
        return std::min(v, g.n[a]);
    };
    if (tile_req[0] > 0 && tile_req[1] > 0 && tile_req[2] > 0) {
        for (int a = 0; a < 3; ++a) o->tile[a] = even(tile_req[a], a);
        build_tiles(s, *o, plans, f16);
        o->device_bytes = need(*o);
        if (o->device_bytes > budget_bytes)
            return fail(PG_ERR_NOMEM, "tiles %dx%dx%d need %lld device bytes, over the budget of %lld", o->tile[0],
                        o->tile[1], o->tile[2], (long long)o->device_bytes, (long long)budget_bytes);
    } else {
        // largest roughly square (in cells) x-y tile with the whole heading
        // axis that fits, by bisection on the fraction of each axis
        o->tile[2] = g.n[2];
        double lo = 0.0, hi = 1.0;
        bool any = false;
        for (int it = 0; it < 24; ++it) {
            const double f = it == 0 ? 1.0 : 0.5 * (lo + hi);
            Ooc trial;
            trial.colv = o->colv;
            trial.tile[0] = even((int32_t)std::lround(f * g.n[0]), 0);
            trial.tile[1] = even((int32_t)std::lround(f * g.n[1]), 1);
            trial.tile[2] = g.n[2];
            build_tiles(s, trial, plans, f16);
            const int64_t b = need(trial);
            if (b <= budget_bytes) {
                any = true;
                lo = f;
                o->tile[0] = trial.tile[0];
                o->tile[1] = trial.tile[1];
                if (it == 0) break;
            } else {
                hi = f;
            }
        }
        if (!any) return fail(PG_ERR_NOMEM, "no tile fits a device budget of %lld bytes", (long long)budget_bytes);
        build_tiles(s, *o, plans, f16);
        o->device_bytes = need(*o);
    }
    int64_t win = 0, in = 0, loaded = 0;
    for (const OocTile& t : o->tiles) {
        win = std::max(win, window_cells(t, o->colv));
        in = std::max(in, interior_cells(t, o->colv));
        loaded += window_cells(t, o->colv);
    }
    // the tile kernels index windows and interiors in 32 bits
    if (win >= ((int64_t)1 << 31))
        return fail(PG_ERR_UNSUPPORTED, "a tile window of %lld cells: tiles hold fewer than 2^31 "
                    "(give a smaller tile)", (long long)win);
    o->amplification = (double)loaded / (double)s->ncell;
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t x) { if (e == hipSuccess) e = x; };
    for (int j = 0; j < 2; ++j) {
        chk(hipHostMalloc((void**)&o->hV[j], (size_t)s->ncell * sizeof(float), hipHostMallocDefault));
        chk(hipMalloc(&o->dwin[j], (size_t)win * sizeof(float)));
        o->win_floats = win;
        chk(hipMalloc(&o->dout[j], (size_t)in * sizeof(float)));
        chk(hipEventCreateWithFlags(&o->ev_h2d[j], hipEventDisableTiming));
        chk(hipEventCreateWithFlags(&o->ev_k[j], hipEventDisableTiming));
        chk(hipEventCreateWithFlags(&o->ev_d2h[j], hipEventDisableTiming));
    }
    if (f16) {
        for (int j = 0; j < 2; ++j) {
            chk(hipHostMalloc((void**)&o->hM[j], (size_t)s->ncell * sizeof(uint16_t), hipHostMallocDefault));
            chk(hipMalloc(&o->dwin16[j], (size_t)win * sizeof(uint16_t)));
            chk(hipMalloc(&o->dself[j], (size_t)in * sizeof(float)));
            chk(hipMalloc(&o->dcode[j], (size_t)in * sizeof(uint16_t)));
        }
        chk(hipMalloc(&o->d_tpn, (size_t)s->Q.ntile * sizeof(float2)));
    }
    chk(hipStreamCreate(&o->h2d));
    chk(hipStreamCreate(&o->d2h));
    chk(hipMalloc(&o->d_miss, sizeof(unsigned)));
    chk(hipMalloc(&o->d_tsum, o->tiles.size() * sizeof(double)));
    int64_t blocks = 0;
    for (const OocTile& t : o->tiles) blocks = std::max(blocks, pgk::box_blocks(t.G.nbox));
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? PG_ERR_NOMEM : PG_ERR_HIP, "out-of-core buffers: %s",
                    hipGetErrorString(e));
    return ensure_partials(s, blocks);
}

void ooc_release(pg_solver* s) {
    Ooc* o = s->ooc;
    if (!o) return;
    if (o->h2d) (void)hipStreamSynchronize(o->h2d);
    if (o->d2h) (void)hipStreamSynchronize(o->d2h);
    for (int j = 0; j < 2; ++j) {
        (void)hipHostFree(o->hV[j]);
        (void)hipFree(o->dwin[j]);
        (void)hipFree(o->dout[j]);
        (void)hipHostFree(o->hM[j]);
        (void)hipFree(o->dwin16[j]);
        (void)hipFree(o->dself[j]);
        (void)hipFree(o->dcode[j]);
        if (o->ev_h2d[j]) (void)hipEventDestroy(o->ev_h2d[j]);
        if (o->ev_k[j]) (void)hipEventDestroy(o->ev_k[j]);
        if (o->ev_d2h[j]) (void)hipEventDestroy(o->ev_d2h[j]);
    }
    if (o->h2d) (void)hipStreamDestroy(o->h2d);
    if (o->d2h) (void)hipStreamDestroy(o->d2h);
    (void)hipFree(o->d_miss);
    (void)hipFree(o->d_tsum);
    (void)hipFree(o->d_tpn);
    (void)hipHostFree(o->hW);
    (void)hipFree(o->dwp[0]);
    (void)hipFree(o->dwp[1]);
    delete o;
    s->ooc = nullptr;
}

namespace {

// The codes of the whole host table from its FP32 values, tile by tile:
// every quantization tile refit from its reached values (as a whole-table
// load in HBM does), then encoded.
int ooc_encode_all(pg_solver* s) {
    Ooc& o = *s->ooc;
    hipStream_t st = s->ctx->stream;
    for (const OocTile& t : o.tiles) {
        HIPCHECK(copy_interior<float>(s, o, t, o.dself[0], o.hV[o.cur], false, st));
        HIPCHECK(pgk_launch_interior_stats(o.dself[0], t.t0, t.tn, s->K, s->Q, s->d_tstat, st));
        HIPCHECK(pgk_launch_interior_refit(s->store, t.t0, t.tn, s->Q, s->d_tstat, s->d_tp, s->d_tp, s->d_trefit, 1,
                                           st));
        HIPCHECK(pgk_launch_interior_encode(s->store, o.dself[0], t.t0, t.tn, s->K, s->Q, s->d_tp, o.dcode[0], st));
        HIPCHECK(copy_interior<uint16_t>(s, o, t, o.dcode[0], o.hM[o.cur], true, st));
        HIPCHECK(hipStreamSynchronize(st));
    }
    s->f16_active = true;
    return PG_OK;
}

// One precision-form sweep over the tiles (the 16-bit twin of the FP32
// pipeline below): window codes and interior values in on one side stream,
// the sweep, refit and encode on the solver stream, values and codes out on
// the other side stream.
int ooc_sweep16(pg_solver* s, double* sum_decrease, float* linf) {
    Ooc& o = *s->ooc;
    hipStream_t st = s->ctx->stream;
    const float* hin = o.hV[o.cur];
    float* hout = o.hV[1 - o.cur];
    const uint16_t* min_ = o.hM[o.cur];
    uint16_t* mout = o.hM[1 - o.cur];
    const int64_t nt = (int64_t)o.tiles.size();
    HIPCHECK(hipMemsetAsync(s->d_linf, 0, sizeof(unsigned), st));
    HIPCHECK(hipMemsetAsync(o.d_miss, 0, sizeof(unsigned), st));
    HIPCHECK(hipEventRecord(s->ev0, st));
    auto load = [&](int64_t k, int j) -> hipError_t {
        const OocTile& t = o.tiles[(size_t)k];
        hipError_t e = copy_window_in<uint16_t>(s, o, t, min_, o.dwin16[j], o.h2d);
        if (e == hipSuccess) e = copy_interior<float>(s, o, t, o.dself[j], const_cast<float*>(hin), false, o.h2d);
        return e;
    };
    HIPCHECK(hipStreamWaitEvent(o.h2d, s->ev0, 0));
    HIPCHECK(load(0, 0));
    HIPCHECK(hipEventRecord(o.ev_h2d[0], o.h2d));
    for (int64_t k = 0; k < nt; ++k) {
        const int j = (int)(k & 1);
        const OocTile& t = o.tiles[(size_t)k];
        HIPCHECK(hipStreamWaitEvent(st, o.ev_h2d[j], 0));
        if (k >= 2) HIPCHECK(hipStreamWaitEvent(st, o.ev_d2h[j], 0));
        HIPCHECK(pgk_launch_sweep_fast_tile(s->shape, s->diag ? 1 : 0, s->store, s->K, s->F, t.G, s->Q, o.dwin16[j],
                                            s->d_tp, t.wo, t.wl, o.d_miss, o.dself[j], o.dout[j], s->d_occ, s->d_bu,
                                            s->d_worig, s->d_ffree, s->d_part, s->d_linf, s->d_tstat, st));
        HIPCHECK(pgk_launch_reduce(s->d_part, pgk::box_blocks(t.G.nbox), o.d_tsum + k, nullptr, st));
        // this tile's quantization tiles: refit into the next sweep's array, encode
        HIPCHECK(pgk_launch_interior_refit(s->store, t.t0, t.tn, s->Q, s->d_tstat, s->d_tp, o.d_tpn, s->d_trefit, 0,
                                           st));
        HIPCHECK(pgk_launch_interior_encode(s->store, o.dout[j], t.t0, t.tn, s->K, s->Q, o.d_tpn, o.dcode[j], st));
        HIPCHECK(hipEventRecord(o.ev_k[j], st));
        if (k + 1 < nt) {
            const int jn = 1 - j;
            if (k >= 1) HIPCHECK(hipStreamWaitEvent(o.h2d, o.ev_k[jn], 0));
            HIPCHECK(load(k + 1, jn));
            HIPCHECK(hipEventRecord(o.ev_h2d[jn], o.h2d));
        }
        HIPCHECK(hipStreamWaitEvent(o.d2h, o.ev_k[j], 0));
        HIPCHECK(copy_interior<float>(s, o, t, o.dout[j], hout, true, o.d2h));
        HIPCHECK(copy_interior<uint16_t>(s, o, t, o.dcode[j], mout, true, o.d2h));
        HIPCHECK(hipEventRecord(o.ev_d2h[j], o.d2h));
    }
    std::vector<double> tsum((size_t)nt);
    unsigned hl = 0, miss = 0;
    HIPCHECK(hipMemcpyAsync(tsum.data(), o.d_tsum, (size_t)nt * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(&hl, s->d_linf, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(&miss, o.d_miss, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    HIPCHECK(hipStreamSynchronize(o.d2h));
    HIPCHECK(hipEventRecord(s->ev1, st));
    HIPCHECK(hipEventSynchronize(s->ev1));
    HIPCHECK(hipEventElapsedTime(&s->last_ms, s->ev0, s->ev1));
    if (miss) return fail(PG_ERR_STATE, "a backup read a cell outside its tile's window");
    // seeds hold 0 in both tables already (a backup never goes below 0)
    std::swap(s->d_tp, o.d_tpn);
    o.cur = 1 - o.cur;
    double d = 0.0;
    for (double v : tsum) d += v;
    float lf;
    std::memcpy(&lf, &hl, 4);
    if (sum_decrease) *sum_decrease = d;
    if (linf) *linf = lf;
    return PG_OK;
}

}  // namespace

namespace {

// FP32 window buffers of `win` cells each, if they hold fewer, within the
// device budget given at creation.
int grow_windows(pg_solver* s, int64_t win, const char* what) {
    Ooc& o = *s->ooc;
    if (win <= o.win_floats) return PG_OK;
    if (win >= ((int64_t)1 << 31))
        return fail(PG_ERR_UNSUPPORTED, "%s tile windows of %lld cells: tiles hold fewer than 2^31 (give a smaller "
                    "tile)", what, (long long)win);
    const int64_t grown = o.device_bytes + 2 * (win - o.win_floats) * (int64_t)sizeof(float);
    if (grown > o.budget)
        return fail(PG_ERR_NOMEM, "%s tile windows need %lld device bytes, over the budget of %lld", what,
                    (long long)grown, (long long)o.budget);
    for (int j = 0; j < 2; ++j) {
        (void)hipFree(o.dwin[j]);
        o.dwin[j] = nullptr;
    }
    hipError_t a = hipMalloc(&o.dwin[0], (size_t)win * sizeof(float));
    if (a == hipSuccess) a = hipMalloc(&o.dwin[1], (size_t)win * sizeof(float));
    if (a != hipSuccess) {
        // leave the solver whole: the other sweeps need their windows back
        (void)hipFree(o.dwin[0]);
        (void)hipFree(o.dwin[1]);
        o.dwin[0] = o.dwin[1] = nullptr;
        hipError_t b = hipMalloc(&o.dwin[0], (size_t)o.win_floats * sizeof(float));
        if (b == hipSuccess) b = hipMalloc(&o.dwin[1], (size_t)o.win_floats * sizeof(float));
        if (b != hipSuccess) return fail(PG_ERR_NOMEM, "out-of-core window buffers lost: %s", hipGetErrorString(b));
        return fail(a == hipErrorOutOfMemory ? PG_ERR_NOMEM : PG_ERR_HIP, "%s window buffers: %s", what,
                    hipGetErrorString(a));
    }
    o.device_bytes = grown;
    o.win_floats = win;
    return PG_OK;
}

// The escape pass's windows, on its first sweep. The escape backups read
// from blocked cells too and without the probes' rejections (a blocked cell
// moves through obstacle space), so the windows come from a plan over an
// empty occupancy -- every state, every successor: a superset of what the
// escape pass reads -- joined with the interior (each cell reads its own
// value). The FP32 window buffers grow if a window needs more room, within
// the device budget.
int ooc_escape_prepare(pg_solver* s) {
    Ooc& o = *s->ooc;
    if (o.escape_ready) return PG_OK;
    int rc = PG_OK;
    const pg_grid& g = s->g;
    hipStream_t st = s->ctx->stream;
    const int64_t nbox = s->G.nbox;
    const size_t nocc = (size_t)g.n[0] * g.n[1] * g.n[2];
    uint8_t* zocc = nullptr;
    int16_t* d_orig = nullptr;
    uint8_t* d_ext = nullptr;
    std::vector<int16_t> orig((size_t)nbox * 6);
    std::vector<uint8_t> ext((size_t)nbox * 6);
    hipError_t e = hipMalloc(&zocc, nocc);
    if (e == hipSuccess) e = hipMalloc(&d_orig, (size_t)nbox * 6 * sizeof(int16_t));
    if (e == hipSuccess) e = hipMalloc(&d_ext, (size_t)nbox * 6);
    if (e == hipSuccess) e = hipMemsetAsync(zocc, 0, nocc, st);
    if (e == hipSuccess)
        e = pgk_launch_plan(s->K, nullptr, s->G, zocc, s->d_bu, d_orig, d_ext, nullptr, s->diag ? 1 : 0, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(orig.data(), d_orig, orig.size() * sizeof(int16_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(ext.data(), d_ext, ext.size(), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(zocc);
    (void)hipFree(d_ext);
    if (e != hipSuccess) {
        (void)hipFree(d_orig);
        return fail(e == hipErrorOutOfMemory ? PG_ERR_NOMEM : PG_ERR_HIP, "escape window plan: %s",
                    hipGetErrorString(e));
    }
    const std::vector<PlanRef> plans{{&orig, &ext}};
    int64_t win = 0, blocks = 0;
    for (OocTile& t : o.tiles) {
        tile_window(s, o, t, plans, true, false, t.ewo, t.ewl);
        win = std::max(win, (int64_t)t.ewl[0] * t.ewl[1] * t.ewl[2] * o.colv);
        blocks = std::max(blocks, std::max(ceil_div(interior_cells(t, o.colv), 256), pgk::box_blocks(t.G.nbox)));
    }
    rc = ensure_partials(s, blocks);
    if (rc == PG_OK) rc = grow_windows(s, win, "the escape pass's");
    if (rc) {
        (void)hipFree(d_orig);
        return rc;
    }
    // The LDS-staged escape kernel stages each box's window from the plan's
    // origins, which then stay on the device (12 bytes per box): within the
    // budget, and with the control table in LDS, the tiles run it; otherwise
    // the global-gather escape kernel, which reads the tile window directly.
    const int64_t orig_bytes = nbox * 6 * (int64_t)sizeof(int16_t);
    o.escape_lds = s->nctl <= kLdsMaxCtl && o.device_bytes + orig_bytes <= o.budget;
    if (o.escape_lds) {
        (void)hipFree(s->d_worig_esc);
        s->d_worig_esc = d_orig;
        o.device_bytes += orig_bytes;
        // the window shape that holds the most boxes' escape reads whole
        const int (*shapes)[6] = nullptr;
        const int nsh = pgk_window_shapes(&shapes);
        int64_t best = -1;
        for (int i = 0; i < nsh; ++i) {
            int64_t cov = 0;
            for (int64_t b = 0; b < nbox; ++b) {
                bool ok = true;
                for (int k = 0; k < 6; ++k) ok = ok && ext[(size_t)b * 6 + k] <= shapes[i][k];
                cov += ok ? 1 : 0;
            }
            if (cov > best) { best = cov; s->esc_shape = i; }
        }
        s->esc_coverage = nbox > 0 ? (double)best / (double)nbox : 1.0;
    } else {
        (void)hipFree(d_orig);
    }
    o.escape_ready = true;
    return PG_OK;
}

// The escape window of tile t in, host FP32 table -> buffer (the value
// pass's copy with the escape window's origin and extent).
hipError_t copy_escape_window_in(const pg_solver* s, const Ooc& o, const OocTile& t, const float* hV, float* dw,
                                 hipStream_t st) {
    OocTile e = t;
    for (int a = 0; a < 3; ++a) {
        e.wo[a] = t.ewo[a];
        e.wl[a] = t.ewl[a];
    }
    return copy_window_in<float>(s, o, e, hV, dw, st);
}

}  // namespace

namespace {
int ooc_prod_prepare(pg_solver* s);
hipError_t copy_prod_window_in(const pg_solver* s, const Ooc& o, const OocTile& t, const float* hV, float* dw,
                               hipStream_t st);
}  // namespace

// One escape sweep over the tiles (Jacobi, as in HBM): the pipeline of the
// FP32 value sweep with the escape kernel.
int ooc_escape_sweep(pg_solver* s, double* sum_improve, float* linf) {
    if (s->f16_active) return fail(PG_ERR_STATE, "a 16-bit store must be polished (pg_polish) before the escape pass");
    // with the production candidates the reads follow the scheme's reach
    const bool prod = s->K.rounds > 0 || s->K.ntau > 1;
    int rc = prod ? ooc_prod_prepare(s) : ooc_escape_prepare(s);
    if (rc) return rc;
    Ooc& o = *s->ooc;
    hipStream_t st = s->ctx->stream;
    const float* hin = o.hV[o.cur];
    float* hout = o.hV[1 - o.cur];
    const int64_t nt = (int64_t)o.tiles.size();
    HIPCHECK(hipMemsetAsync(s->d_linf, 0, sizeof(unsigned), st));
    HIPCHECK(hipMemsetAsync(o.d_miss, 0, sizeof(unsigned), st));
    HIPCHECK(hipEventRecord(s->ev0, st));
    HIPCHECK(hipStreamWaitEvent(o.h2d, s->ev0, 0));
    auto load = [&](const OocTile& t, float* dw) {
        return prod ? copy_prod_window_in(s, o, t, hin, dw, o.h2d) : copy_escape_window_in(s, o, t, hin, dw, o.h2d);
    };
    HIPCHECK(load(o.tiles[0], o.dwin[0]));
    HIPCHECK(hipEventRecord(o.ev_h2d[0], o.h2d));
    for (int64_t k = 0; k < nt; ++k) {
        const int j = (int)(k & 1);
        const OocTile& t = o.tiles[(size_t)k];
        HIPCHECK(hipStreamWaitEvent(st, o.ev_h2d[j], 0));
        if (k >= 2) HIPCHECK(hipStreamWaitEvent(st, o.ev_d2h[j], 0));
        int64_t nblk = 0;
        if (!prod && o.escape_lds) {
            HIPCHECK(pgk_launch_escape_lds_tile(s->esc_shape, s->diag ? 1 : 0, s->K, t.G, o.dwin[j], t.ewo, t.ewl,
                                                o.dout[j], s->d_occ, s->d_bu, s->d_worig_esc, s->d_pfree, s->d_part,
                                                s->d_linf, o.d_miss, st));
            nblk = pgk::box_blocks(t.G.nbox);
        } else {
            HIPCHECK(pgk_launch_escape_tile(s->K, t.t0, t.tn, o.dwin[j], prod ? t.pwo : t.ewo,
                                            prod ? t.pwl : t.ewl, o.dout[j], s->d_occ, s->d_bu, s->d_uraw,
                                            s->d_part, s->d_linf, o.d_miss, st, &nblk));
        }
        HIPCHECK(pgk_launch_reduce(s->d_part, nblk, o.d_tsum + k, nullptr, st));
        HIPCHECK(hipEventRecord(o.ev_k[j], st));
        if (k + 1 < nt) {
            const int jn = 1 - j;
            if (k >= 1) HIPCHECK(hipStreamWaitEvent(o.h2d, o.ev_k[jn], 0));
            HIPCHECK(load(o.tiles[(size_t)k + 1], o.dwin[jn]));
            HIPCHECK(hipEventRecord(o.ev_h2d[jn], o.h2d));
        }
        HIPCHECK(hipStreamWaitEvent(o.d2h, o.ev_k[j], 0));
        HIPCHECK(copy_interior_out(s, o, t, o.dout[j], hout, o.d2h));
        HIPCHECK(hipEventRecord(o.ev_d2h[j], o.d2h));
    }
    std::vector<double> tsum((size_t)nt);
    unsigned hl = 0, miss = 0;
    HIPCHECK(hipMemcpyAsync(tsum.data(), o.d_tsum, (size_t)nt * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(&hl, s->d_linf, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(&miss, o.d_miss, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    HIPCHECK(hipStreamSynchronize(o.d2h));
    HIPCHECK(hipEventRecord(s->ev1, st));
    HIPCHECK(hipEventSynchronize(s->ev1));
    HIPCHECK(hipEventElapsedTime(&s->last_ms, s->ev0, s->ev1));
    if (miss) return fail(PG_ERR_STATE, "an escape backup read a cell outside its tile's window");
    o.cur = 1 - o.cur;
    double d = 0.0;
    for (double v : tsum) d += v;
    float lf;
    std::memcpy(&lf, &hl, 4);
    if (sum_improve) *sum_improve = d;
    if (linf) *linf = lf;
    return PG_OK;
}

namespace {

// The production scheme's windows, on its first sweep: interior plus a halo
// from the scheme's reach (every grid heading, velocity and spin, every
// lattice command, commands sampled over the octahedron -- the pattern
// search and the warm start evaluate commands on it -- and every horizon),
// two cells of margin each side for the interpolation corners. Pattern-search
// and warm-start commands are data-dependent, so the windows are a bound, not
// a plan: a read outside one raises the miss flag and fails the sweep.
int ooc_prod_prepare(pg_solver* s) {
    Ooc& o = *s->ooc;
    if (o.prod_ready) return PG_OK;
    const pg_grid& g = s->g;
    hipStream_t st = s->ctx->stream;
    // octahedron |u1| + |u2| + |u3| = 1: every face, barycentric step 1/12
    std::vector<float> raw;
    const int M = 12;
    for (int f = 0; f < 8; ++f)
        for (int a = 0; a <= M; ++a)
            for (int b = 0; a + b <= M; ++b) {
                const int c = M - a - b;
                raw.push_back((f & 1 ? -1.0f : 1.0f) * (float)a / M);
                raw.push_back((f & 2 ? -1.0f : 1.0f) * (float)b / M);
                raw.push_back((f & 4 ? -1.0f : 1.0f) * (float)c / M);
            }
    const int ns = (int)(raw.size() / 3);
    float* d_raw = nullptr;
    int* d_out = nullptr;
    int h_out[6] = {0, 0, 0, 0, 0, 0};
    hipError_t e = hipMalloc(&d_raw, raw.size() * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&d_out, sizeof h_out);
    if (e == hipSuccess) e = hipMemcpyAsync(d_raw, raw.data(), raw.size() * sizeof(float), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_out, h_out, sizeof h_out, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = pgk_launch_reach(s->K, s->d_bu, d_raw, ns, d_out, st);
    if (e == hipSuccess) e = hipMemcpyAsync(h_out, d_out, sizeof h_out, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d_raw);
    (void)hipFree(d_out);
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? PG_ERR_NOMEM : PG_ERR_HIP, "reach pass: %s", hipGetErrorString(e));
    const int n3 = g.n[2];
    int64_t win = 0, blocks = 0;
    for (OocTile& t : o.tiles) {
        for (int a = 0; a < 2; ++a) {
            const int64_t lo = std::max<int64_t>(0, (int64_t)t.t0[a] + h_out[2 * a] - 2);
            const int64_t hi = std::min<int64_t>(g.n[a], (int64_t)t.t0[a] + t.tn[a] + h_out[2 * a + 1] + 2);
            t.pwo[a] = (int32_t)lo;
            t.pwl[a] = (int32_t)(hi - lo);
        }
        const int64_t lo = (int64_t)t.t0[2] + h_out[4] - 2, hi = (int64_t)t.t0[2] + t.tn[2] + h_out[5] + 2;
        if (hi - lo >= n3) {
            t.pwo[2] = 0;
            t.pwl[2] = n3;
        } else {
            t.pwo[2] = (int32_t)(((lo % n3) + n3) % n3);
            t.pwl[2] = (int32_t)(hi - lo);
        }
        win = std::max(win, (int64_t)t.pwl[0] * t.pwl[1] * t.pwl[2] * o.colv);
        blocks = std::max(blocks, ceil_div(interior_cells(t, o.colv), 256));
    }
    int rc = ensure_partials(s, blocks);
    if (rc) return rc;
    rc = grow_windows(s, win, "the production scheme's");
    if (rc) return rc;
    o.prod_ready = true;
    return PG_OK;
}

// A tile's warm-start entries (3 bytes per cell) between the host store and
// an interior buffer: one strided block.
hipError_t copy_warm(const pg_solver* s, const Ooc& o, const OocTile& t, int8_t* dbuf, bool to_host,
                     hipStream_t st) {
    const pg_grid& g = s->g;
    const size_t cell = 3 * (size_t)o.colv;   // bytes per (x, y, heading)
    int8_t* h = o.hW + ((((int64_t)t.t0[0] * g.n[1] + t.t0[1]) * g.n[2] + t.t0[2]) * (int64_t)cell);
    const size_t row = (size_t)t.tn[2] * cell;
    hipPitchedPtr hp = make_hipPitchedPtr(h, (size_t)g.n[2] * cell, row, (size_t)g.n[1]);
    hipPitchedPtr dp = make_hipPitchedPtr(dbuf, row, row, (size_t)t.tn[1]);
    hipMemcpy3DParms p{};
    p.srcPtr = to_host ? dp : hp;
    p.dstPtr = to_host ? hp : dp;
    p.extent = make_hipExtent(row, (size_t)t.tn[1], (size_t)t.tn[0]);
    p.kind = to_host ? hipMemcpyDeviceToHost : hipMemcpyHostToDevice;
    return hipMemcpy3DAsync(&p, st);
}

hipError_t copy_prod_window_in(const pg_solver* s, const Ooc& o, const OocTile& t, const float* hV, float* dw,
                               hipStream_t st) {
    OocTile e = t;
    for (int a = 0; a < 3; ++a) {
        e.wo[a] = t.pwo[a];
        e.wl[a] = t.pwl[a];
    }
    return copy_window_in<float>(s, o, e, hV, dw, st);
}

// One production-scheme sweep over the tiles (Jacobi, as in HBM): the FP32
// pipeline with the global-gather kernel, the tile's warm-start entries in
// and out beside its window and interior (each cell reads and writes only
// its own entry).
int ooc_sweep_prod(pg_solver* s, double* sum_decrease, float* linf) {
    int rc = ooc_prod_prepare(s);
    if (rc) return rc;
    Ooc& o = *s->ooc;
    hipStream_t st = s->ctx->stream;
    const float* hin = o.hV[o.cur];
    float* hout = o.hV[1 - o.cur];
    const int64_t nt = (int64_t)o.tiles.size();
    const bool warm = o.hW != nullptr;
    HIPCHECK(hipMemsetAsync(s->d_linf, 0, sizeof(unsigned), st));
    HIPCHECK(hipMemsetAsync(o.d_miss, 0, sizeof(unsigned), st));
    HIPCHECK(hipEventRecord(s->ev0, st));
    auto load = [&](int64_t k, int j) -> hipError_t {
        const OocTile& t = o.tiles[(size_t)k];
        hipError_t e = copy_prod_window_in(s, o, t, hin, o.dwin[j], o.h2d);
        if (e == hipSuccess && warm) e = copy_warm(s, o, t, o.dwp[j], false, o.h2d);
        return e;
    };
    HIPCHECK(hipStreamWaitEvent(o.h2d, s->ev0, 0));
    HIPCHECK(load(0, 0));
    HIPCHECK(hipEventRecord(o.ev_h2d[0], o.h2d));
    for (int64_t k = 0; k < nt; ++k) {
        const int j = (int)(k & 1);
        const OocTile& t = o.tiles[(size_t)k];
        HIPCHECK(hipStreamWaitEvent(st, o.ev_h2d[j], 0));
        if (k >= 2) HIPCHECK(hipStreamWaitEvent(st, o.ev_d2h[j], 0));
        int64_t nblk = 0;
        HIPCHECK(pgk_launch_sweep_global_tile(s->K, t.t0, t.tn, o.dwin[j], t.pwo, t.pwl, o.dout[j], s->d_occ, s->d_bu,
                                              s->d_uraw, warm ? o.dwp[j] : nullptr, s->d_part, s->d_linf, o.d_miss,
                                              st, &nblk));
        HIPCHECK(pgk_launch_reduce(s->d_part, nblk, o.d_tsum + k, nullptr, st));
        HIPCHECK(hipEventRecord(o.ev_k[j], st));
        if (k + 1 < nt) {
            const int jn = 1 - j;
            if (k >= 1) HIPCHECK(hipStreamWaitEvent(o.h2d, o.ev_k[jn], 0));
            // the warm-start buffer jn is also waiting on its copy out
            if (k >= 1 && warm) HIPCHECK(hipStreamWaitEvent(o.h2d, o.ev_d2h[jn], 0));
            HIPCHECK(load(k + 1, jn));
            HIPCHECK(hipEventRecord(o.ev_h2d[jn], o.h2d));
        }
        HIPCHECK(hipStreamWaitEvent(o.d2h, o.ev_k[j], 0));
        HIPCHECK(copy_interior_out(s, o, t, o.dout[j], hout, o.d2h));
        if (warm) HIPCHECK(copy_warm(s, o, t, o.dwp[j], true, o.d2h));
        HIPCHECK(hipEventRecord(o.ev_d2h[j], o.d2h));
    }
    std::vector<double> tsum((size_t)nt);
    unsigned hl = 0, miss = 0;
    HIPCHECK(hipMemcpyAsync(tsum.data(), o.d_tsum, (size_t)nt * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(&hl, s->d_linf, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(&miss, o.d_miss, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    HIPCHECK(hipStreamSynchronize(o.d2h));
    HIPCHECK(hipEventRecord(s->ev1, st));
    HIPCHECK(hipEventSynchronize(s->ev1));
    HIPCHECK(hipEventElapsedTime(&s->last_ms, s->ev0, s->ev1));
    if (miss) return fail(PG_ERR_STATE, "a backup read a cell outside its tile's window (reach bound)");
    for (int64_t i : s->h_seeds) hout[i] = 0.0f;
    o.cur = 1 - o.cur;
    double d = 0.0;
    for (double v : tsum) d += v;
    float lf;
    std::memcpy(&lf, &hl, 4);
    if (sum_decrease) *sum_decrease = d;
    if (linf) *linf = lf;
    return PG_OK;
}

}  // namespace

bool ooc_production(const pg_solver* s) {
    return s->ooc && (s->K.rounds > 0 || s->K.ntau > 1 || s->ooc->hW != nullptr);
}

int ooc_set_warm_start(pg_solver* s, bool on) {
    Ooc& o = *s->ooc;
    if (!on) {
        (void)hipHostFree(o.hW);
        (void)hipFree(o.dwp[0]);
        (void)hipFree(o.dwp[1]);
        o.hW = nullptr;
        o.dwp[0] = o.dwp[1] = nullptr;
        return PG_OK;
    }
    if (o.hW) return PG_OK;
    int64_t in = 0;
    for (const OocTile& t : o.tiles) in = std::max(in, interior_cells(t, o.colv));
    hipError_t e = hipHostMalloc((void**)&o.hW, 3 * (size_t)s->ncell, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(&o.dwp[0], 3 * (size_t)in);
    if (e == hipSuccess) e = hipMalloc(&o.dwp[1], 3 * (size_t)in);
    if (e != hipSuccess) {
        ooc_set_warm_start(s, false);
        return fail(e == hipErrorOutOfMemory ? PG_ERR_NOMEM : PG_ERR_HIP, "warm-start store allocation failed: %s",
                    hipGetErrorString(e));
    }
    std::memset(o.hW, 0, 3 * (size_t)s->ncell);
    return PG_OK;
}

int ooc_download_warm_start(const pg_solver* s, int8_t* dst) {
    if (!s->ooc->hW) return fail(PG_ERR_STATE, "the warm start is off");
    std::memcpy(dst, s->ooc->hW, 3 * (size_t)s->ncell);
    return PG_OK;
}

int ooc_fill_seed(pg_solver* s) {
    Ooc& o = *s->ooc;
    o.cur = 0;
    if (o.hW) std::memset(o.hW, 0, 3 * (size_t)s->ncell);
    std::fill(o.hV[0], o.hV[0] + s->ncell, s->p.cap);
    for (int64_t i : s->h_seeds) o.hV[0][i] = 0.0f;
    if (s->vstore != PG_VSTORE_F32) return ooc_encode_all(s);
    return PG_OK;
}

int ooc_sweep(pg_solver* s, double* sum_decrease, float* linf) {
    if (s->f16_active) return ooc_sweep16(s, sum_decrease, linf);
    if (ooc_production(s)) return ooc_sweep_prod(s, sum_decrease, linf);
    Ooc& o = *s->ooc;
    hipStream_t st = s->ctx->stream;
    const float* hin = o.hV[o.cur];
    float* hout = o.hV[1 - o.cur];
    const int64_t nt = (int64_t)o.tiles.size();
    HIPCHECK(hipMemsetAsync(s->d_linf, 0, sizeof(unsigned), st));
    HIPCHECK(hipMemsetAsync(o.d_miss, 0, sizeof(unsigned), st));
    HIPCHECK(hipEventRecord(s->ev0, st));
    // the first window; later ones are issued one tile ahead
    HIPCHECK(hipStreamWaitEvent(o.h2d, s->ev0, 0));
    HIPCHECK(copy_window_in<float>(s, o, o.tiles[0], hin, o.dwin[0], o.h2d));
    HIPCHECK(hipEventRecord(o.ev_h2d[0], o.h2d));
    for (int64_t k = 0; k < nt; ++k) {
        const int j = (int)(k & 1);
        const OocTile& t = o.tiles[(size_t)k];
        // sweep tile k once its window is in and interior buffer j is free
        HIPCHECK(hipStreamWaitEvent(st, o.ev_h2d[j], 0));
        if (k >= 2) HIPCHECK(hipStreamWaitEvent(st, o.ev_d2h[j], 0));
        HIPCHECK(pgk_launch_sweep_tile(s->shape, s->diag ? 1 : 0, s->K, t.G, o.dwin[j], t.wo, t.wl, o.dout[j],
                                       nullptr, s->d_occ, s->d_bu, s->d_worig, s->d_pfree, s->d_part, s->d_linf,
                                       o.d_miss, st));
        HIPCHECK(pgk_launch_reduce(s->d_part, pgk::box_blocks(t.G.nbox), o.d_tsum + k, nullptr, st));
        HIPCHECK(hipEventRecord(o.ev_k[j], st));
        // tile k+1's window, into the buffer tile k-1 has finished with
        if (k + 1 < nt) {
            const int jn = 1 - j;
            if (k >= 1) HIPCHECK(hipStreamWaitEvent(o.h2d, o.ev_k[jn], 0));
            HIPCHECK(copy_window_in<float>(s, o, o.tiles[(size_t)k + 1], hin, o.dwin[jn], o.h2d));
            HIPCHECK(hipEventRecord(o.ev_h2d[jn], o.h2d));
        }
        // tile k's interior out
        HIPCHECK(hipStreamWaitEvent(o.d2h, o.ev_k[j], 0));
        HIPCHECK(copy_interior_out(s, o, t, o.dout[j], hout, o.d2h));
        HIPCHECK(hipEventRecord(o.ev_d2h[j], o.d2h));
    }
    std::vector<double> tsum((size_t)nt);
    unsigned hl = 0, miss = 0;
    HIPCHECK(hipMemcpyAsync(tsum.data(), o.d_tsum, (size_t)nt * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(&hl, s->d_linf, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(&miss, o.d_miss, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    HIPCHECK(hipStreamSynchronize(o.d2h));
    HIPCHECK(hipEventRecord(s->ev1, st));
    HIPCHECK(hipEventSynchronize(s->ev1));
    HIPCHECK(hipEventElapsedTime(&s->last_ms, s->ev0, s->ev1));
    if (miss) return fail(PG_ERR_STATE, "a backup read a cell outside its tile's window");
    for (int64_t i : s->h_seeds) hout[i] = 0.0f;
    o.cur = 1 - o.cur;
    double d = 0.0;
    for (double v : tsum) d += v;
    float lf;
    std::memcpy(&lf, &hl, 4);
    if (sum_decrease) *sum_decrease = d;
    if (linf) *linf = lf;
    return PG_OK;
}

int ooc_download(pg_solver* s, float* dst, int64_t off, int64_t n) {
    std::memcpy(dst, s->ooc->hV[s->ooc->cur] + off, (size_t)n * sizeof(float));
    return PG_OK;
}

int ooc_upload(pg_solver* s, const float* src, int64_t off, int64_t n) {
    Ooc& o = *s->ooc;
    if (!s->seeded) std::fill(o.hV[o.cur], o.hV[o.cur] + s->ncell, s->p.cap);
    std::memcpy(o.hV[o.cur] + off, src, (size_t)n * sizeof(float));
    if (s->vstore != PG_VSTORE_F32 && (s->f16_active || !s->seeded)) return ooc_encode_all(s);
    return PG_OK;
}

const uint16_t* ooc_host_codes(const pg_solver* s) { return s->ooc->hM[s->ooc->cur]; }

// The precision form's end: the host FP32 table is the table, the codes
// are dropped and the FP32 tiles sweep from here on.
int ooc_polish_start(pg_solver* s) {
    s->f16_active = false;
    s->kernel = PG_KERNEL_LDS;
    return PG_OK;
}

int ooc_info(pg_solver* s, int32_t tile[3], int64_t* ntiles, int64_t* device_bytes, double* amplification) {
    const Ooc& o = *s->ooc;
    if (tile)
        for (int a = 0; a < 3; ++a) tile[a] = o.tile[a];
    if (ntiles) *ntiles = (int64_t)o.tiles.size();
    if (device_bytes) *device_bytes = o.device_bytes;
    if (amplification) *amplification = o.amplification;
    return PG_OK;
}

int ooc_tile(pg_solver* s, int64_t k, int32_t t0[3], int32_t tn[3], int32_t wo[3], int32_t wl[3]) {
    const Ooc& o = *s->ooc;
    if (k < 0 || k >= (int64_t)o.tiles.size()) return fail(PG_ERR_ARG, "tile %lld out of range", (long long)k);
    const OocTile& t = o.tiles[(size_t)k];
    for (int a = 0; a < 3; ++a) {
        t0[a] = t.t0[a];
        tn[a] = t.tn[a];
        wo[a] = t.wo[a];
        wl[a] = t.wl[a];
    }
    return PG_OK;
}

}  // namespace pgh
