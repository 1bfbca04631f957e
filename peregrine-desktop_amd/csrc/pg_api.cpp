// C-ABI implementation of libperegrine_hip (declared in include/peregrine_hip.h).
//
// Host side of the solver: argument validation, the per-solve constant
// block, device allocations, the Jacobi buffer pair and the sweep loop. The
// command saturation and the first three terms of each acceleration row
// (B * u_sat, solver/src/PeregrineSolver.jl:139-161) are folded here once per
// control in float, in the reference's order, so that the kernels start each
// row's sum from the identical partial value.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "pg_internal.hpp"
#include "pg_tanh.hpp"

namespace pgh {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int64_t grid_cells(const pg_grid& g) {
    int64_t n = 1;
    for (int k = 0; k < 6; ++k) n *= (int64_t)g.n[k];
    return n;
}

}  // namespace pgh

using pgh::fail;
using pgh::g_err;
using pgh::grid_cells;

namespace {

// traction_gain (PeregrineSolver.jl:126-136), the device's own definition
float traction_gain(float u1, float u2, float u3, float knee) {
    return pgt::traction_gain_f(u1, u2, u3, knee);
}

// substeps_for (PeregrineSolver.jl:985-989)
int32_t substeps_for(float tau, const pg_params& p) {
    if (p.hmax <= 0.0f) return p.nsub;
    float q = tau / p.hmax;
    if (!(q < 1.0e4f)) q = 1.0e4f;
    int32_t k = (int32_t)q + 1;
    if (k > 64) k = 64;
    if (k < 1) k = 1;
    return k;
}

// traction_gain and substeps_for carried in double, as the FP64 tables are.
double traction_gain64(double u1, double u2, double u3, float knee) {
    if (std::isnan(knee) || knee <= 0.0f) return 1.0;
    const double m = std::sqrt(u1 * u1 + u2 * u2 + u3 * u3);
    const double r = m / (double)knee;
    if (r < (double)1.0e-6f) return 1.0;
    return std::tanh(r) / r;
}

int32_t substeps_for64(double tau, const pg_params& p) {
    if (p.hmax <= 0.0f) return p.nsub;
    double q = tau / (double)p.hmax;
    if (!(q < (double)1.0e4f)) q = (double)1.0e4f;
    int32_t k = (int32_t)q + 1;
    if (k > 64) k = 64;
    if (k < 1) k = 1;
    return k;
}

}  // namespace

namespace {

using pgh::kLdsMaxCtl;

void release(pg_solver* s) {
    if (!s) return;
    pgh::ooc_release(s);
    (void)hipFree(s->dV[0]);
    (void)hipFree(s->dV[1]);
    (void)hipFree(s->d_V16);
    (void)hipFree(s->d_tp);
    (void)hipFree(s->d_bu64);
    (void)hipFree(s->dV64[0]);
    (void)hipFree(s->dV64[1]);
    (void)hipFree(s->d_seedmask);
    (void)hipFree(s->d_tstat);
    (void)hipFree(s->d_trefit);
    (void)hipFree(s->d_occ);
    (void)hipFree(s->d_bu);
    (void)hipFree(s->d_uraw);
    (void)hipFree(s->d_wpol);
    (void)hipFree(s->d_seeds);
    (void)hipFree(s->d_part);
    (void)hipFree(s->d_sum);
    (void)hipFree(s->d_linf);
    (void)hipFree(s->d_pol);
    (void)hipFree(s->d_worig);
    (void)hipFree(s->d_wext);
    (void)hipFree(s->d_pfree);
    (void)hipFree(s->d_forig);
    (void)hipFree(s->d_fext);
    (void)hipFree(s->d_ffree);
    (void)hipFree(s->d_worig_esc);
    (void)hipFree(s->d_occ0);
    (void)hipFree(s->d_act[0]);
    (void)hipFree(s->d_act[1]);
    (void)hipFree(s->d_nwork);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    delete s;
}

int check_grid(const pg_grid* g) {
    for (int k = 0; k < 6; ++k) {
        if (g->n[k] < 2) return fail(PG_ERR_ARG, "axis %d needs at least 2 samples, got %d", k, g->n[k]);
        if (g->n[k] > 32000) return fail(PG_ERR_ARG, "axis %d has %d samples (at most 32000)", k, g->n[k]);
        if (!(g->step[k] > 0.0f) || !std::isfinite(g->step[k]) || !std::isfinite(g->lo[k]))
            return fail(PG_ERR_ARG, "axis %d has a non-positive or non-finite step", k);
    }
    return PG_OK;
}

int check_params(const pg_params* p, int32_t nctl) {
    if (p->cfl > 0.0f && !(p->tau_min > 0.0f && p->tau_min <= p->tau_max))
        return fail(PG_ERR_ARG, "cfl > 0 needs 0 < tau_min <= tau_max");
    if (p->ntau < 1 || p->ntau > 16) return fail(PG_ERR_ARG, "tau_levels must be in [1, 16], got %d", p->ntau);
    if (p->rounds < 0 || p->rounds > 16) return fail(PG_ERR_ARG, "refine_rounds must be in [0, 16], got %d", p->rounds);
    if (p->ntau > 1 && !(p->tau_ratio > 0.0f && std::isfinite(p->tau_ratio)))
        return fail(PG_ERR_ARG, "tau_ratio must be positive");
    if (p->ntau > 1 && !(p->tau_min > 0.0f && p->tau_min <= p->tau_max))
        return fail(PG_ERR_ARG, "tau_levels > 1 needs 0 < tau_min <= tau_max");
    if (p->rounds > 0 && !std::isfinite(p->delta0)) return fail(PG_ERR_ARG, "refine_delta must be finite");
    if (p->checks < 1 || p->checks > 24) return fail(PG_ERR_ARG, "sweep_checks must be in [1, 24], got %d", p->checks);
    if (!(p->dt > 0.0f) || !std::isfinite(p->dt)) return fail(PG_ERR_ARG, "dt must be positive");
    if (!(p->cap > 0.0f) || !std::isfinite(p->cap)) return fail(PG_ERR_ARG, "value_cap must be positive");
    if (p->hmax <= 0.0f && p->nsub < 1) return fail(PG_ERR_ARG, "substeps must be >= 1");
    return PG_OK;
}

void build_consts(pg_solver* s) {
    pgk::Consts& K = s->K;
    std::memset(&K, 0, sizeof K);
    const pg_grid& g = s->g;
    for (int k = 0; k < 6; ++k) {
        K.n[k] = g.n[k];
        K.lo[k] = g.lo[k];
        K.step[k] = g.step[k];
        K.rstep[k] = 1.0 / (double)g.step[k];
        K.lim[k] = (float)(g.n[k] - 1);
        s->F.rstep[k] = 1.0f / g.step[k];
    }
    K.n3f = (float)g.n[2];
    K.lim[2] = K.n3f;
    K.rn3 = 1.0 / (double)K.n3f;
    s->F.rn3 = 1.0f / K.n3f;
    K.ncell = s->ncell;
    int64_t st = 1;
    for (int k = 5; k >= 0; --k) { K.stride[k] = st; st *= (int64_t)g.n[k]; }
    std::memcpy(K.A, s->m.A, sizeof K.A);
    std::memcpy(K.q, s->m.q, sizeof K.q);
    std::memcpy(K.S, s->m.S, sizeof K.S);
    std::memcpy(K.D, s->m.D, sizeof K.D);
    std::memcpy(K.c, s->m.c, sizeof K.c);
    for (int k = 0; k < 3; ++k) {
        K.reps[k] = 1.0 / (double)s->m.eps[k];
        s->F.reps[k] = 1.0f / s->m.eps[k];
    }
    // plain scheme: cfl <= 0, so the horizon is dt (PeregrineSolver.jl:955)
    K.tau = s->p.dt;
    K.nsub = substeps_for(K.tau, s->p);
    K.hs = K.tau / (float)K.nsub;
    K.cap = s->p.cap;
    K.nchk = s->p.checks;
    for (int sc = 1; sc <= K.nchk; ++sc) K.probe_a[sc - 1] = (float)sc / (float)K.nchk;
    K.nctl = s->nctl;
    K.rk2 = s->p.rk2 ? 1 : 0;
    // Kuhn lookups take precedence over nearest-node ones (value_at,
    // PeregrineSolver.jl:1141-1146)
    K.simplex = s->p.simplex ? 1 : 0;
    K.nearest = (s->p.nearest && !s->p.simplex) ? 1 : 0;
    // per-cell horizon: the bound from tau_max and the acceleration rows,
    // folded here in float as cfl_tau does (PeregrineSolver.jl:954-976)
    K.cfl = s->p.cfl > 0.0f ? 1 : 0;
    {
        const float* B = s->m.B;
        auto amax3 = [](float a, float b, float c) {
            const float ab = a > b ? a : b;
            return ab > c ? ab : c;
        };
        const float ax = amax3(std::fabs(B[0]), std::fabs(B[1]), std::fabs(B[2]));
        const float ay = amax3(std::fabs(B[3]), std::fabs(B[4]), std::fabs(B[5]));
        const float aw = amax3(std::fabs(B[6]), std::fabs(B[7]), std::fabs(B[8]));
        float t = s->p.tau_max;
        const float cfl = s->p.cfl;
        if (ax > 1.0e-6f) { const float c = cfl * g.step[3] / ax; t = t < c ? t : c; }
        if (ay > 1.0e-6f) { const float c = cfl * g.step[4] / ay; t = t < c ? t : c; }
        if (aw > 1.0e-6f) { const float c = cfl * g.step[5] / aw; t = t < c ? t : c; }
        K.cfl_tb = t;
        K.cfl_xy = cfl * (g.step[0] < g.step[1] ? g.step[0] : g.step[1]);
        K.cfl_h = cfl * g.step[2];
        K.tau_min = s->p.tau_min;
        K.tau_max = s->p.tau_max;
        K.hmax = s->p.hmax;
    }
    std::memcpy(K.B, s->m.B, sizeof K.B);
    K.knee = s->m.knee;
    K.rounds = s->p.rounds > 0 ? s->p.rounds : 0;
    K.ntau = s->p.ntau > 1 ? s->p.ntau : 1;
    K.delta0 = s->p.delta0;
    K.tau_ratio = s->p.tau_ratio;
    K.vclamp = s->p.vclamp ? 1 : 0;
    K.adaptive = s->p.adaptive_checks ? 1 : 0;
    K.half_xy = 0.5f * (g.step[0] < g.step[1] ? g.step[0] : g.step[1]);
    K.half_h = 0.5f * g.step[2];
    K.nnc = s->nctl - 1 > 1 ? s->nctl - 1 : 1;
    K.span = s->p.ncoarse <= 0 ? K.nnc : (s->p.ncoarse < K.nnc ? s->p.ncoarse : K.nnc);
    K.off = 0;
    int64_t nt = 1;
    for (int k = 0; k < 6; ++k) {
        s->Q.nt[k] = (g.n[k] + (1 << pgk::kQuantShift) - 1) >> pgk::kQuantShift;
        nt *= s->Q.nt[k];
    }
    s->Q.ntile = nt;

    pgk::Consts64& L = s->K64;
    std::memset(&L, 0, sizeof L);
    for (int k = 0; k < 6; ++k) {
        L.n[k] = g.n[k];
        L.lo[k] = (double)g.lo[k];
        L.step[k] = (double)g.step[k];
        L.lim[k] = (double)(float)(g.n[k] - 1);
        L.stride[k] = K.stride[k];
    }
    L.n3 = (double)(float)g.n[2];
    L.lim[2] = L.n3;
    L.ncell = s->ncell;
    for (int j = 0; j < 9; ++j) {
        L.A[j] = s->m.A[j]; L.S[j] = s->m.S[j]; L.D[j] = s->m.D[j];
    }
    for (int j = 0; j < 3; ++j) {
        L.q[j] = s->m.q[j]; L.c[j] = s->m.c[j]; L.eps[j] = s->m.eps[j];
    }
    L.tau = (double)s->p.dt;
    L.nsub = substeps_for64(L.tau, s->p);
    L.hs = L.tau / (double)(float)L.nsub;
    L.cap = (double)s->p.cap;
    L.nchk = s->p.checks;
    for (int sc = 1; sc <= L.nchk; ++sc) L.probe_a[sc - 1] = (double)(float)sc / (double)(float)L.nchk;
    L.nctl = s->nctl;
}

// A model whose A, S, D blocks are diagonal and q is zero (every other
// coefficient exactly +-0) runs the kernels' diagonal specialisation.
bool model_is_diagonal(const pg_model& m) {
    for (int r = 0; r < 3; ++r) {
        if (m.q[r] != 0.0f) return false;
        for (int c = 0; c < 3; ++c) {
            if (r == c) continue;
            if (m.A[3 * r + c] != 0.0f || m.S[3 * r + c] != 0.0f || m.D[3 * r + c] != 0.0f) return false;
        }
    }
    return true;
}

void box_geometry(pg_solver* s) {
    const int want[6] = {2, 2, 2, 2, 2, 4};
    // a box must lie inside one quantization tile (its statistics are per tile)
    static_assert((1 << pgk::kQuantShift) % 4 == 0, "tile edge must be a multiple of the box edges");
    int64_t nbox = 1;
    for (int k = 0; k < 6; ++k) {
        s->G.b[k] = want[k] < s->g.n[k] ? want[k] : s->g.n[k];
        s->G.nb[k] = (s->g.n[k] + s->G.b[k] - 1) / s->G.b[k];
        s->G.nbg[k] = s->G.nb[k];
        s->G.org[k] = 0;
        s->G.rn[k] = s->g.n[k];
        nbox *= s->G.nb[k];
    }
    s->G.nbox = nbox;
}

int64_t lds_blocks(const pg_solver* s) { return pgk::box_blocks(s->G.nbox); }

}  // namespace

namespace pgh {

int ensure_partials(pg_solver* s, int64_t nb) {
    if (nb <= s->part_cap) return PG_OK;
    (void)hipFree(s->d_part);
    s->d_part = nullptr;
    HIPCHECK(hipMalloc(&s->d_part, (size_t)nb * sizeof(double)));
    s->part_cap = nb;
    return PG_OK;
}

// One-time pass over every (state, control) pair: each box's window origin,
// and the window shape that holds the most boxes' successors whole. With
// keep_host the origins and extents are also returned (the out-of-core
// tiler builds its tile windows from them).
int plan(pg_solver* s, bool keep_host, std::vector<int16_t>* h_orig, std::vector<uint8_t>* h_ext) {
    if (s->planned && !keep_host) return PG_OK;
    hipStream_t st = s->ctx->stream;
    box_geometry(s);
    const int64_t nbox = s->G.nbox;
    if (!s->d_worig) HIPCHECK(hipMalloc(&s->d_worig, (size_t)nbox * 6 * sizeof(int16_t)));
    if (!s->d_wext) HIPCHECK(hipMalloc(&s->d_wext, (size_t)nbox * 6));
    if (!s->d_pfree) HIPCHECK(hipMalloc(&s->d_pfree, (size_t)nbox));
    uint8_t* d_ext = s->d_wext;
    hipError_t e = pgk_launch_plan(s->K, nullptr, s->G, s->d_occ, s->d_bu, s->d_worig, d_ext, s->d_pfree,
                                   s->diag ? 1 : 0, st);
    std::vector<uint8_t> ext_local;
    std::vector<uint8_t>& ext = h_ext ? *h_ext : ext_local;
    ext.resize((size_t)nbox * 6);
    if (e == hipSuccess) e = hipMemcpyAsync(ext.data(), d_ext, ext.size(), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && keep_host && h_orig) {
        h_orig->resize((size_t)nbox * 6);
        e = hipMemcpyAsync(h_orig->data(), s->d_worig, h_orig->size() * sizeof(int16_t), hipMemcpyDeviceToHost, st);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(PG_ERR_HIP, "window plan failed: %s", hipGetErrorString(e));
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
        if (cov > best) { best = cov; s->shape = i; }
    }
    s->coverage = nbox > 0 ? (double)best / (double)nbox : 1.0;
    s->planned = true;
    return ensure_partials(s, lds_blocks(s));
}

// The escape sweep's window plan, once per solver: the exact plan pass over
// an empty occupancy (every state counts as free and no probe rejects), so
// each box's window holds every corner any escape backup of it reads.
int plan_escape(pg_solver* s) {
    if (s->esc_planned) return PG_OK;
    hipStream_t st = s->ctx->stream;
    box_geometry(s);
    const int64_t nbox = s->G.nbox;
    const int64_t nocc = (int64_t)s->g.n[0] * s->g.n[1] * s->g.n[2];
    if (!s->d_occ0) {
        HIPCHECK(hipMalloc(&s->d_occ0, (size_t)nocc));
        HIPCHECK(hipMemsetAsync(s->d_occ0, 0, (size_t)nocc, st));
    }
    if (!s->d_worig_esc) HIPCHECK(hipMalloc(&s->d_worig_esc, (size_t)nbox * 6 * sizeof(int16_t)));
    uint8_t* d_ext = nullptr;
    HIPCHECK(hipMalloc(&d_ext, (size_t)nbox * 6));
    hipError_t e = pgk_launch_plan(s->K, nullptr, s->G, s->d_occ0, s->d_bu, s->d_worig_esc, d_ext, nullptr,
                                   s->diag ? 1 : 0, st);
    std::vector<uint8_t> ext((size_t)nbox * 6);
    if (e == hipSuccess) e = hipMemcpyAsync(ext.data(), d_ext, ext.size(), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d_ext);
    if (e != hipSuccess) return fail(PG_ERR_HIP, "escape window plan failed: %s", hipGetErrorString(e));
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
    s->esc_planned = true;
    return ensure_partials(s, lds_blocks(s));
}

}  // namespace pgh

using pgh::ensure_partials;
using pgh::plan;
using pgh::plan_escape;

namespace {

const char* global_only_option(const pg_solver* s);

int auto_kernel(const pg_solver* s) {
    return (s->nctl <= kLdsMaxCtl && !global_only_option(s)) ? PG_KERNEL_LDS : PG_KERNEL_GLOBAL;
}

// FP32 table in dV[0] -> 16-bit mirror. A whole-table load (force) fits
// every tile afresh; after a sweep, tiles are refit only when their values
// left the representable range and only changed codes are written.
int encode_from_scratch(pg_solver* s, bool force, uint32_t* mark = nullptr) {
    hipStream_t st = s->ctx->stream;
    if (force) HIPCHECK(pgk_launch_tile_stats(s->dV[0], s->K, s->Q, s->d_tstat, st));
    else HIPCHECK(pgk_launch_seed_stats(s->d_seeds, s->nseeds, s->K, s->Q, s->d_tstat, st));
    HIPCHECK(pgk_launch_tile_refit(s->store, s->d_tstat, s->d_tp, s->d_trefit, s->Q.ntile, force ? 1 : 0, st));
    HIPCHECK(pgk_launch_encode(s->store, s->dV[0], s->K, s->Q, s->d_tp, s->d_trefit, s->d_V16, mark, s->G, st));
    return PG_OK;
}

// The current FP32 table on the device. With a 16-bit store, dV[0] holds
// every state's own value in FP32 and the codes mirror it for the gathers.
float* current_f32(pg_solver* s) { return s->f16_active ? s->dV[0] : s->dV[s->cur]; }

// Scheme options that only the exact FP32 kernels carry (the precision form,
// the FP64 tables and the adjoint integrate the plain default): a message
// naming the first one set, or null.
const char* exact_only_option(const pg_solver* s) {
    if (s->p.rk2) return "rk2";
    if (s->p.nearest) return "nearest";
    if (s->p.vclamp) return "velocity_clamp";
    if (s->p.adaptive_checks) return "adaptive_checks";
    if (s->p.simplex) return "simplex";
    if (s->p.cfl > 0.0f) return "cfl";
    if (s->K.span < s->K.nnc) return "control_scan (rotating slice)";
    if (s->p.rounds > 0) return "refine_rounds";
    if (s->p.ntau > 1) return "tau_levels";
    if (s->d_wpol) return "warm_start";
    return nullptr;
}

// Options only the global-gather kernel implements: the pattern search, the
// horizon ladder and the warm start evaluate commands and horizons that are
// not known before the sweep, so no LDS window can be planned for them.
const char* global_only_option(const pg_solver* s) {
    if (s->p.rounds > 0) return "refine_rounds";
    if (s->p.ntau > 1) return "tau_levels";
    if (s->d_wpol) return "warm_start";
    return nullptr;
}

// The scan offset of a sweep's rotating slice (PeregrineSolver.jl:1263-1275):
// (phase * span) mod nnc, 0 for the whole lattice.
int32_t slice_offset(const pg_solver* s, int32_t phase) {
    if (s->p.ncoarse <= 0) return 0;
    const int64_t nnc = s->K.nnc;
    return (int32_t)((((int64_t)phase * s->K.span) % nnc + nnc) % nnc);
}

// Consecutive quiet sweeps that end a solve: a whole turn of a rotating slice.
int32_t quiet_needed(const pg_solver* s) {
    if (s->p.ncoarse <= 0) return 1;
    const int32_t n = (s->nctl - 1 + s->p.ncoarse - 1) / s->p.ncoarse;
    return n > 1 ? n : 1;
}

// The precision form's plan (once per solver, after the exact one): its
// read ranges and the boxes whose probes it skips.
int plan_fast(pg_solver* s) {
    if (s->d_ffree) return PG_OK;
    hipStream_t st = s->ctx->stream;
    HIPCHECK(hipMalloc(&s->d_forig, (size_t)s->G.nbox * 6 * sizeof(int16_t)));
    HIPCHECK(hipMalloc(&s->d_fext, (size_t)s->G.nbox * 6));
    HIPCHECK(hipMalloc(&s->d_ffree, (size_t)s->G.nbox));
    HIPCHECK(pgk_launch_plan(s->K, &s->F, s->G, s->d_occ, s->d_bu, s->d_forig, s->d_fext, s->d_ffree,
                             s->diag ? 1 : 0, st));
    return PG_OK;
}

// The active set of the coming sweep in `regime` (see pg_solver::act_for):
// A.prev holds the previous sweep's change bits when they were made in the
// same regime, A.next the cleared bits this sweep marks. Returns false (A
// empty, every box computed) when the set is off or cannot be kept.
int active_setup(pg_solver* s, int regime, bool inplace, pgk::ActiveG& A, bool& track) {
    A = pgk::ActiveG{};
    // a rotating slice backs up with other controls each sweep: unchanged
    // inputs do not imply an unchanged result
    track = s->active_on && !inplace && regime >= 0 && s->G.nbg[5] <= 32 && s->K.span == s->K.nnc;
    if (!track) return PG_OK;
    hipStream_t st = s->ctx->stream;
    const int64_t words = s->G.nbox / s->G.nbg[5];
    if (!s->d_act[0]) {
        HIPCHECK(hipMalloc(&s->d_act[0], (size_t)words * sizeof(uint32_t)));
        HIPCHECK(hipMalloc(&s->d_act[1], (size_t)words * sizeof(uint32_t)));
        HIPCHECK(hipMalloc(&s->d_nwork, sizeof(unsigned)));
        s->act_for = -1;
    }
    if (regime != 0) {
        const int rc = plan_fast(s);
        if (rc) return rc;
    }
    uint32_t* next = s->d_act[1 - s->act_cur];
    HIPCHECK(hipMemsetAsync(next, 0, (size_t)words * sizeof(uint32_t), st));
    HIPCHECK(hipMemsetAsync(s->d_nwork, 0, sizeof(unsigned), st));
    A.prev = s->act_for == regime ? s->d_act[s->act_cur] : nullptr;
    A.next = regime == 2 ? nullptr : next;   // the 16-bit store marks in its encode pass
    A.rorig = regime == 0 ? s->d_worig : s->d_forig;
    A.rext = regime == 0 ? s->d_wext : s->d_fext;
    A.nwork = s->d_nwork;
    return PG_OK;
}

}  // namespace

extern "C" {

const char* pg_version(void) { return "peregrine_hip 0.2 (gfx950: exact FP32 Jacobi sweeps, FP16-tile precision form)"; }
int pg_abi_version(void) { return PG_ABI_VERSION; }
const char* pg_last_error(void) { return g_err.c_str(); }

int pg_device_count(int32_t* n) {
    if (!n) return fail(PG_ERR_ARG, "null output");
    int c = 0;
    HIPCHECK(hipGetDeviceCount(&c));
    *n = c;
    return PG_OK;
}

int pg_create(int32_t device, pg_ctx** out) {
    if (!out) return fail(PG_ERR_ARG, "null output");
    *out = nullptr;
    int c = 0;
    HIPCHECK(hipGetDeviceCount(&c));
    if (device < 0 || device >= c) return fail(PG_ERR_ARG, "device %d out of range (%d visible)", device, c);
    HIPCHECK(hipSetDevice(device));
    pg_ctx* ctx = new pg_ctx;
    ctx->device = device;
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return fail(PG_ERR_HIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
    }
    // this device's copy of the no-allocation scratch, looked up once here
    void* sp = nullptr;
    e = pgk_noalloc_scratch(&sp);
    if (e != hipSuccess || !sp) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return fail(PG_ERR_HIP, "scratch address lookup failed: %s", hipGetErrorString(e));
    }
    ctx->noalloc = static_cast<unsigned char*>(sp);
    *out = ctx;
    return PG_OK;
}

int pg_destroy(pg_ctx* ctx) {
    if (!ctx) return PG_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PG_OK;
}

int pg_device_budget(pg_ctx* ctx, double headroom, int64_t* bytes) {
    if (!ctx || !bytes) return fail(PG_ERR_ARG, "null argument");
    HIPCHECK(hipSetDevice(ctx->device));
    size_t fr = 0, tot = 0;
    HIPCHECK(hipMemGetInfo(&fr, &tot));
    *bytes = (int64_t)std::floor((double)fr * headroom);
    return PG_OK;
}

int pg_grid_fits(pg_ctx* ctx, const pg_grid* g, int32_t vstore, double headroom, int32_t* fits,
                 double* need_gib, double* free_gib) {
    if (!ctx || !g || !fits) return fail(PG_ERR_ARG, "null argument");
    int rc = check_grid(g);
    if (rc) return rc;
    HIPCHECK(hipSetDevice(ctx->device));
    size_t fr = 0, tot = 0;
    HIPCHECK(hipMemGetInfo(&fr, &tot));
    // FP32: the Jacobi pair; 16-bit: the codes plus the FP32 sweep scratch
    const double per_cell = (vstore == PG_VSTORE_F32) ? 8.0 : 6.0;
    const double need = (double)grid_cells(*g) * per_cell / 1073741824.0;
    const double free = (double)fr / 1073741824.0;
    *fits = need <= free * headroom ? 1 : 0;
    if (need_gib) *need_gib = need;
    if (free_gib) *free_gib = free;
    return PG_OK;
}

}  // extern "C"

namespace {

// pg_solver_create and pg_solver_create_ooc: with host_tables the Jacobi
// pair is not allocated on the device (the out-of-core driver keeps it in
// pinned host memory).
int create_solver(pg_ctx* ctx, const pg_grid* g, const pg_model* m, const pg_params* p, const float* ctl_soa,
                  int32_t nctl, const uint8_t* occ, int32_t vstore, bool host_tables, pg_solver** out) {
    if (!out) return fail(PG_ERR_ARG, "null output");
    *out = nullptr;
    if (!ctx || !g || !m || !p || !ctl_soa || !occ) return fail(PG_ERR_ARG, "null argument");
    int rc = check_grid(g);
    if (rc) return rc;
    if (nctl < 2 || nctl > 255) return fail(PG_ERR_ARG, "nctl must be in [2, 255], got %d", nctl);
    if (ctl_soa[0] != 0.0f || ctl_soa[nctl] != 0.0f || ctl_soa[2 * nctl] != 0.0f)
        return fail(PG_ERR_ARG, "control 0 must be coast (0, 0, 0)");
    rc = check_params(p, nctl);
    if (rc) return rc;
    if (vstore != PG_VSTORE_F32 && vstore != PG_VSTORE_F16_TILE && vstore != PG_VSTORE_U16_TILE)
        return fail(PG_ERR_ARG, "unknown vstore %d", vstore);
    if (vstore != PG_VSTORE_F32 &&
        (p->rk2 || p->nearest || p->vclamp || p->adaptive_checks || p->simplex || p->cfl > 0.0f ||
         (p->ncoarse > 0 && p->ncoarse < nctl - 1) || p->rounds > 0 || p->ntau > 1))
        return fail(PG_ERR_UNSUPPORTED,
                    "the 16-bit stores (precision form) run the plain default scheme only (no rk2, nearest, "
                    "velocity_clamp, adaptive_checks, simplex, cfl, control_scan slice, refine_rounds or "
                    "tau_levels)");
    if (vstore != PG_VSTORE_F32 && nctl > kLdsMaxCtl)
        return fail(PG_ERR_UNSUPPORTED, "the 16-bit stores sweep at most %d controls", kLdsMaxCtl);
    for (int k = 0; k < 3; ++k)
        if (!(m->eps[k] > 0.0f)) return fail(PG_ERR_ARG, "coulomb eps[%d] must be positive", k);
    // element-wise kernels run one work-item per cell, and a dispatch counts
    // work-items in 32 bits; larger grids are solved out of core
    if (!host_tables && grid_cells(*g) >= ((int64_t)1 << 32))
        return fail(PG_ERR_UNSUPPORTED, "%lld cells: an in-HBM solver holds fewer than 2^32 cells "
                    "(use pg_solver_create_ooc)", (long long)grid_cells(*g));
    HIPCHECK(hipSetDevice(ctx->device));

    pg_solver* s = new pg_solver;
    s->ctx = ctx;
    s->g = *g;
    s->m = *m;
    s->p = *p;
    s->nctl = nctl;
    s->vstore = vstore;
    s->ncell = grid_cells(*g);
    build_consts(s);
    s->diag = model_is_diagonal(*m);
    s->store = vstore == PG_VSTORE_F16_TILE ? PG_STORE_F16
             : vstore == PG_VSTORE_U16_TILE ? PG_STORE_U16 : PG_STORE_F32;
    s->kernel = (vstore != PG_VSTORE_F32) ? PG_KERNEL_FAST : auto_kernel(s);

    // B * u_sat per control, in the reference's summation order
    std::vector<float> bu(3 * (size_t)nctl);
    s->bu64.assign(3 * (size_t)nctl, 0.0);
    s->usat64.assign(3 * (size_t)nctl, 0.0);
    for (int k = 0; k < nctl; ++k) {
        const float u1r = ctl_soa[k], u2r = ctl_soa[k + nctl], u3r = ctl_soa[k + 2 * nctl];
        const float gn = traction_gain(u1r, u2r, u3r, m->knee);
        const float u1 = u1r * gn, u2 = u2r * gn, u3 = u3r * gn;
        for (int r = 0; r < 3; ++r) {
            float a = m->B[3 * r] * u1;
            a = a + m->B[3 * r + 1] * u2;
            a = a + m->B[3 * r + 2] * u3;
            bu[3 * k + r] = a;
        }
        // the same in double for the FP64 tables
        const double g64 = traction_gain64(u1r, u2r, u3r, m->knee);
        const double w1 = u1r * g64, w2 = u2r * g64, w3 = u3r * g64;
        s->usat64[3 * k] = w1;
        s->usat64[3 * k + 1] = w2;
        s->usat64[3 * k + 2] = w3;
        for (int r = 0; r < 3; ++r) {
            double a = (double)m->B[3 * r] * w1;
            a = a + (double)m->B[3 * r + 1] * w2;
            a = a + (double)m->B[3 * r + 2] * w3;
            s->bu64[3 * k + r] = a;
        }
    }

    const int64_t nocc = (int64_t)g->n[0] * g->n[1] * g->n[2];
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t x) { if (e == hipSuccess) e = x; };
    if (!host_tables) chk(hipMalloc(&s->dV[0], (size_t)s->ncell * sizeof(float)));
    if (vstore == PG_VSTORE_F32) {
        if (!host_tables) chk(hipMalloc(&s->dV[1], (size_t)s->ncell * sizeof(float)));
    } else {
        if (!host_tables) chk(hipMalloc(&s->d_V16, (size_t)s->ncell * sizeof(uint16_t)));
        chk(hipMalloc(&s->d_tp, (size_t)s->Q.ntile * sizeof(float2)));
        chk(hipMalloc(&s->d_tstat, (size_t)s->Q.ntile * 2 * sizeof(unsigned)));
        chk(hipMalloc(&s->d_trefit, (size_t)s->Q.ntile));
    }
    chk(hipMalloc(&s->d_occ, (size_t)nocc));
    chk(hipMalloc(&s->d_bu, bu.size() * sizeof(float)));
    chk(hipMalloc(&s->d_uraw, 3 * (size_t)nctl * sizeof(float)));
    chk(hipMalloc(&s->d_sum, (1 + kReduceBlocks) * sizeof(double)));   // the sum, then reduction scratch
    chk(hipMalloc(&s->d_linf, sizeof(unsigned)));
    if (e != hipSuccess) {
        release(s);
        return fail(e == hipErrorOutOfMemory ? PG_ERR_NOMEM : PG_ERR_HIP,
                    "device allocation for %lld cells failed: %s", (long long)grid_cells(*g),
                    hipGetErrorString(e));
    }
    std::vector<uint8_t> occ01(nocc);
    for (int64_t i = 0; i < nocc; ++i) occ01[i] = occ[i] ? 1 : 0;
    chk(hipMemcpy(s->d_occ, occ01.data(), (size_t)nocc, hipMemcpyHostToDevice));
    chk(hipMemcpy(s->d_bu, bu.data(), bu.size() * sizeof(float), hipMemcpyHostToDevice));
    {
        std::vector<float> ur(3 * (size_t)nctl);
        for (int k = 0; k < nctl; ++k)
            for (int a = 0; a < 3; ++a) ur[3 * k + a] = ctl_soa[k + a * nctl];
        chk(hipMemcpy(s->d_uraw, ur.data(), ur.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    chk(hipEventCreate(&s->ev0));
    chk(hipEventCreate(&s->ev1));
    if (vstore != PG_VSTORE_F32) {
        // tile statistics start empty and every tile unset (refit before any use)
        chk(hipMemsetAsync(s->d_tp, 0, (size_t)s->Q.ntile * sizeof(float2), ctx->stream));
        chk(hipMemsetAsync(s->d_tstat, 0, (size_t)s->Q.ntile * 2 * sizeof(unsigned), ctx->stream));
        chk(pgk_launch_tile_refit(s->store, s->d_tstat, s->d_tp, s->d_trefit, s->Q.ntile, 0, ctx->stream));
        chk(hipStreamSynchronize(ctx->stream));
    }
    if (e != hipSuccess) {
        release(s);
        return fail(PG_ERR_HIP, "solver setup failed: %s", hipGetErrorString(e));
    }
    *out = s;
    return PG_OK;
}

}  // namespace

extern "C" {

int pg_solver_create(pg_ctx* ctx, const pg_grid* g, const pg_model* m, const pg_params* p,
                     const float* ctl_soa, int32_t nctl, const uint8_t* occ, int32_t vstore,
                     pg_solver** out) {
    return create_solver(ctx, g, m, p, ctl_soa, nctl, occ, vstore, false, out);
}

int pg_solver_create_ooc(pg_ctx* ctx, const pg_grid* g, const pg_model* m, const pg_params* p,
                         const float* ctl_soa, int32_t nctl, const uint8_t* occ, int32_t vstore, int64_t budget_bytes,
                         const int32_t tile[3], pg_solver** out) {
    if (nctl > kLdsMaxCtl)
        return fail(PG_ERR_UNSUPPORTED, "the out-of-core sweep holds at most %d controls", kLdsMaxCtl);
    if (p && (p->rounds > 0 || p->ntau > 1) && vstore != PG_VSTORE_F32)
        return fail(PG_ERR_UNSUPPORTED, "the out-of-core precision form does not run refine_rounds or tau_levels "
                    "(%s set)", p->rounds > 0 ? "refine_rounds" : "tau_levels");
    if (!tile) return fail(PG_ERR_ARG, "null tile");
    int rc = create_solver(ctx, g, m, p, ctl_soa, nctl, occ, vstore, true, out);
    if (rc) return rc;
    rc = pgh::ooc_create(*out, budget_bytes, tile);
    if (rc) {
        const std::string msg = g_err;
        pg_solver_destroy(*out);
        *out = nullptr;
        g_err = msg;
    }
    return rc;
}

int pg_ooc_tile(pg_solver* s, int64_t k, int32_t* t0, int32_t* tn, int32_t* wo, int32_t* wl) {
    if (!s || !t0 || !tn || !wo || !wl) return fail(PG_ERR_ARG, "null argument");
    if (!s->ooc) return fail(PG_ERR_STATE, "not an out-of-core solver");
    return pgh::ooc_tile(s, k, t0, tn, wo, wl);
}

int pg_ooc_info(pg_solver* s, int32_t tile[3], int64_t* ntiles, int64_t* device_bytes, double* amplification) {
    if (!s) return fail(PG_ERR_ARG, "null solver");
    if (!s->ooc) return fail(PG_ERR_STATE, "not an out-of-core solver");
    return pgh::ooc_info(s, tile, ntiles, device_bytes, amplification);
}

int pg_solver_destroy(pg_solver* s) {
    if (!s) return PG_OK;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    release(s);
    return PG_OK;
}

int pg_solver_ncells(pg_solver* s, int64_t* n) {
    if (!s || !n) return fail(PG_ERR_ARG, "null argument");
    *n = s->ncell;
    return PG_OK;
}

int pg_fill_seed(pg_solver* s, const int64_t* seeds, int64_t nseeds) {
    if (!s) return fail(PG_ERR_ARG, "null solver");
    if (nseeds < 0 || (nseeds > 0 && !seeds)) return fail(PG_ERR_ARG, "bad seed list");
    for (int64_t i = 0; i < nseeds; ++i)
        if (seeds[i] < 0 || seeds[i] >= s->ncell)
            return fail(PG_ERR_ARG, "seed %lld out of range [0, %lld)", (long long)seeds[i], (long long)s->ncell);
    HIPCHECK(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    (void)hipFree(s->d_seeds);
    s->d_seeds = nullptr;
    s->nseeds = nseeds;
    s->seedmask_ok = false;
    if (nseeds > 0) {
        HIPCHECK(hipMalloc(&s->d_seeds, (size_t)nseeds * sizeof(int64_t)));
        HIPCHECK(hipMemcpyAsync(s->d_seeds, seeds, (size_t)nseeds * sizeof(int64_t), hipMemcpyHostToDevice, st));
    }
    s->h_seeds.assign(seeds, seeds + nseeds);
    s->cur = 0;
    s->act_for = -1;
    if (s->ooc) {
        HIPCHECK(hipStreamSynchronize(st));
        int rc = pgh::ooc_fill_seed(s);
        if (rc) return rc;
        s->seeded = true;
        return PG_OK;
    }
    HIPCHECK(pgk_launch_fill(s->dV[0], s->ncell, s->p.cap, st));
    HIPCHECK(pgk_launch_seed(s->dV[0], s->d_seeds, s->nseeds, st));
    if (s->d_wpol) HIPCHECK(hipMemsetAsync(s->d_wpol, 0, 3 * (size_t)s->ncell, st));
    if (s->vstore != PG_VSTORE_F32) {
        s->f16_active = true;
        s->kernel = PG_KERNEL_FAST;
        int rc = encode_from_scratch(s, true);
        if (rc) return rc;
    }
    HIPCHECK(hipStreamSynchronize(st));
    s->seeded = true;
    return PG_OK;
}

int pg_sweep(pg_solver* s, int32_t phase, int32_t rev, double* sum_decrease, float* linf) {
    if (!s) return fail(PG_ERR_ARG, "null solver");
    if (!s->seeded) return fail(PG_ERR_STATE, "pg_fill_seed (or pg_upload_V) must come before a sweep");
    HIPCHECK(hipSetDevice(s->ctx->device));
    s->K.off = slice_offset(s, phase);
    if (s->ooc) return pgh::ooc_sweep(s, sum_decrease, linf);
    hipStream_t st = s->ctx->stream;
    const int kern = s->f16_active ? PG_KERNEL_FAST : s->kernel;
    if (kern != PG_KERNEL_GLOBAL) {
        int rc = plan(s);
        if (rc == PG_OK && kern == PG_KERNEL_FAST) rc = plan_fast(s);
        if (rc) return rc;
    }
    const int64_t nb = (kern == PG_KERNEL_GLOBAL) ? (s->ncell + 255) / 256 : lds_blocks(s);
    int rc = ensure_partials(s, nb);
    if (rc) return rc;
    HIPCHECK(hipMemsetAsync(s->d_linf, 0, sizeof(unsigned), st));
    // in place: one buffer, boxes walked in alternating directions
    const bool inplace = s->sweep_mode == PG_SWEEP_IN_PLACE;
    float* Vout = s->f16_active ? s->dV[0] : (inplace ? s->dV[s->cur] : s->dV[1 - s->cur]);
    s->G.rev = (inplace && rev) ? 1 : 0;
    // boxes whose inputs did not change since they were last computed are skipped
    const int regime = s->f16_active ? 2 : (kern == PG_KERNEL_FAST ? 1 : (kern == PG_KERNEL_LDS ? 0 : -1));
    pgk::ActiveG A;
    bool track = false;
    rc = active_setup(s, regime, inplace, A, track);
    if (rc) return rc;
    uint32_t* mark = track ? s->d_act[1 - s->act_cur] : nullptr;
    int64_t nblk = nb;
    HIPCHECK(hipEventRecord(s->ev0, st));
    if (s->f16_active) {
        HIPCHECK(pgk_launch_sweep_fast(s->shape, s->diag ? 1 : 0, s->store, s->K, s->F, s->G, s->Q,
                                       s->d_V16, s->d_tp, s->dV[0], Vout, inplace ? s->d_V16 : nullptr, s->d_occ,
                                       s->d_bu, s->d_worig, s->d_ffree, nullptr, s->d_part, s->d_linf, s->d_tstat, 1,
                                       A, st));
    } else if (kern == PG_KERNEL_FAST) {
        HIPCHECK(pgk_launch_sweep_fast(s->shape, s->diag ? 1 : 0, PG_STORE_F32, s->K, s->F, s->G, s->Q,
                                       s->dV[s->cur], nullptr, s->dV[s->cur], Vout, nullptr, s->d_occ, s->d_bu,
                                       s->d_worig, s->d_ffree, nullptr, s->d_part, s->d_linf, nullptr, 1, A, st));
    } else if (kern == PG_KERNEL_LDS) {
        HIPCHECK(pgk_launch_sweep_lds(s->shape, s->diag ? 1 : 0, s->K, s->G, s->dV[s->cur], Vout, s->d_occ,
                                      s->d_bu, s->d_worig, s->d_pfree, nullptr, s->d_part, s->d_linf, 1, A, st));
    } else {
        HIPCHECK(pgk_launch_sweep(s->K, s->dV[s->cur], Vout, s->d_occ, s->d_bu, s->d_uraw, nullptr, s->d_wpol,
                                  s->d_part, s->d_linf, 1, st, &nblk));
    }
    HIPCHECK(hipEventRecord(s->ev1, st));
    HIPCHECK(pgk_launch_reduce(s->d_part, nblk, s->d_sum, s->d_sum + 1, st));
    if (mark && !s->f16_active) HIPCHECK(pgk_launch_seed_mark(Vout, s->d_seeds, s->nseeds, s->K, s->G, mark, st));
    else HIPCHECK(pgk_launch_seed(Vout, s->d_seeds, s->nseeds, st));
    if (s->f16_active) {
        rc = encode_from_scratch(s, false, mark);
        if (rc) return rc;
    } else if (!inplace) {
        s->cur = 1 - s->cur;
    }
    s->G.rev = 0;
    double hsum = 0.0;
    unsigned hl = 0, hw = 0;
    HIPCHECK(hipMemcpyAsync(&hsum, s->d_sum, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(&hl, s->d_linf, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    if (track) HIPCHECK(hipMemcpyAsync(&hw, s->d_nwork, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    if (track) {
        s->act_cur = 1 - s->act_cur;
        s->act_for = regime;
        s->last_work = hw;
    } else {
        s->act_for = -1;
        s->last_work = -1;
    }
    HIPCHECK(hipEventElapsedTime(&s->last_ms, s->ev0, s->ev1));
    float lf;
    std::memcpy(&lf, &hl, 4);
    if (sum_decrease) *sum_decrease = hsum;
    if (linf) *linf = lf;
    return PG_OK;
}

int pg_solve(pg_solver* s, int32_t max_iters, double tol_sum, float tol_linf, int32_t* iters,
             double* last_sum, float* last_linf) {
    if (!s) return fail(PG_ERR_ARG, "null solver");
    int32_t it = 0;
    double d = INFINITY;
    float l = INFINITY;
    // with a rotating slice, one quiet sweep only proves that its slice
    // helps nobody: a whole turn of the lattice must be quiet (Run.jl:1449-1453)
    const int32_t need = quiet_needed(s);
    int32_t quiet = 0;
    for (it = 1; it <= max_iters; ++it) {
        int rc = pg_sweep(s, it - 1, it & 1, &d, &l);
        if (rc) return rc;
        const bool quiet_sum = tol_sum >= 0.0 && d <= tol_sum;
        const bool quiet_linf = tol_linf > 0.0f && l < tol_linf;
        quiet = (quiet_sum || quiet_linf) ? quiet + 1 : 0;
        if (quiet >= need) break;
    }
    if (it > max_iters) it = max_iters;
    if (iters) *iters = it;
    if (last_sum) *last_sum = d;
    if (last_linf) *last_linf = l;
    return PG_OK;
}

int pg_escape_sweep(pg_solver* s, int32_t phase, int32_t rev, double* sum_improve, float* linf) {
    (void)rev;
    if (!s) return fail(PG_ERR_ARG, "null solver");
    if (!s->seeded) return fail(PG_ERR_STATE, "no value table yet");
    if (s->f16_active) return fail(PG_ERR_STATE, "a 16-bit store must be polished (pg_polish) before the escape pass");
    HIPCHECK(hipSetDevice(s->ctx->device));
    s->K.off = slice_offset(s, phase);
    if (s->ooc) return pgh::ooc_escape_sweep(s, sum_improve, linf);
    hipStream_t st = s->ctx->stream;
    const int64_t nb = (s->ncell + 255) / 256;
    int rc = ensure_partials(s, nb);
    if (rc) return rc;
    if (!s->dV[1]) {
        hipError_t e = hipMalloc(&s->dV[1], (size_t)s->ncell * sizeof(float));
        if (e != hipSuccess)
            return fail(e == hipErrorOutOfMemory ? PG_ERR_NOMEM : PG_ERR_HIP, "escape buffer allocation failed: %s",
                        hipGetErrorString(e));
    }
    // LDS-staged unless a setting needs the global-gather kernel (the
    // production candidates, more controls than the LDS table holds) or the
    // caller chose that kernel
    const bool lds = s->kernel != PG_KERNEL_GLOBAL && s->p.rounds <= 0 && s->p.ntau <= 1 && s->nctl <= kLdsMaxCtl;
    if (lds) {
        rc = plan_escape(s);
        if (rc) return rc;
    }
    HIPCHECK(hipMemsetAsync(s->d_linf, 0, sizeof(unsigned), st));
    int64_t nblk = lds ? lds_blocks(s) : nb;
    HIPCHECK(hipEventRecord(s->ev0, st));
    if (lds)
        HIPCHECK(pgk_launch_escape_lds(s->esc_shape, s->diag ? 1 : 0, s->K, s->G, s->dV[s->cur], s->dV[1 - s->cur],
                                       s->d_occ, s->d_bu, s->d_worig_esc, s->planned ? s->d_pfree : nullptr,
                                       s->d_part, s->d_linf, st));
    else
        HIPCHECK(pgk_launch_escape(s->K, s->dV[s->cur], s->dV[1 - s->cur], s->d_occ, s->d_bu, s->d_uraw, nullptr,
                                   s->d_part, s->d_linf, st, &nblk));
    HIPCHECK(hipEventRecord(s->ev1, st));
    HIPCHECK(pgk_launch_reduce(s->d_part, nblk, s->d_sum, s->d_sum + 1, st));
    s->cur = 1 - s->cur;
    s->act_for = -1;
    s->last_work = -1;
    double hsum = 0.0;
    unsigned hl = 0;
    HIPCHECK(hipMemcpyAsync(&hsum, s->d_sum, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(&hl, s->d_linf, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    HIPCHECK(hipEventElapsedTime(&s->last_ms, s->ev0, s->ev1));
    float lf;
    std::memcpy(&lf, &hl, 4);
    if (sum_improve) *sum_improve = hsum;
    if (linf) *linf = lf;
    return PG_OK;
}

int pg_solve_escape(pg_solver* s, int32_t max_iters, double tol_sum, float tol_linf, int32_t* iters,
                    double* last_sum, float* last_linf) {
    if (!s) return fail(PG_ERR_ARG, "null solver");
    int32_t it = 0;
    double d = INFINITY;
    float l = INFINITY;
    const int32_t need = quiet_needed(s);
    int32_t quiet = 0;
    for (it = 1; it <= max_iters; ++it) {
        int rc = pg_escape_sweep(s, it - 1, it & 1, &d, &l);
        if (rc) return rc;
        const bool quiet_sum = tol_sum >= 0.0 && d <= tol_sum;
        const bool quiet_linf = tol_linf > 0.0f && l < tol_linf;
        quiet = (quiet_sum || quiet_linf) ? quiet + 1 : 0;
        if (quiet >= need) break;
    }
    if (it > max_iters) it = max_iters;
    if (iters) *iters = it;
    if (last_sum) *last_sum = d;
    if (last_linf) *last_linf = l;
    return PG_OK;
}

int pg_polish(pg_solver* s, int32_t nsweeps, double* last_sum, float* last_linf) {
    if (!s) return fail(PG_ERR_ARG, "null solver");
    if (!s->seeded) return fail(PG_ERR_STATE, "no value table yet");
    if (nsweeps < 0) return fail(PG_ERR_ARG, "nsweeps must be >= 0");
    HIPCHECK(hipSetDevice(s->ctx->device));
    if (s->ooc) {
        if (s->f16_active) pgh::ooc_polish_start(s);
    } else if (s->f16_active) {
        if (!s->dV[1]) {
            hipError_t e = hipMalloc(&s->dV[1], (size_t)s->ncell * sizeof(float));
            if (e != hipSuccess)
                return fail(e == hipErrorOutOfMemory ? PG_ERR_NOMEM : PG_ERR_HIP,
                            "polish buffer allocation failed: %s", hipGetErrorString(e));
        }
        // dV[0] already holds the table in FP32; the codes are dropped
        s->cur = 0;
        s->f16_active = false;
    }
    s->kernel = auto_kernel(s);
    double d = 0.0;
    float l = 0.0f;
    for (int32_t i = 0; i < nsweeps; ++i) {
        int rc = pg_sweep(s, i, i & 1, &d, &l);
        if (rc) return rc;
    }
    if (last_sum) *last_sum = d;
    if (last_linf) *last_linf = l;
    return PG_OK;
}

int pg_policy(pg_solver* s, uint8_t* pol_host) {
    if (!s || !pol_host) return fail(PG_ERR_ARG, "null argument");
    if (s->ooc) return fail(PG_ERR_UNSUPPORTED, "pg_policy is not built for an out-of-core solver");
    if (!s->seeded) return fail(PG_ERR_STATE, "no value table yet");
    HIPCHECK(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    if (!s->d_pol) HIPCHECK(hipMalloc(&s->d_pol, (size_t)s->ncell));
    int64_t nblk = 0;
    // the argmin over the whole lattice, whatever slice the last sweep used
    // (no pattern search or ladder: the policy is a lattice index)
    const int32_t span = s->K.span, off = s->K.off, rounds = s->K.rounds, ntau = s->K.ntau;
    s->K.span = s->K.nnc;
    s->K.off = 0;
    s->K.rounds = 0;
    s->K.ntau = 1;
    struct Restore {
        pg_solver* s;
        int32_t span, off, rounds, ntau;
        ~Restore() { s->K.span = span; s->K.off = off; s->K.rounds = rounds; s->K.ntau = ntau; }
    } restore{s, span, off, rounds, ntau};
    const int kern = s->f16_active ? PG_KERNEL_FAST : s->kernel;
    if (kern != PG_KERNEL_GLOBAL) {
        int rc = plan(s);
        if (rc == PG_OK && kern == PG_KERNEL_FAST) rc = plan_fast(s);
        if (rc) return rc;
    }
    if (s->f16_active) {
        HIPCHECK(pgk_launch_sweep_fast(s->shape, s->diag ? 1 : 0, s->store, s->K, s->F, s->G, s->Q,
                                       s->d_V16, s->d_tp, s->dV[0], nullptr, nullptr, s->d_occ, s->d_bu,
                                       s->d_worig, s->d_ffree, s->d_pol, nullptr, nullptr, nullptr, 0, pgk::ActiveG{},
                                       st));
    } else if (kern == PG_KERNEL_FAST) {
        HIPCHECK(pgk_launch_sweep_fast(s->shape, s->diag ? 1 : 0, PG_STORE_F32, s->K, s->F, s->G, s->Q,
                                       s->dV[s->cur], nullptr, s->dV[s->cur], nullptr, nullptr, s->d_occ, s->d_bu,
                                       s->d_worig, s->d_ffree, s->d_pol, nullptr, nullptr, nullptr, 0, pgk::ActiveG{},
                                       st));
    } else if (kern == PG_KERNEL_LDS) {
        HIPCHECK(pgk_launch_sweep_lds(s->shape, s->diag ? 1 : 0, s->K, s->G, s->dV[s->cur], nullptr, s->d_occ,
                                      s->d_bu, s->d_worig, s->d_pfree, s->d_pol, nullptr, nullptr, 0, pgk::ActiveG{},
                                      st));
    } else {
        HIPCHECK(pgk_launch_sweep(s->K, s->dV[s->cur], nullptr, s->d_occ, s->d_bu, s->d_uraw, s->d_pol, nullptr,
                                  nullptr, nullptr, 0, st, &nblk));
    }
    HIPCHECK(hipMemcpyAsync(pol_host, s->d_pol, (size_t)s->ncell, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    return PG_OK;
}

int pg_download_V(pg_solver* s, float* dst, int64_t off, int64_t n) {
    if (!s || (!dst && n > 0)) return fail(PG_ERR_ARG, "null argument");
    if (off < 0 || n < 0 || off + n > s->ncell) return fail(PG_ERR_ARG, "range [%lld, %lld) outside the grid", (long long)off, (long long)(off + n));
    if (n == 0) return PG_OK;
    if (s->ooc) return pgh::ooc_download(s, dst, off, n);
    HIPCHECK(hipSetDevice(s->ctx->device));
    const float* src = current_f32(s);
    HIPCHECK(hipMemcpyAsync(dst, src + off, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s->ctx->stream));
    HIPCHECK(hipStreamSynchronize(s->ctx->stream));
    return PG_OK;
}

int pg_download_V16(pg_solver* s, float* dst, int64_t off, int64_t n) {
    if (!s || (!dst && n > 0)) return fail(PG_ERR_ARG, "null argument");
    if (off < 0 || n < 0 || off + n > s->ncell) return fail(PG_ERR_ARG, "range [%lld, %lld) outside the grid", (long long)off, (long long)(off + n));
    if (!s->f16_active) return fail(PG_ERR_STATE, "no 16-bit mirror is active (FP32 store, or polished)");
    if (n == 0) return PG_OK;
    HIPCHECK(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    float* tmp = nullptr;
    uint16_t* codes = nullptr;
    hipError_t e = hipMalloc(&tmp, (size_t)n * sizeof(float));
    if (e == hipSuccess && s->ooc) e = hipMalloc(&codes, (size_t)n * sizeof(uint16_t));
    if (e != hipSuccess) {
        (void)hipFree(tmp);
        return fail(PG_ERR_NOMEM, "decode buffer allocation failed: %s", hipGetErrorString(e));
    }
    // out of core the codes are on the host: cells [off, off + n) go to a
    // device buffer addressed from cell off
    if (s->ooc)
        e = hipMemcpyAsync(codes, pgh::ooc_host_codes(s) + off, (size_t)n * sizeof(uint16_t), hipMemcpyHostToDevice,
                           st);
    if (e == hipSuccess)
        e = pgk_launch_decode(s->store, s->ooc ? codes - off : s->d_V16, s->K, s->Q, s->d_tp, off, n, tmp, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dst, tmp, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(tmp);
    (void)hipFree(codes);
    if (e != hipSuccess) return fail(PG_ERR_HIP, "decode failed: %s", hipGetErrorString(e));
    return PG_OK;
}

int pg_upload_V(pg_solver* s, const float* src, int64_t off, int64_t n) {
    if (!s || (!src && n > 0)) return fail(PG_ERR_ARG, "null argument");
    if (off < 0 || n < 0 || off + n > s->ncell) return fail(PG_ERR_ARG, "range [%lld, %lld) outside the grid", (long long)off, (long long)(off + n));
    if (s->ooc) {
        int rc = pgh::ooc_upload(s, src, off, n);
        if (rc == PG_OK) s->seeded = true;
        return rc;
    }
    HIPCHECK(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    s->act_for = -1;
    if (s->vstore != PG_VSTORE_F32 && (s->f16_active || !s->seeded)) {
        if (!s->seeded && (off > 0 || n < s->ncell)) {
            HIPCHECK(pgk_launch_fill(s->dV[0], s->ncell, s->p.cap, st));
        }
        if (n > 0) HIPCHECK(hipMemcpyAsync(s->dV[0] + off, src, (size_t)n * sizeof(float), hipMemcpyHostToDevice, st));
        s->f16_active = true;
        s->kernel = PG_KERNEL_FAST;
        int rc = encode_from_scratch(s, true);
        if (rc) return rc;
    } else if (n > 0) {
        HIPCHECK(hipMemcpyAsync(s->dV[s->cur] + off, src, (size_t)n * sizeof(float), hipMemcpyHostToDevice, st));
    }
    HIPCHECK(hipStreamSynchronize(st));
    s->seeded = true;
    return PG_OK;
}

// ---------------------------------------------------------------------------
// FP64 tables
// ---------------------------------------------------------------------------

namespace {

int f64_setup(pg_solver* s) {
    if (s->ooc) return fail(PG_ERR_UNSUPPORTED, "the FP64 tables are not built for an out-of-core solver");
    if (const char* opt = exact_only_option(s))
        return fail(PG_ERR_UNSUPPORTED, "the FP64 tables and the adjoint run the plain default scheme only (%s set)",
                    opt);
    if (s->nctl > kLdsMaxCtl)
        return fail(PG_ERR_UNSUPPORTED, "the FP64 sweeps hold at most %d controls", kLdsMaxCtl);
    int rc = plan(s);
    if (rc) return rc;
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t x) { if (e == hipSuccess) e = x; };
    if (!s->d_bu64) {
        chk(hipMalloc(&s->d_bu64, s->bu64.size() * sizeof(double)));
        chk(hipMemcpy(s->d_bu64, s->bu64.data(), s->bu64.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    for (int j = 0; j < 2; ++j)
        if (!s->dV64[j]) chk(hipMalloc(&s->dV64[j], (size_t)s->ncell * sizeof(double)));
    const size_t mask_words = (size_t)((s->ncell + 31) / 32);
    if (!s->d_seedmask) chk(hipMalloc(&s->d_seedmask, mask_words * sizeof(uint32_t)));
    if (e == hipSuccess && !s->seedmask_ok) {
        chk(hipMemsetAsync(s->d_seedmask, 0, mask_words * sizeof(uint32_t), s->ctx->stream));
        chk(pgk_launch_seed_mask(s->d_seedmask, s->d_seeds, s->nseeds, s->ctx->stream));
        if (e == hipSuccess) s->seedmask_ok = true;
    }
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? PG_ERR_NOMEM : PG_ERR_HIP, "FP64 table allocation failed: %s",
                    hipGetErrorString(e));
    return PG_OK;
}

// One FP64 sweep (or residual pass) of Vin; results in *linf / *sum.
int f64_pass(pg_solver* s, int mode, const double* Vin, double* Vout, uint8_t* pol, float* linf, double* sum) {
    hipStream_t st = s->ctx->stream;
    HIPCHECK(hipMemsetAsync(s->d_linf, 0, sizeof(unsigned), st));
    HIPCHECK(pgk_launch_sweep_f64(mode, s->shape, s->K64, s->G, Vin, Vout, s->d_occ, s->d_bu64, s->d_worig,
                                  s->d_seedmask, pol, s->d_part, s->d_linf, st));
    HIPCHECK(pgk_launch_reduce(s->d_part, lds_blocks(s), s->d_sum, s->d_sum + 1, st));
    unsigned hl = 0;
    double hs = 0.0;
    HIPCHECK(hipMemcpyAsync(&hs, s->d_sum, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(&hl, s->d_linf, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    std::memcpy(linf, &hl, 4);
    *sum = hs;
    return PG_OK;
}

}  // namespace

int pg_solve_f64(pg_solver* s, int32_t mode, int32_t max_iters, double tol_linf, int32_t* iters,
                 double* last_linf) {
    if (!s) return fail(PG_ERR_ARG, "null solver");
    if (mode != PG_F64_SWEEP_PLAIN && mode != PG_F64_SWEEP_DIAGONAL) return fail(PG_ERR_ARG, "unknown FP64 sweep mode %d", mode);
    if (!s->seeded || (s->nseeds > 0 && !s->d_seeds)) return fail(PG_ERR_STATE, "pg_fill_seed must come first");
    HIPCHECK(hipSetDevice(s->ctx->device));
    int rc = f64_setup(s);
    if (rc) return rc;
    hipStream_t st = s->ctx->stream;
    HIPCHECK(pgk_launch_fill_f64(s->dV64[0], s->ncell, s->K64.cap, st));
    HIPCHECK(pgk_launch_seed_f64(s->dV64[0], s->d_seeds, s->nseeds, st));
    s->cur64 = 0;
    s->f64_valid = true;
    const int kmode = mode == PG_F64_SWEEP_DIAGONAL ? PG_F64_DIAGONAL : PG_F64_PLAIN;
    float l = INFINITY;
    double d = 0.0;
    int32_t it = 0;
    for (it = 1; it <= max_iters; ++it) {
        double* in = s->dV64[s->cur64];
        double* out = s->dV64[1 - s->cur64];
        rc = f64_pass(s, kmode, in, out, nullptr, &l, &d);
        if (rc) return rc;
        HIPCHECK(pgk_launch_seed_f64(out, s->d_seeds, s->nseeds, st));
        s->cur64 = 1 - s->cur64;
        if ((double)l < tol_linf) break;
    }
    if (it > max_iters) it = max_iters;
    if (iters) *iters = it;
    if (last_linf) *last_linf = l;
    return PG_OK;
}

int pg_sweep_f64(pg_solver* s, int32_t mode, double* linf, double* sum_decrease) {
    if (!s) return fail(PG_ERR_ARG, "null solver");
    if (mode != PG_F64_SWEEP_PLAIN && mode != PG_F64_SWEEP_DIAGONAL) return fail(PG_ERR_ARG, "unknown FP64 sweep mode %d", mode);
    if (!s->f64_valid) return fail(PG_ERR_STATE, "no FP64 table yet (pg_solve_f64)");
    HIPCHECK(hipSetDevice(s->ctx->device));
    int rc = f64_setup(s);
    if (rc) return rc;
    float l = 0.0f;
    double d = 0.0;
    double* out = s->dV64[1 - s->cur64];
    rc = f64_pass(s, mode == PG_F64_SWEEP_DIAGONAL ? PG_F64_DIAGONAL : PG_F64_PLAIN, s->dV64[s->cur64], out,
                  nullptr, &l, &d);
    if (rc) return rc;
    HIPCHECK(pgk_launch_seed_f64(out, s->d_seeds, s->nseeds, s->ctx->stream));
    s->cur64 = 1 - s->cur64;
    if (linf) *linf = l;
    if (sum_decrease) *sum_decrease = d;
    return PG_OK;
}

int pg_residual_f64(pg_solver* s, int32_t which, double* linf, double* sum) {
    if (!s) return fail(PG_ERR_ARG, "null solver");
    if (which != 0 && which != 1) return fail(PG_ERR_ARG, "which must be 0 (FP64 table) or 1 (current table)");
    if (which == 0 && !s->f64_valid) return fail(PG_ERR_STATE, "no FP64 table yet (pg_solve_f64)");
    if (which == 1 && !s->seeded) return fail(PG_ERR_STATE, "no value table yet");
    HIPCHECK(hipSetDevice(s->ctx->device));
    int rc = f64_setup(s);
    if (rc) return rc;
    const double* V = s->dV64[s->cur64];
    if (which == 1) {
        double* tmp = s->dV64[s->f64_valid ? 1 - s->cur64 : 0];
        HIPCHECK(pgk_launch_widen(current_f32(s), tmp, s->ncell, s->ctx->stream));
        V = tmp;
    }
    float l = 0.0f;
    double d = 0.0;
    rc = f64_pass(s, PG_F64_RESIDUAL, V, nullptr, nullptr, &l, &d);
    if (rc) return rc;
    if (linf) *linf = l;
    if (sum) *sum = d;
    return PG_OK;
}

int pg_download_V64(pg_solver* s, double* dst, int64_t off, int64_t n) {
    if (!s || (!dst && n > 0)) return fail(PG_ERR_ARG, "null argument");
    if (off < 0 || n < 0 || off + n > s->ncell) return fail(PG_ERR_ARG, "range [%lld, %lld) outside the grid", (long long)off, (long long)(off + n));
    if (!s->f64_valid) return fail(PG_ERR_STATE, "no FP64 table yet (pg_solve_f64)");
    if (n == 0) return PG_OK;
    HIPCHECK(hipSetDevice(s->ctx->device));
    HIPCHECK(hipMemcpyAsync(dst, s->dV64[s->cur64] + off, (size_t)n * sizeof(double), hipMemcpyDeviceToHost,
                            s->ctx->stream));
    HIPCHECK(hipStreamSynchronize(s->ctx->stream));
    return PG_OK;
}

int pg_policy_f64(pg_solver* s, uint8_t* pol_host) {
    if (!s || !pol_host) return fail(PG_ERR_ARG, "null argument");
    if (!s->f64_valid) return fail(PG_ERR_STATE, "no FP64 table yet (pg_solve_f64)");
    HIPCHECK(hipSetDevice(s->ctx->device));
    int rc = f64_setup(s);
    if (rc) return rc;
    if (!s->d_pol) HIPCHECK(hipMalloc(&s->d_pol, (size_t)s->ncell));
    float l = 0.0f;
    double d = 0.0;
    rc = f64_pass(s, PG_F64_RESIDUAL, s->dV64[s->cur64], nullptr, s->d_pol, &l, &d);
    if (rc) return rc;
    HIPCHECK(hipMemcpy(pol_host, s->d_pol, (size_t)s->ncell, hipMemcpyDeviceToHost));
    return PG_OK;
}

int pg_adjoint(pg_solver* s, const double* starts, int64_t nstarts, int32_t max_iters, double tol,
               double* dL_dtheta, double* L, int32_t* iters) {
    if (!s || !dL_dtheta || !L || (!starts && nstarts > 0)) return fail(PG_ERR_ARG, "null argument");
    if (nstarts <= 0) return fail(PG_ERR_ARG, "nstarts must be positive");
    if (!s->f64_valid) return fail(PG_ERR_STATE, "no FP64 table yet (pg_solve_f64)");
    HIPCHECK(hipSetDevice(s->ctx->device));
    int rc = f64_setup(s);
    if (rc) return rc;
    hipStream_t st = s->ctx->stream;
    const int64_t n = s->ncell;
    const double* V = s->dV64[s->cur64];
    if (!s->d_pol) HIPCHECK(hipMalloc(&s->d_pol, (size_t)n));
    float l32 = 0.0f;
    double d = 0.0;
    rc = f64_pass(s, PG_F64_RESIDUAL, V, nullptr, s->d_pol, &l32, &d);   // the FP64 argmin
    if (rc) return rc;
    const int64_t nbg = pgk_adj_grad_blocks(s->K64);
    double *tcoord = nullptr, *self = nullptr, *g = nullptr, *lam = nullptr, *nxt = nullptr, *dst = nullptr,
           *vst = nullptr, *part = nullptr, *usat = nullptr;
    int16_t* bcoord = nullptr;
    unsigned long long* stat = nullptr;
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t x) { if (e == hipSuccess) e = x; };
    chk(hipMalloc(&tcoord, (size_t)n * 6 * sizeof(double)));
    chk(hipMalloc(&bcoord, (size_t)n * 6 * sizeof(int16_t)));
    chk(hipMalloc(&self, (size_t)n * sizeof(double)));
    chk(hipMalloc(&g, (size_t)n * sizeof(double)));
    chk(hipMalloc(&lam, (size_t)n * sizeof(double)));
    chk(hipMalloc(&nxt, (size_t)n * sizeof(double)));
    chk(hipMalloc(&dst, (size_t)nstarts * 6 * sizeof(double)));
    chk(hipMalloc(&vst, (size_t)nstarts * sizeof(double)));
    chk(hipMalloc(&part, (size_t)nbg * 30 * sizeof(double)));
    chk(hipMalloc(&usat, s->usat64.size() * sizeof(double)));
    chk(hipMalloc(&stat, 2 * sizeof(unsigned long long)));
    int32_t it = 0;
    bool done = false;
    std::vector<double> vstart((size_t)nstarts), parts((size_t)nbg * 30);
    if (e == hipSuccess) {
        chk(hipMemcpyAsync(dst, starts, (size_t)nstarts * 6 * sizeof(double), hipMemcpyHostToDevice, st));
        chk(hipMemcpyAsync(usat, s->usat64.data(), s->usat64.size() * sizeof(double), hipMemcpyHostToDevice, st));
        chk(hipMemsetAsync(g, 0, (size_t)n * sizeof(double), st));
        chk(hipMemsetAsync(nxt, 0, (size_t)n * sizeof(double), st));
        chk(pgk_launch_adj_linearize(V, s->d_pol, s->d_seedmask, s->d_occ, s->d_bu64, s->K64, tcoord, bcoord, self, st));
        chk(pgk_launch_adj_starts(V, dst, nstarts, s->K64, g, vst, st));
        chk(hipMemcpyAsync(lam, g, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, st));
        for (it = 1; it <= max_iters && e == hipSuccess; ++it) {
            unsigned long long h[2] = {0, 0};
            chk(hipMemsetAsync(stat, 0, sizeof h, st));
            chk(pgk_launch_adj_iteration(lam, nxt, g, self, tcoord, bcoord, s->K64, stat, st));
            chk(hipMemcpyAsync(h, stat, sizeof h, hipMemcpyDeviceToHost, st));
            chk(hipStreamSynchronize(st));
            double dch, mag;
            std::memcpy(&dch, &h[0], 8);
            std::memcpy(&mag, &h[1], 8);
            if (dch <= tol * mag) { done = true; break; }
        }
        chk(pgk_launch_adj_grad(V, lam, s->d_pol, tcoord, s->d_occ, s->d_bu64, usat, s->K64, part, st));
        chk(hipMemcpyAsync(parts.data(), part, parts.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        chk(hipMemcpyAsync(vstart.data(), vst, vstart.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        chk(hipStreamSynchronize(st));
    }
    (void)hipFree(tcoord); (void)hipFree(bcoord); (void)hipFree(self); (void)hipFree(g); (void)hipFree(lam);
    (void)hipFree(nxt); (void)hipFree(dst); (void)hipFree(vst); (void)hipFree(part); (void)hipFree(usat);
    (void)hipFree(stat);
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? PG_ERR_NOMEM : PG_ERR_HIP, "adjoint failed: %s", hipGetErrorString(e));
    for (int j = 0; j < 30; ++j) dL_dtheta[j] = 0.0;
    for (int64_t b = 0; b < nbg; ++b)
        for (int j = 0; j < 30; ++j) dL_dtheta[j] += parts[(size_t)b * 30 + j];
    double ls = 0.0;
    for (double v : vstart) ls += v;
    *L = ls / (double)nstarts;
    if (iters) *iters = done ? (it > max_iters ? max_iters : it) : -1;
    return PG_OK;
}

int pg_release_f64(pg_solver* s) {
    if (!s) return fail(PG_ERR_ARG, "null solver");
    HIPCHECK(hipSetDevice(s->ctx->device));
    for (int j = 0; j < 2; ++j) {
        (void)hipFree(s->dV64[j]);
        s->dV64[j] = nullptr;
    }
    s->f64_valid = false;
    return PG_OK;
}

int pg_solver_set_sweep_mode(pg_solver* s, int32_t mode) {
    if (!s) return fail(PG_ERR_ARG, "null solver");
    if (mode != PG_SWEEP_JACOBI && mode != PG_SWEEP_IN_PLACE) return fail(PG_ERR_ARG, "unknown sweep mode %d", mode);
    if (s->ooc && mode != PG_SWEEP_JACOBI)
        return fail(PG_ERR_UNSUPPORTED, "an out-of-core solver sweeps Jacobi (its tiles read the old table)");
    s->sweep_mode = mode;
    return PG_OK;
}

int pg_solver_set_kernel(pg_solver* s, int32_t kernel) {
    if (!s) return fail(PG_ERR_ARG, "null solver");
    if (kernel == PG_KERNEL_AUTO) {
        s->kernel = s->f16_active ? PG_KERNEL_FAST : auto_kernel(s);
        return PG_OK;
    }
    if (kernel != PG_KERNEL_GLOBAL && kernel != PG_KERNEL_LDS && kernel != PG_KERNEL_FAST)
        return fail(PG_ERR_ARG, "unknown kernel %d", kernel);
    if (kernel != PG_KERNEL_GLOBAL && s->nctl > kLdsMaxCtl)
        return fail(PG_ERR_UNSUPPORTED, "the LDS-staged sweeps hold at most %d controls", kLdsMaxCtl);
    if (kernel != PG_KERNEL_GLOBAL && global_only_option(s))
        return fail(PG_ERR_UNSUPPORTED, "%s runs on the global-gather kernel only", global_only_option(s));
    if (kernel == PG_KERNEL_FAST && exact_only_option(s))
        return fail(PG_ERR_UNSUPPORTED, "the precision-form kernel runs the plain default scheme only (%s set)",
                    exact_only_option(s));
    if (s->f16_active && kernel != PG_KERNEL_FAST)
        return fail(PG_ERR_STATE, "a 16-bit store sweeps with the fast kernel until pg_polish");
    if (s->ooc && kernel != PG_KERNEL_LDS)
        return fail(PG_ERR_UNSUPPORTED, "an out-of-core solver sweeps with the LDS-staged kernel");
    s->kernel = kernel;
    return PG_OK;
}

int pg_solver_kernel_info(pg_solver* s, int32_t* kernel, double* window_coverage) {
    if (!s) return fail(PG_ERR_ARG, "null solver");
    const int kern = s->f16_active ? PG_KERNEL_FAST : s->kernel;
    if (kernel) *kernel = kern;
    if (window_coverage) *window_coverage = (kern != PG_KERNEL_GLOBAL && s->planned) ? s->coverage : 0.0;
    return PG_OK;
}

int pg_solver_probe_info(pg_solver* s, double* chords_free, double* probes_pass) {
    if (!s) return fail(PG_ERR_ARG, "null solver");
    double f1 = 0.0, f2 = 0.0;
    if (s->planned && s->d_pfree && s->G.nbox > 0) {
        std::vector<uint8_t> h((size_t)s->G.nbox);
        hipStream_t st = s->ctx->stream;
        HIPCHECK(hipMemcpyAsync(h.data(), s->d_pfree, h.size(), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        int64_t n1 = 0, n2 = 0;
        for (uint8_t v : h) {
            n1 += v & 1;
            n2 += (v >> 1) & 1;
        }
        f1 = (double)n1 / (double)h.size();
        f2 = (double)n2 / (double)h.size();
    }
    if (chords_free) *chords_free = f1;
    if (probes_pass) *probes_pass = f2;
    return PG_OK;
}

int pg_last_sweep_ms(pg_solver* s, float* ms) {
    if (!s || !ms) return fail(PG_ERR_ARG, "null argument");
    *ms = s->last_ms;
    return PG_OK;
}

int pg_sweep_noalloc(pg_ctx* ctx, const pg_grid* g, const pg_model* m, const pg_params* p,
                     const float* ctl_soa, int32_t nctl, const uint8_t* occ_dev, const float* Vin_dev,
                     float* Vout_dev, int32_t phase, double* sum_decrease, float* linf) {
    if (!ctx || !g || !m || !p || !ctl_soa || !occ_dev || !Vin_dev || !Vout_dev)
        return fail(PG_ERR_ARG, "null argument");
    if (Vin_dev == Vout_dev) return fail(PG_ERR_ARG, "Vin_dev and Vout_dev must be distinct (a Jacobi sweep)");
    int rc = check_grid(g);
    if (rc) return rc;
    if (nctl < 2 || nctl > 255) return fail(PG_ERR_ARG, "nctl must be in [2, 255], got %d", nctl);
    if (ctl_soa[0] != 0.0f || ctl_soa[nctl] != 0.0f || ctl_soa[2 * nctl] != 0.0f)
        return fail(PG_ERR_ARG, "control 0 must be coast (0, 0, 0)");
    rc = check_params(p, nctl);
    if (rc) return rc;
    for (int k = 0; k < 3; ++k)
        if (!(m->eps[k] > 0.0f)) return fail(PG_ERR_ARG, "coulomb eps[%d] must be positive", k);
    const int64_t ncell = grid_cells(*g);
    if (ncell > PG_NOALLOC_MAX_CELLS)
        return fail(PG_ERR_ARG, "%lld cells: the no-allocation sweep's scratch holds the partial sums of at most "
                    "%lld cells", (long long)ncell, (long long)PG_NOALLOC_MAX_CELLS);
    if (!ctx->noalloc) return fail(PG_ERR_STATE, "the context has no scratch address");
    HIPCHECK(hipSetDevice(ctx->device));

    // the constants a solver of this grid would hold; no device memory
    pg_solver t;
    t.ctx = ctx;
    t.g = *g;
    t.m = *m;
    t.p = *p;
    t.nctl = nctl;
    t.ncell = ncell;
    build_consts(&t);
    t.K.off = slice_offset(&t, phase);
    // B * u_sat and the raw commands, as pg_solver_create folds them
    float bu[3 * 255], ur[3 * 255];
    for (int k = 0; k < nctl; ++k) {
        const float u1r = ctl_soa[k], u2r = ctl_soa[k + nctl], u3r = ctl_soa[k + 2 * nctl];
        const float gn = traction_gain(u1r, u2r, u3r, m->knee);
        const float u1 = u1r * gn, u2 = u2r * gn, u3 = u3r * gn;
        for (int r = 0; r < 3; ++r) {
            float a = m->B[3 * r] * u1;
            a = a + m->B[3 * r + 1] * u2;
            a = a + m->B[3 * r + 2] * u3;
            bu[3 * k + r] = a;
        }
        for (int a = 0; a < 3; ++a) ur[3 * k + a] = ctl_soa[k + a * nctl];
    }
    unsigned char* S = ctx->noalloc;
    double* d_sum = reinterpret_cast<double*>(S);              // + kReduceBlocks partials
    unsigned* d_linf = reinterpret_cast<unsigned*>(S + 8448);
    float* d_bu = reinterpret_cast<float*>(S + 8704);
    float* d_uraw = reinterpret_cast<float*>(S + 11776);
    double* d_part = reinterpret_cast<double*>(S + 16384);
    static_assert(8 * (1 + kReduceBlocks) <= 8448, "reduction scratch overlaps");
    static_assert(8704 + 3 * 255 * 4 <= 11776 && 11776 + 3 * 255 * 4 <= 16384, "control tables overlap");
    hipStream_t st = ctx->stream;
    HIPCHECK(hipMemcpyAsync(d_bu, bu, 3 * (size_t)nctl * sizeof(float), hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(d_uraw, ur, 3 * (size_t)nctl * sizeof(float), hipMemcpyHostToDevice, st));
    HIPCHECK(pgk_launch_zero_u32(d_linf, st));
    int64_t nblk = 0;
    HIPCHECK(pgk_launch_sweep(t.K, Vin_dev, Vout_dev, occ_dev, d_bu, d_uraw, nullptr, nullptr, d_part, d_linf, 1,
                              st, &nblk));
    HIPCHECK(pgk_launch_reduce(d_part, nblk, d_sum, d_sum + 1, st));
    double hsum = 0.0;
    unsigned hl = 0;
    HIPCHECK(hipMemcpyAsync(&hsum, d_sum, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(&hl, d_linf, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    float lf;
    std::memcpy(&lf, &hl, 4);
    if (sum_decrease) *sum_decrease = hsum;
    if (linf) *linf = lf;
    return PG_OK;
}

int pg_solver_set_active(pg_solver* s, int32_t on) {
    if (!s) return fail(PG_ERR_ARG, "null solver");
    s->active_on = on != 0;
    s->act_for = -1;
    return PG_OK;
}

int pg_solver_set_warm_start(pg_solver* s, int32_t on) {
    if (!s) return fail(PG_ERR_ARG, "null solver");
    if (on && s->vstore != PG_VSTORE_F32)
        return fail(PG_ERR_UNSUPPORTED, "the warm start runs on an FP32 solver only");
    HIPCHECK(hipSetDevice(s->ctx->device));
    if (s->ooc) return pgh::ooc_set_warm_start(s, on != 0);
    if (!on) {
        if (s->d_wpol) HIPCHECK(hipFree(s->d_wpol));
        s->d_wpol = nullptr;
        s->kernel = auto_kernel(s);
        return PG_OK;
    }
    if (!s->d_wpol) {
        hipError_t e = hipMalloc(&s->d_wpol, 3 * (size_t)s->ncell);
        if (e != hipSuccess) {
            s->d_wpol = nullptr;
            return fail(e == hipErrorOutOfMemory ? PG_ERR_NOMEM : PG_ERR_HIP, "warm-start store allocation failed: %s",
                        hipGetErrorString(e));
        }
        HIPCHECK(hipMemsetAsync(s->d_wpol, 0, 3 * (size_t)s->ncell, s->ctx->stream));
        HIPCHECK(hipStreamSynchronize(s->ctx->stream));
    }
    s->kernel = PG_KERNEL_GLOBAL;
    return PG_OK;
}

int pg_download_warm_start(pg_solver* s, int8_t* dst) {
    if (!s || !dst) return fail(PG_ERR_ARG, "null argument");
    if (s->ooc) return pgh::ooc_download_warm_start(s, dst);
    if (!s->d_wpol) return fail(PG_ERR_STATE, "the warm start is off");
    HIPCHECK(hipSetDevice(s->ctx->device));
    HIPCHECK(hipMemcpyAsync(dst, s->d_wpol, 3 * (size_t)s->ncell, hipMemcpyDeviceToHost, s->ctx->stream));
    HIPCHECK(hipStreamSynchronize(s->ctx->stream));
    return PG_OK;
}

int pg_last_sweep_boxes(pg_solver* s, int64_t* computed, int64_t* total) {
    if (!s) return fail(PG_ERR_ARG, "null solver");
    const int64_t nbox = s->G.nbox;
    if (total) *total = nbox;
    if (computed) *computed = s->last_work >= 0 ? s->last_work : nbox;
    return PG_OK;
}

}  // extern "C"
