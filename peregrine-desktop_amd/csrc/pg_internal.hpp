// Internals of libperegrine_hip shared by its host source files.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include "../../include/peregrine_hip.h"
#include "pg_kernels.h"

namespace pgh {

extern thread_local std::string g_err;
// Records a formatted message for pg_last_error and returns code.
int fail(int code, const char* fmt, ...);
int64_t grid_cells(const pg_grid& g);

constexpr int kLdsMaxCtl = 128;   // controls the LDS-staged sweeps hold

struct Ooc;

}  // namespace pgh

#define HIPCHECK(expr)                                                                       \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return pgh::fail(PG_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));    \
    } while (0)

struct pg_ctx {
    int32_t device = 0;
    hipStream_t stream = nullptr;
    // device address of this device's pg_sweep_noalloc scratch (pg_noalloc.hip)
    unsigned char* noalloc = nullptr;
};

struct pg_solver {
    pg_ctx* ctx = nullptr;
    pg_grid g{};
    pg_model m{};
    pg_params p{};
    pgk::Consts K{};
    pgk::FastConsts F{};
    int32_t nctl = 0;
    int32_t vstore = 0;
    int64_t ncell = 0;
    // FP32 tables: the Jacobi pair (FP16 store: [0] is the scratch the fast
    // sweep writes, [1] exists once polishing starts)
    float* dV[2] = {nullptr, nullptr};
    int cur = 0;
    // FP16 store
    bool f16_active = false;      // V lives in the 16-bit codes (until polished)
    int store = PG_STORE_F32;     // code type of the 16-bit store
    uint16_t* d_V16 = nullptr;
    float2* d_tp = nullptr;
    unsigned* d_tstat = nullptr;
    uint8_t* d_trefit = nullptr;
    pgk::QuantGeom Q{};
    uint8_t* d_occ = nullptr;
    float* d_bu = nullptr;
    float* d_uraw = nullptr;     // raw lattice commands, 3 per control
    int8_t* d_wpol = nullptr;    // warm-start store, Int8 x 3 per cell (null: off)
    int64_t* d_seeds = nullptr;
    int64_t nseeds = 0;
    std::vector<int64_t> h_seeds;
    bool seeded = false;
    double* d_part = nullptr;
    int64_t part_cap = 0;
    double* d_sum = nullptr;
    unsigned* d_linf = nullptr;
    uint8_t* d_pol = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0.0f;
    // kernel selection and the LDS window plan
    int32_t kernel = PG_KERNEL_GLOBAL;
    bool diag = false;
    pgk::BoxGeom G{};
    int16_t* d_worig = nullptr;
    bool planned = false;
    int shape = 0;
    double coverage = 0.0;
    int32_t sweep_mode = PG_SWEEP_JACOBI;
    // active boxes of Jacobi sweeps: the change bits of the last sweep in
    // d_act[act_cur], valid for the sweep regime act_for (0 exact FP32,
    // 1 precision-form FP32, 2 16-bit store; -1: none, every box computed)
    bool active_on = true;
    uint32_t* d_act[2] = {nullptr, nullptr};
    int act_cur = 0;
    int act_for = -1;
    uint8_t* d_wext = nullptr;      // exact read ranges: cells per axis from d_worig
    uint8_t* d_pfree = nullptr;     // exact plan: boxes whose probe chords meet free cells only
    int16_t* d_forig = nullptr;     // precision-form read ranges
    uint8_t* d_fext = nullptr;
    uint8_t* d_ffree = nullptr;     // precision form: boxes whose probes only meet free cells
    // the LDS-staged escape sweep: window origins from a plan over an empty
    // occupancy (every state's successors, no probe rejections)
    int16_t* d_worig_esc = nullptr;
    uint8_t* d_occ0 = nullptr;
    bool esc_planned = false;
    int esc_shape = 0;
    double esc_coverage = 0.0;
    unsigned* d_nwork = nullptr;
    int64_t last_work = -1;         // boxes the last sweep computed (-1: every box)
    // FP64 tables (allocated on first use)
    pgk::Consts64 K64{};
    std::vector<double> bu64;
    std::vector<double> usat64;   // saturated commands per control, in double
    double* d_bu64 = nullptr;
    double* dV64[2] = {nullptr, nullptr};
    uint32_t* d_seedmask = nullptr;   // one bit per cell: the pg_fill_seed targets
    bool seedmask_ok = false;
    int cur64 = 0;
    bool f64_valid = false;       // dV64[cur64] holds an FP64 table
    // out-of-core: both tables of the Jacobi pair in pinned host memory,
    // swept tile by tile (pg_ooc.cpp)
    pgh::Ooc* ooc = nullptr;
};


namespace pgh {

// The LDS window plan of the whole grid (pg_api.cpp): per-box window origins
// on the device; with keep_host, origins and extents also on the host.
int plan(pg_solver* s, bool keep_host = false, std::vector<int16_t>* h_orig = nullptr,
         std::vector<uint8_t>* h_ext = nullptr);
int ensure_partials(pg_solver* s, int64_t nb);

// Out-of-core residency (pg_ooc.cpp).
int ooc_create(pg_solver* s, int64_t budget_bytes, const int32_t tile[3]);
void ooc_release(pg_solver* s);
int ooc_fill_seed(pg_solver* s);
int ooc_sweep(pg_solver* s, double* sum_decrease, float* linf);
// Whether a tiled solver's sweeps take the production candidates (pattern
// search, horizon ladder, warm start): then they run the global-gather kernel
// through windows from the scheme's reach.
bool ooc_production(const pg_solver* s);
int ooc_set_warm_start(pg_solver* s, bool on);
int ooc_download_warm_start(const pg_solver* s, int8_t* dst);
// One escape sweep over the tiles (the phase's slice offset already in s->K).
int ooc_escape_sweep(pg_solver* s, double* sum_improve, float* linf);
int ooc_download(pg_solver* s, float* dst, int64_t off, int64_t n);
int ooc_upload(pg_solver* s, const float* src, int64_t off, int64_t n);
int ooc_polish_start(pg_solver* s);
const uint16_t* ooc_host_codes(const pg_solver* s);
int ooc_info(pg_solver* s, int32_t tile[3], int64_t* ntiles, int64_t* device_bytes, double* amplification);
int ooc_tile(pg_solver* s, int64_t k, int32_t t0[3], int32_t tn[3], int32_t wo[3], int32_t wl[3]);

}  // namespace pgh
