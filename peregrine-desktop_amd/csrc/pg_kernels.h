// Launch wrappers shared by the kernels and the C-ABI implementation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "pg_fast.hpp"
#include "pg_f64.hpp"

// One Jacobi sweep of the whole grid, global-gather kernel. uraw: the raw
// lattice commands (3 per control); wpol: the warm-start store (Int8, 3 per
// cell) or null.
hipError_t pgk_launch_sweep(const pgk::Consts& K, const float* Vin, float* Vout,
                            const uint8_t* occ, const float* bu, const float* uraw, uint8_t* pol, int8_t* wpol,
                            double* part, unsigned* linf_bits, int write_v, hipStream_t st, int64_t* nblocks);
// Sum of part[0, n) into *out; with mid (room for kReduceBlocks doubles)
// in two levels.
constexpr int kReduceBlocks = 1024;
// One Jacobi escape sweep (gpu_escape_kernel!, Gpu.jl:51-67) of the whole
// grid, global-gather kernel.
hipError_t pgk_launch_escape(const pgk::Consts& K, const float* Vin, float* Vout, const uint8_t* occ,
                             const float* bu, const float* uraw, uint8_t* pol, double* part, unsigned* linf_bits,
                             hipStream_t st, int64_t* nblocks);
// The same over an out-of-core tile's interior (t0, tn: x, y, heading),
// reading V from the tile's window (origin wo, extent wl; heading modulo n3)
// and writing the interior to `out`; a read outside the window raises *miss.
// With the production candidates when K asks for them (as pgk_launch_escape).
hipError_t pgk_launch_escape_tile(const pgk::Consts& K, const int32_t t0[3], const int32_t tn[3],
                                  const float* win, const int32_t wo[3], const int32_t wl[3], float* out,
                                  const uint8_t* occ, const float* bu, const float* uraw, double* part,
                                  unsigned* linf, unsigned* miss, hipStream_t st, int64_t* nblocks);
// One value sweep of the global-gather kernel over an out-of-core tile's
// interior through its window (as pgk_launch_escape_tile), with the
// production candidates when K or wpol asks for them; wpol is the interior's
// warm-start store (3 per cell, interior order) or null.
hipError_t pgk_launch_sweep_global_tile(const pgk::Consts& K, const int32_t t0[3], const int32_t tn[3],
                                        const float* win, const int32_t wo[3], const int32_t wl[3], float* out,
                                        const uint8_t* occ, const float* bu, const float* uraw, int8_t* wpol,
                                        double* part, unsigned* linf, unsigned* miss, hipStream_t st,
                                        int64_t* nblocks);
// The production scheme's reach in cells (lo/hi for x, y, heading into out[6],
// which the caller pre-sets to 0): every lattice command and the ns raw
// commands in raw (3 each), at every horizon the scheme uses.
hipError_t pgk_launch_reach(const pgk::Consts& K, const float* bu, const float* raw, int ns, int* out,
                            hipStream_t st);
hipError_t pgk_launch_reduce(const double* part, int64_t n, double* out, double* mid, hipStream_t st);
hipError_t pgk_launch_fill(float* V, int64_t n, float v, hipStream_t st);
hipError_t pgk_launch_seed(float* V, const int64_t* seeds, int64_t n, hipStream_t st);
// The same, marking in `mark` the box of every seed whose value it changes.
hipError_t pgk_launch_seed_mark(float* V, const int64_t* seeds, int64_t n, const pgk::Consts& K,
                                const pgk::BoxGeom& G, uint32_t* mark, hipStream_t st);

// LDS-staged path (pg_sweep_lds.hip).
int pgk_window_shapes(const int (**shapes)[6]);
// Per-box corner ranges (origin, cells per axis): with F null those of the
// exact successor, else those of the precision form's. With F, pfree[box] is
// 1 where every probe chord of the box meets free cells only (the precision
// form then skips the box's probes). With F null and pfree given, bit 1 says
// the same (the exact sweeps skip their occupancy lookups) and bit 2 that
// every pair of the box passed its probes (the exact sweeps run none).
hipError_t pgk_launch_plan(const pgk::Consts& K, const pgk::FastConsts* F, const pgk::BoxGeom& G,
                           const uint8_t* occ, const float* bu, int16_t* worig, uint8_t* wext, uint8_t* pfree,
                           int diag, hipStream_t st);
hipError_t pgk_launch_sweep_lds(int shape, int diag, const pgk::Consts& K, const pgk::BoxGeom& G,
                                const float* Vin, float* Vout, const uint8_t* occ, const float* bu,
                                const int16_t* worig, const uint8_t* pfree, uint8_t* pol, double* part,
                                unsigned* linf, int write_v, const pgk::ActiveG& A, hipStream_t st);
// One LDS-staged Jacobi escape sweep of the whole grid (gpu_escape_kernel!,
// Gpu.jl:51-67), window origins from a plan over an empty occupancy; pfree
// (may be null) is the value plan's per-box flag of probe chords that meet
// free cells only.
hipError_t pgk_launch_escape_lds(int shape, int diag, const pgk::Consts& K, const pgk::BoxGeom& G, const float* Vin,
                                 float* Vout, const uint8_t* occ, const float* bu, const int16_t* worig,
                                 const uint8_t* pfree, double* part, unsigned* linf, hipStream_t st);
// The same over one out-of-core tile: G is the tile's interior region, the
// window (as pgk_launch_sweep_tile's) holds the old V of the interior and of
// every corner the tile's escape backups read; results go to the interior
// buffer `out`; *miss is raised if a backup needed a cell outside the window.
hipError_t pgk_launch_escape_lds_tile(int shape, int diag, const pgk::Consts& K, const pgk::BoxGeom& G,
                                      const float* win, const int32_t wo[3], const int32_t wl[3], float* out,
                                      const uint8_t* occ, const float* bu, const int16_t* worig,
                                      const uint8_t* pfree, double* part, unsigned* linf, unsigned* miss,
                                      hipStream_t st);
// One tile of an out-of-core sweep: G is the tile's interior region; the
// window (cells wo[a] .. wo[a] + wl[a] - 1 on x, y and heading modulo n3,
// every velocity cell) holds the old V; results go to the interior buffer
// `out` (and the argmin to `pol`, when given). *miss is raised if a backup
// needed a cell outside the window.
hipError_t pgk_launch_sweep_tile(int shape, int diag, const pgk::Consts& K, const pgk::BoxGeom& G,
                                 const float* win, const int32_t wo[3], const int32_t wl[3], float* out,
                                 uint8_t* pol, const uint8_t* occ, const float* bu, const int16_t* worig,
                                 const uint8_t* pfree, double* part, unsigned* linf, unsigned* miss, hipStream_t st);

// Precision form (pg_sweep_fast.hip). pfree: the fast plan's per-box flags
// (probes skipped where set).
hipError_t pgk_launch_sweep_fast(int shape, int diag, int store, const pgk::Consts& K,
                                 const pgk::FastConsts& F, const pgk::BoxGeom& G,
                                 const pgk::QuantGeom& Q, const void* Vin, const float2* tp,
                                 const float* vself, float* Vout, uint16_t* codes_out, const uint8_t* occ,
                                 const float* bu, const int16_t* worig, const uint8_t* pfree, uint8_t* pol,
                                 double* part, unsigned* linf, unsigned* tstat, int write_v, const pgk::ActiveG& A,
                                 hipStream_t st);
hipError_t pgk_launch_tile_stats(const float* V, const pgk::Consts& K, const pgk::QuantGeom& Q,
                                 unsigned* tstat, hipStream_t st);
hipError_t pgk_launch_seed_stats(const int64_t* seeds, int64_t nseeds, const pgk::Consts& K,
                                 const pgk::QuantGeom& Q, unsigned* tstat, hipStream_t st);
// Refit tiles whose reached values left their range (every tile with force).
hipError_t pgk_launch_tile_refit(int store, unsigned* tstat, float2* tp, uint8_t* trefit, int64_t ntile,
                                 int force, hipStream_t st);
// FP32 table -> 16-bit mirror (changed codes and refit tiles only); with
// `mark`, the boxes whose decoded values change are marked there.
hipError_t pgk_launch_encode(int store, const float* V32, const pgk::Consts& K, const pgk::QuantGeom& Q,
                             const float2* tp, const uint8_t* trefit, uint16_t* codes, uint32_t* mark,
                             const pgk::BoxGeom& G, hipStream_t st);
// Out of core, 16-bit stores: one tile's sweep reading its window of codes
// (win: x [wo0, wo0 + wl0), y [wo1, wo1 + wl1), heading [wo2, wo2 + wl2)
// modulo n3; a read outside raises *miss) with the grid's offsets/scales tp,
// the interior's own FP32 values from self and its results into out (both
// interior-shaped, G's interior); tile statistics into tstat.
hipError_t pgk_launch_sweep_fast_tile(int shape, int diag, int store, const pgk::Consts& K, const pgk::FastConsts& F,
                                      const pgk::BoxGeom& G, const pgk::QuantGeom& Q, const uint16_t* win,
                                      const float2* tp, const int32_t wo[3], const int32_t wl[3], unsigned* miss,
                                      const float* self, float* out, const uint8_t* occ, const float* bu,
                                      const int16_t* worig, const uint8_t* pfree, double* part, unsigned* linf,
                                      unsigned* tstat, hipStream_t st);
// An interior buffer (x, y, heading cells t0 .. t0 + tn, every velocity
// cell): statistics of its reached values; refit of its tiles from tpo into
// tpn; its codes under tp.
hipError_t pgk_launch_interior_stats(const float* V, const int32_t t0[3], const int32_t tn[3], const pgk::Consts& K,
                                     const pgk::QuantGeom& Q, unsigned* tstat, hipStream_t st);
hipError_t pgk_launch_interior_refit(int store, const int32_t t0[3], const int32_t tn[3], const pgk::QuantGeom& Q,
                                     unsigned* tstat, const float2* tpo, float2* tpn, uint8_t* trefit, int force,
                                     hipStream_t st);
hipError_t pgk_launch_interior_encode(int store, const float* V, const int32_t t0[3], const int32_t tn[3],
                                      const pgk::Consts& K, const pgk::QuantGeom& Q, const float2* tp,
                                      uint16_t* codes, hipStream_t st);
// Codes of cells [off, off + n) -> V32[0, n).
hipError_t pgk_launch_decode(int store, const uint16_t* codes, const pgk::Consts& K, const pgk::QuantGeom& Q,
                             const float2* tp, int64_t off, int64_t n, float* V32, hipStream_t st);

// FP64 sweeps (pg_sweep_f64.hip).
#define PG_F64_PLAIN 0      // V' = min(V, T V)
#define PG_F64_DIAGONAL 1   // V' = min(V, C V), each control's own-cell term solved
#define PG_F64_RESIDUAL 2   // no write: |T V - V| where either is below cap
hipError_t pgk_launch_sweep_f64(int mode, int shape, const pgk::Consts64& K, const pgk::BoxGeom& G,
                                const double* Vin, double* Vout, const uint8_t* occ, const double* bu,
                                const int16_t* worig, const uint32_t* seedmask, uint8_t* pol, double* part,
                                unsigned* linf, hipStream_t st);
// Sets bit seeds[i] of mask (one bit per cell; the mask must start cleared).
hipError_t pgk_launch_seed_mask(uint32_t* mask, const int64_t* seeds, int64_t n, hipStream_t st);
hipError_t pgk_launch_fill_f64(double* V, int64_t n, double v, hipStream_t st);
hipError_t pgk_launch_seed_f64(double* V, const int64_t* seeds, int64_t n, hipStream_t st);
hipError_t pgk_launch_widen(const float* V32, double* V64, int64_t n, hipStream_t st);

// Adjoint (pg_adjoint.hip): rows of the policy-fixed system, the starts'
// weights, one iteration of the transposed solve, and per-block gradient
// partials (30 per block).
hipError_t pgk_launch_adj_linearize(const double* V, const uint8_t* pol, const uint32_t* seedmask, const uint8_t* occ,
                                    const double* bu, const pgk::Consts64& K, double* tcoord, int16_t* bcoord,
                                    double* self, hipStream_t st);
hipError_t pgk_launch_adj_starts(const double* V, const double* starts, int64_t n, const pgk::Consts64& K, double* g,
                                 double* vstart, hipStream_t st);
hipError_t pgk_launch_adj_iteration(double* lam, double* nxt, const double* g, const double* self, const double* tcoord,
                                    const int16_t* bcoord, const pgk::Consts64& K, unsigned long long* stat,
                                    hipStream_t st);
int64_t pgk_adj_grad_blocks(const pgk::Consts64& K);
hipError_t pgk_launch_adj_grad(const double* V, const double* lam, const uint8_t* pol, const double* tcoord,
                               const uint8_t* occ, const double* bu, const double* usat, const pgk::Consts64& K,
                               double* part, hipStream_t st);

// The no-allocation sweep's scratch (pg_noalloc.hip): the current device's
// address of the static array, and a one-word clear that is a kernel (a
// memset node misbehaves in captured graphs).
hipError_t pgk_noalloc_scratch(void** ptr);
hipError_t pgk_launch_zero_u32(unsigned* p, hipStream_t st);
