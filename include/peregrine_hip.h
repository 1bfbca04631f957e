/*
 * libperegrine_hip -- C ABI of the MI355X value-iteration solver.
 *
 * This library replaces the in-process GPU boundary of Peregrine's solver:
 * every kernel launch of `solver/src/Gpu.jl` and the device-array plumbing
 * that `solver/src/Run.jl` / `solver/src/OutOfCore.jl` do around it.  The
 * caller owns host memory; the library owns device memory.  Indices are
 * 0-based.  Every entry point returns an int status (0 = ok, < 0 = error)
 * and never throws; `pg_last_error()` gives the message of the last failure
 * on the calling thread.  A solver handle is used by one thread at a time.
 *
 * Conventions kept from the reference (file:line into the reference tree):
 *   - flat index  ((((ix*Ny+iy)*Nh+ih)*Nvx+ivx)*Nvy+ivy)*Nw+iw
 *     (solver/src/PeregrineSolver.jl:78-81, docs/TABLE_FORMAT.md:225-242)
 *   - occupancy   one byte per (x, y, h) cell, index (ix*Ny+iy)*Nh+ih
 *     (solver/src/Pipeline.jl:392-397)
 *   - controls    structure of arrays, ctl[k], ctl[k+nctl], ctl[k+2*nctl],
 *     entry 0 is coast (solver/src/Run.jl:283-285,
 *     solver/src/PeregrineSolver.jl:1263-1275)
 */
#ifndef PEREGRINE_HIP_H
#define PEREGRINE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PG_ABI_VERSION 1

/* status codes */
#define PG_OK               0
#define PG_ERR_ARG         -1   /* a bad argument: null, size, range        */
#define PG_ERR_HIP         -2   /* a HIP runtime call failed                */
#define PG_ERR_UNSUPPORTED -3   /* a parameter combination not built here   */
#define PG_ERR_NOMEM       -4   /* device memory budget exceeded            */
#define PG_ERR_STATE       -5   /* call out of order (e.g. sweep before seed) */

/* Uniform 6-D grid over (x, y, h, vx, vy, w); axis 2 (h) is periodic.
 * Mirrors `Grid6` (solver/src/PeregrineSolver.jl:47-60): lo/step in float,
 * heading lo = Float32(-pi), step = Float32(2pi/n). */
typedef struct pg_grid {
    int32_t n[6];
    float   lo[6];
    float   step[6];
} pg_grid;

/* Body-frame drivetrain regression, row-major 3x3 blocks.
 * Mirrors `Model` (solver/src/PeregrineSolver.jl:100-109). knee <= 0 or NaN
 * disables traction saturation (PeregrineSolver.jl:126-136). */
typedef struct pg_model {
    float B[9];
    float A[9];
    float q[3];
    float S[9];
    float D[9];
    float c[3];
    float eps[3];
    float knee;
} pg_model;

/* Backup parameters, field for field `Params`
 * (solver/src/PeregrineSolver.jl:897-916). Booleans are int32 0/1. */
typedef struct pg_params {
    float   dt;
    int32_t nsub;
    int32_t checks;
    int32_t adaptive_checks;
    int32_t nearest;
    float   cap;
    int32_t ntau;
    float   tau_ratio;
    float   cfl;
    float   tau_min;
    float   tau_max;
    float   hmax;
    int32_t ncoarse;
    int32_t rounds;   /* pattern-search rounds; with ntau > 1 (horizon ladder) runs on
                         the global-gather in-HBM FP32 kernel only */
    float   delta0;
    int32_t rk2;      /* explicit midpoint substeps (step_state_rk2); exact FP32
                         kernels only: the 16-bit stores, the FP64 tables and the
                         adjoint integrate with forward Euler and refuse it */
    int32_t vclamp;
    int32_t simplex;
} pg_params;

/* How V is held on the device. The 16-bit stores hold one code per cell and
 * an FP32 (offset, scale) per 4^6-cell tile, value = offset + scale * code;
 * unreached cells (value cap) hold an out-of-band code. */
#define PG_VSTORE_F32       0   /* plain FP32, Jacobi double buffer             */
#define PG_VSTORE_F16_TILE  1   /* IEEE binary16 codes, per-tile offset/scale    */
#define PG_VSTORE_U16_TILE  2   /* 16-bit fixed-point codes, per-tile offset/scale */

typedef struct pg_ctx pg_ctx;
typedef struct pg_solver pg_solver;

/* Library version string and ABI number. */
const char* pg_version(void);
int pg_abi_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char* pg_last_error(void);

/* Number of visible HIP devices. */
int pg_device_count(int32_t* n);

/* Bind a device and create its streams. Replaces `CUDA.functional()` and the
 * implicit CUDA.jl context (solver/src/Run.jl:1653-1654). */
int pg_create(int32_t device, pg_ctx** out);
int pg_destroy(pg_ctx* ctx);

/* Device memory this run may use: free * headroom.
 * Replaces `gpu_budget` (solver/src/Gpu.jl:181-185). */
int pg_device_budget(pg_ctx* ctx, double headroom, int64_t* bytes);

/* Would this grid fit whole, before anything is allocated?
 * Replaces `gpu_fits` (solver/src/Gpu.jl:204-209); the per-cell byte count is
 * the Jacobi pair of V buffers in the chosen store. */
int pg_grid_fits(pg_ctx* ctx, const pg_grid* g, int32_t vstore, double headroom,
                 int32_t* fits, double* need_gib, double* free_gib);

/* Allocate a resident solver for one grid: V (two buffers), occupancy and the
 * control table go to the device. Replaces the uploads at
 * solver/src/Run.jl:1776-1784. `occ` holds n[0]*n[1]*n[2] bytes (0 free,
 * nonzero blocked); `ctl_soa` holds 3*nctl floats with entry 0 = coast. */
int pg_solver_create(pg_ctx* ctx, const pg_grid* g, const pg_model* m,
                     const pg_params* p, const float* ctl_soa, int32_t nctl,
                     const uint8_t* occ, int32_t vstore, pg_solver** out);
int pg_solver_destroy(pg_solver* s);

/* fill!(V, cap); V[seeds] = 0 (solver/src/Run.jl:1437-1438). The seeds are
 * kept and reasserted after every sweep (Run.jl:1464). */
int pg_fill_seed(pg_solver* s, const int64_t* seeds, int64_t nseeds);

/* One Jacobi value-iteration sweep: V' = min(V, T V), seeds reasserted.
 * Replaces `sweep_gpu!` (solver/src/Gpu.jl:5-37). Returns the sum of the
 * per-cell decreases (FP64) and their maximum, the sup-norm of the change.
 * `phase` and `rev` are accepted for interface parity: the plain scheme
 * scans the whole control table (phase unused) and a Jacobi sweep has no
 * order (rev unused). */
int pg_sweep(pg_solver* s, int32_t phase, int32_t rev, double* sum_decrease,
             float* linf);

/* Sweep until max_iters, or until the change is quiet: sum <= tol_sum (the
 * reference's rule, Run.jl:1449-1471) when tol_sum >= 0, or sup-norm <
 * tol_linf when tol_linf > 0 (both may be set; either stops).
 * Replaces `solve_value!` (solver/src/Run.jl:1434-1474) after the seeding. */
int pg_solve(pg_solver* s, int32_t max_iters, double tol_sum, float tol_linf,
             int32_t* iters, double* last_sum, float* last_linf);

/* The escape pass over a solved table (in place of `sweep_escape_gpu!` and
 * `solve_escape!`, solver/src/Gpu.jl:39-80, solver/src/Run.jl:1503-1530).
 * V keeps its sign convention (PeregrineSolver.jl:756-856): 0 <= v < cap is a
 * real route (the boundary, never written), v < 0 an escape taking -v
 * seconds, v >= cap no way out yet. A sweep backs up every other cell, blocked
 * ones included (they may move through obstacle space, free cells may not
 * enter it), reading corners as the escape time they hold (routes as 0,
 * obstacle corners as cap for a free cell), and stores -c where the backup c
 * beats the cell's escape time. Jacobi, global-gather kernel, bit-exact with
 * the CPU restatement. Returns the summed and the largest improvement.
 * pg_solve_escape stops as pg_solve does. A 16-bit store must be polished
 * first. Out of core (solve_escape_ooc!, solver/src/OutOfCore.jl:609-700) the
 * tiles sweep the same kernel through windows planned for the escape reads
 * (every state, no probe rejections) on the first escape sweep -- with the
 * production candidates, through the production scheme's windows -- bit for
 * bit equal to the in-HBM pass. */
int pg_escape_sweep(pg_solver* s, int32_t phase, int32_t rev, double* sum_improve, float* linf);
int pg_solve_escape(pg_solver* s, int32_t max_iters, double tol_sum, float tol_linf, int32_t* iters,
                    double* last_sum, float* last_linf);

/* Finish a precision-form solve: decode the FP16 store into an FP32 Jacobi
 * pair (allocated here) and run nsweeps exact FP32 sweeps on it; afterwards
 * the solver behaves as an FP32 solver. With an FP32 store, runs nsweeps
 * exact sweeps. No counterpart in the reference (its only FP16 is the card's
 * output dtype, solver/src/Pipeline.jl:737-746). */
int pg_polish(pg_solver* s, int32_t nsweeps, double* last_sum, float* last_linf);

/* Argmin control index (0 = coast) of the Bellman backup of the current V,
 * one byte per cell; blocked cells report 0 (PeregrineSolver.jl:1226-1228).
 * Ties resolve to the lowest index, as the strict `<` scan does
 * (PeregrineSolver.jl:1245-1275). nctl must be <= 255. */
int pg_policy(pg_solver* s, uint8_t* pol_host);

/* Copy n values of the current V (FP32) starting at flat index off. With a
 * 16-bit store this is the FP32 value every state keeps of its own. */
int pg_download_V(pg_solver* s, float* dst, int64_t off, int64_t n);
/* No counterpart in the reference (diagnostics): the 16-bit mirror that the
 * precision form's gathers read, decoded to FP32, for cells [off, off + n).
 * PG_ERR_STATE unless a 16-bit store is active. */
int pg_download_V16(pg_solver* s, float* dst, int64_t off, int64_t n);
int pg_upload_V(pg_solver* s, const float* src, int64_t off, int64_t n);

/* Sweep kernel selection. GLOBAL and LDS are the exact FP32 sweeps (the
 * reference's unfused float arithmetic; identical bits). FAST is the
 * precision form's contracted FP32 sweep (fused multiply-adds, reciprocal
 * multiplies, hardware sin/cos, tensor-product interpolation). AUTO (the
 * default) takes LDS for an FP32 store with at most 128 controls, GLOBAL
 * otherwise; an FP16 store sweeps with FAST until pg_polish. */
/* Sweep order. JACOBI (the default) reads the previous table and writes the
 * next: deterministic, and for the exact kernels equal bit for bit to the
 * CPU restatement. IN_PLACE updates one table as the reference's GPU sweep
 * does (Gpu.jl:5-25, direction alternating with `rev`, Run.jl:1457-1458):
 * backups may read values already lowered in the same sweep, so fewer sweeps
 * reach a tolerance, and the result depends on the order the GPU schedules
 * boxes in (each sweep is still monotone and reaches the same fixed point). */
#define PG_SWEEP_JACOBI   0
#define PG_SWEEP_IN_PLACE 1
int pg_solver_set_sweep_mode(pg_solver* s, int32_t mode);

#define PG_KERNEL_AUTO   0
#define PG_KERNEL_GLOBAL 1
#define PG_KERNEL_LDS    2
#define PG_KERNEL_FAST   3
int pg_solver_set_kernel(pg_solver* s, int32_t kernel);

/* The kernel in use, and for the LDS-staged one the fraction of boxes whose
 * successors all fall inside the staged window (0 before the first sweep;
 * the rest read global memory for the corners outside). */
int pg_solver_kernel_info(pg_solver* s, int32_t* kernel, double* window_coverage);

/* No counterpart in the reference (its probes run for every pair, each
 * sweep): of the exact plan's boxes, the fraction whose probe chords meet no
 * obstacle (their sweeps make no occupancy lookups) and the fraction whose
 * every (state, control) pair passes its probes (their sweeps run no probes:
 * a pair's successor, and so the probes' outcome, is the same in every
 * sweep). Both 0 before the LDS-staged kernel's first sweep. */
int pg_solver_probe_info(pg_solver* s, double* chords_free, double* probes_pass);

/* Duration of the last sweep kernel, from HIP events on the solver stream. */
int pg_last_sweep_ms(pg_solver* s, float* ms);

/* Active boxes (on by default; no counterpart in the reference, whose
 * sweep_gpu! launches every cell each sweep, Gpu.jl:27-37). A Jacobi sweep of
 * the LDS-staged or precision-form kernels skips a box of states when no
 * cell that its backups can read (its read range, from a plan pass with the
 * sweep's own arithmetic), and none of its own cells, changed in the
 * previous sweep: min(V, TV) would leave it as it is. Tables, sums and
 * sup-norms are bit for bit those of computing every box. The set restarts
 * (every box computed once) after pg_fill_seed, pg_upload_V, pg_polish, a
 * kernel change, or a sweep it does not cover (global kernel, in-place
 * order, out-of-core tiles, omega cut into more than 32 boxes). */
int pg_solver_set_active(pg_solver* s, int32_t on);

/* Boxes the last sweep computed, and boxes of the grid. */
int pg_last_sweep_boxes(pg_solver* s, int64_t* computed, int64_t* total);

/* The warm-start policy store (replaces `pol = warm ? CUDA.zeros(Int8,
 * 3 * cells) : nothing`, Run.jl:1782, its reset in solve_value!,
 * Run.jl:1439, and pol_get / pol_set!, PeregrineSolver.jl:728-744): one
 * Int8 command per cell and axis, x 1/127, zeroed by pg_fill_seed. Each
 * sweep evaluates the cell's stored command after coast and writes the
 * winning command back when the cell's value improves. With the store on,
 * and with refine_rounds > 0 or tau_levels > 1, sweeps run on the
 * global-gather kernel (in-HBM FP32 solvers only). on = 0 frees it. */
int pg_solver_set_warm_start(pg_solver* s, int32_t on);

/* Copy the warm-start store (3 * ncells bytes) to the host. */
int pg_download_warm_start(pg_solver* s, int8_t* dst);

/* Number of cells of the solver's grid. */
int pg_solver_ncells(pg_solver* s, int64_t* n);

/* Out-of-core residency (replaces solve_value_ooc!, OutOfCore.jl:412-584,
 * and the tile planning of Tiles.jl:311-512). Both FP32 tables of the
 * Jacobi pair live in pinned host memory; a sweep streams one tile's window
 * of the old table to the device (side stream), sweeps the tile's interior
 * with the LDS-staged kernel and streams the interior back (second side
 * stream), overlapping copies with the sweep of the neighbouring tiles.
 * Tiles cut x, y and heading; the velocity axes stay whole. A tile's window
 * is the union of the cells its backups read, taken from the one-time
 * window plan, so each sweep equals the in-HBM sweep bit for bit.
 *
 * tile = {nx, ny, nheading} interior cells (rounded up to the box edge), or
 * {0, 0, 0} for the largest x-y tile with the whole heading axis whose
 * buffers fit budget_bytes of device memory. PG_ERR_NOMEM if they do not.
 * The solver then takes pg_fill_seed, pg_sweep, pg_solve, pg_download_V,
 * pg_upload_V, pg_escape_sweep, pg_solve_escape, pg_solver_set_warm_start,
 * pg_download_warm_start, pg_last_sweep_ms as usual (at most 128 controls;
 * no policy or FP64 tables). With the FP32 store the production scheme
 * (refine_rounds, tau_levels, the warm start) runs tiled on the global-gather
 * kernel through windows from the scheme's reach (every grid heading,
 * velocity and spin, every lattice command and commands sampled over the
 * octahedron, every horizon, two cells of margin), the warm-start store held
 * on the host and streamed beside each interior; a read outside a window
 * fails the sweep (PG_ERR_STATE) rather than read a wrong value. Those
 * windows (and the escape pass's) are sized on their first sweep, within
 * budget_bytes: PG_ERR_NOMEM there asks for a smaller tile.
 *
 * vstore = PG_VSTORE_U16_TILE or PG_VSTORE_F16_TILE runs the precision form
 * out of core: the host also holds the Jacobi pair of 16-bit codes, windows
 * stream in as codes (half the bytes) with the interior's own FP32 values,
 * and each tile's quantization tiles are refit and its interior encoded
 * before values and codes stream out; tiles are cut on multiples of 4 cells.
 * Every sweep equals the in-HBM precision-form sweep bit for bit; pg_polish
 * then continues with FP32 tiles. */
int pg_solver_create_ooc(pg_ctx* ctx, const pg_grid* g, const pg_model* m, const pg_params* p,
                         const float* ctl_soa, int32_t nctl, const uint8_t* occ, int32_t vstore,
                         int64_t budget_bytes, const int32_t tile[3], pg_solver** out);
/* The tile interior used, the tile count, the device bytes the tiled sweep
 * holds and the cells loaded per cell updated. */
int pg_ooc_info(pg_solver* s, int32_t tile[3], int64_t* ntiles, int64_t* device_bytes, double* amplification);
/* Tile k: interior origin t0 and extent tn, window origin wo (heading modulo
 * n3) and extent wl, each on x, y, heading. */
int pg_ooc_tile(pg_solver* s, int64_t k, int32_t* t0, int32_t* tn, int32_t* wo, int32_t* wl);

/* FP64 tables. No counterpart in the reference, whose tables are Float32
 * only: these are the FP64 yardstick the FP32 and 16-bit tables are measured
 * against, and the FP64 residual check of a finished table. The arithmetic
 * is the plain scheme carried in double (grid values, forward Euler, probes,
 * 64-corner multilinear lookup, PeregrineSolver.jl:208-223, 399-466,
 * 1100-1153); two extra FP64 tables live on the device once used.
 *
 * pg_solve_f64 starts from cap with the pg_fill_seed targets at 0 and runs
 * FP64 sweeps until ||dV||_inf < tol_linf or max_iters sweeps: PLAIN sweeps are
 * V' = min(V, T V); DIAGONAL sweeps solve each control's own-cell term of the
 * interpolation (V = tau + w V + rest  =>  V = (tau + rest) / (1 - w)),
 * which has T's fixed point and reaches it in fewer sweeps. At most 128
 * controls. */
#define PG_F64_SWEEP_PLAIN    0
#define PG_F64_SWEEP_DIAGONAL 1
int pg_solve_f64(pg_solver* s, int32_t mode, int32_t max_iters, double tol_linf, int32_t* iters,
                 double* last_linf);
/* One more FP64 sweep of the FP64 table (pg_solve_f64 with max_iters = 0
 * only sets up the seeded start). */
int pg_sweep_f64(pg_solver* s, int32_t mode, double* linf, double* sum_decrease);
/* max and sum over states of |T V - V| with T one FP64 backup, over the
 * states where T V or V is below cap; which = 0: the FP64 table, 1: the
 * current FP32 table (widened to double). */
int pg_residual_f64(pg_solver* s, int32_t which, double* linf, double* sum);
int pg_download_V64(pg_solver* s, double* dst, int64_t off, int64_t n);
/* argmin control of one FP64 backup of the FP64 table (ties within 2^-44
 * relative resolve to the lowest index); blocked cells report 0. */
int pg_policy_f64(pg_solver* s, uint8_t* pol_host);
/* Frees the FP64 tables. */
int pg_release_f64(pg_solver* s);

/* Adjoint. No counterpart in the reference (its only derivative is the Kuhn
 * interpolant's gradient in the state, PeregrineSolver.jl:533-570). For
 * L = mean over the nstarts states (row-major, 6 doubles each: x, y,
 * heading, field vx, vy, omega) of the multilinear V at the state, with V
 * the FP64 table of pg_solve_f64, returns L and dL/dtheta for the 30
 * drivetrain parameters theta = B (9, row-major), A (9), q (3), diagonal of
 * S (3), diagonal of D (3), c (3), the argmin control held fixed: the
 * transposed policy-fixed interpolation system (I - P)^T lambda = g is
 * iterated (FP64) until max |change| <= tol * max |lambda| (iters = -1 if
 * max_iters ran out), then lambda is contracted with the successors'
 * forward-mode derivatives. A start whose V is cap contributes cap to L and
 * nothing to the gradient. */
int pg_adjoint(pg_solver* s, const double* starts, int64_t nstarts, int32_t max_iters, double tol,
               double* dL_dtheta, double* L, int32_t* iters);

/* One Jacobi sweep V' = min(V, T V) over caller-owned device buffers that
 * allocates nothing. No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). It runs the global-gather sweep of
 * pg_sweep with the same results bit for bit (V, summed decrease, sup-norm),
 * for any scheme pg_sweep runs on that kernel except the warm-start store;
 * `phase` picks the rotating slice's offset as in pg_sweep. The caller
 * passes, as device pointers, occ_dev (n0*n1*n2 bytes, nonzero = blocked),
 * Vin_dev and Vout_dev (ncell floats each, distinct buffers); ctl_soa is a
 * host array (3*nctl floats, entry 0 coast). Seeds are not reasserted: a
 * Jacobi backup never raises a seed held at 0. The scratch it needs (control
 * tables, per-workgroup partial sums, the reduction) is a static device array
 * of PG_NOALLOC_SCRATCH_BYTES on each device, held until the process ends and
 * shared by every call on that device: calls that overlap on one device race.
 * Grids above PG_NOALLOC_MAX_CELLS cells are refused (PG_ERR_ARG) before
 * anything runs. Returns when the sweep is done. */
#define PG_NOALLOC_SCRATCH_BYTES (32 << 20)
#define PG_NOALLOC_MAX_CELLS ((int64_t)(((PG_NOALLOC_SCRATCH_BYTES) - 16384) / 8) * 256)
int pg_sweep_noalloc(pg_ctx* ctx, const pg_grid* g, const pg_model* m, const pg_params* p,
                     const float* ctl_soa, int32_t nctl, const uint8_t* occ_dev, const float* Vin_dev,
                     float* Vout_dev, int32_t phase, double* sum_decrease, float* linf);

#ifdef __cplusplus
}
#endif

#endif /* PEREGRINE_HIP_H */
