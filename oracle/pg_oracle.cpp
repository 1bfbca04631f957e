// CPU ORACLE -- TEST INFRASTRUCTURE ONLY (see pg_oracle.hpp).
//
// C entry points over the templated restatement, for ctypes. Sweeps:
//   mode 0  Jacobi:  Vout = min(Vin, T Vin), every cell from Vin.
//   mode 1  in-place, serial, in index order (reversed when rev != 0): the
//           reference's asynchronous sweep (solver/src/Pipeline.jl:552-568)
//           with the scheduling fixed so that it is deterministic.
//   mode 2  in-place, OpenMP-threaded: the reference's `sweep_cpu!` as it
//           runs, races included (the timed CPU baseline).
#include "pg_oracle.hpp"
#include <omp.h>
#include <algorithm>
#include <vector>

using namespace orc;

namespace {

bool params_supported(const pg_params* p) {
    return p->rounds >= 0 && p->ntau >= 1 && p->ntau <= 64;
}

template <class T>
int sweep_impl(const T* Vin, T* Vout, const uint8_t* occ, const pg_grid* g, const pg_model* m,
               const float* ctl, int32_t nctl, const pg_params* p, int32_t mode, int32_t rev,
               int32_t phase, int32_t nthreads, uint8_t* pol, double* sum_out, double* linf_out,
               int32_t esc, int8_t* wpol = nullptr, int64_t lo = 0, int64_t hi = -1) {
    if (!params_supported(p) || nctl < 2) return -3;
    if (hi < 0) hi = ncells(*g);
    if (lo < 0 || lo > hi || hi > ncells(*g)) return -2;
    const int64_t n = hi - lo;
    double sum = 0.0, linf = 0.0;
    if (mode == 0) {
        if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 512) reduction(+ : sum) reduction(max : linf)
        for (int64_t i = lo; i < hi; ++i) {
            T old = Vin[i];
            if (esc) {
                // gpu_escape_kernel! update rule (solver/src/Gpu.jl:51-67)
                bool terminal = old >= (T)0 && old < (T)p->cap;
                T e_old = old < (T)0 ? -old : (T)p->cap;
                T out = old;
                int32_t k = 0;
                if (!terminal) {
                    T c = cell_update<T>(i, Vin, occ, *g, *m, ctl, nctl, *p, phase, true, &k);
                    if (c < e_old) {
                        out = -c;
                        sum += (double)(e_old - c);
                        linf = std::max(linf, (double)(e_old - c));
                    }
                }
                Vout[i] = out;
                if (pol) pol[i] = (uint8_t)k;
                continue;
            }
            int32_t k = 0;
            T b[3];
            T c = cell_update<T>(i, Vin, occ, *g, *m, ctl, nctl, *p, phase, false, &k, wpol, b);
            if (c < old) {
                Vout[i] = c;
                sum += (double)(old - c);
                linf = std::max(linf, (double)(old - c));
                if (wpol)
                    for (int a = 0; a < 3; ++a) wpol[3 * i + a] = pol_quant((float)b[a]);
            } else {
                Vout[i] = old;
            }
            if (pol) pol[i] = (uint8_t)k;
        }
    } else if (mode == 1 || mode == 2) {
        // in place: Vout is the one array; Vin must equal Vout
        T* V = Vout;
        if (mode == 2 && nthreads > 0) omp_set_num_threads(nthreads);
        auto body = [&](int64_t ii, double& s, double& l) {
            int64_t idx = rev ? (hi - 1 - ii) : lo + ii;
            T old = V[idx];
            int32_t k = 0;
            if (esc) {
                if (old >= (T)0 && old < (T)p->cap) return;
                T e_old = old < (T)0 ? -old : (T)p->cap;
                T c = cell_update<T>(idx, V, occ, *g, *m, ctl, nctl, *p, phase, true, &k);
                if (c < e_old) {
                    V[idx] = -c;
                    s += (double)(e_old - c);
                    l = std::max(l, (double)(e_old - c));
                }
                return;
            }
            T b[3];
            T c = cell_update<T>(idx, V, occ, *g, *m, ctl, nctl, *p, phase, false, &k, wpol, b);
            if (c < old) {
                V[idx] = c;
                s += (double)(old - c);
                l = std::max(l, (double)(old - c));
                if (pol) pol[idx] = (uint8_t)k;
                if (wpol)
                    for (int a = 0; a < 3; ++a) wpol[3 * idx + a] = pol_quant((float)b[a]);
            }
        };
        if (mode == 1) {
            for (int64_t ii = 0; ii < n; ++ii) body(ii, sum, linf);
        } else {
#pragma omp parallel for schedule(dynamic, 512) reduction(+ : sum) reduction(max : linf)
            for (int64_t ii = 0; ii < n; ++ii) body(ii, sum, linf);
        }
    } else {
        return -1;
    }
    if (sum_out) *sum_out = sum;
    if (linf_out) *linf_out = linf;
    return 0;
}

}  // namespace

namespace {

struct Row {
    int64_t base[6];   // corner indices j0 per axis (j1 derived)
    double t[6];       // fractional parts
    double self;       // weight of the state's own cell
    bool live;
};

// The successor of grid state idx under control k in double, as lookahead
// computes it (:1100-1153); false when the control is rejected.
bool successor_f64(int64_t idx, int32_t k, const uint8_t* occ, const pg_grid& g, const pg_model& m,
                   const float* ctl, int32_t nctl, const pg_params& p, double s0[6], double f[6]) {
    int32_t i[6];
    unflatten(g, idx, i);
    double h = axisvalue<double>(g, 2, i[2]);
    double vfx = axisvalue<double>(g, 3, i[3]), vfy = axisvalue<double>(g, 4, i[4]);
    double sh0, ch0;
    sincos_t(h, sh0, ch0);
    s0[0] = axisvalue<double>(g, 0, i[0]);
    s0[1] = axisvalue<double>(g, 1, i[1]);
    s0[2] = h;
    s0[3] = ch0 * vfx + sh0 * vfy;
    s0[4] = -sh0 * vfx + ch0 * vfy;
    s0[5] = axisvalue<double>(g, 5, i[5]);
    double s[6];
    for (int q = 0; q < 6; ++q) s[q] = s0[q];
    const double tau = (double)p.dt;
    step_state<double>(m, s, (double)ctl[k], (double)ctl[k + nctl], (double)ctl[k + 2 * nctl], tau,
                       substeps_for<double>(tau, p));
    const double dx = s[0] - s0[0], dy = s[1] - s0[1], dh = s[2] - s0[2];
    if (!(isfin(dx) && isfin(dy) && isfin(dh) && isfin(s[3]) && isfin(s[4]) && isfin(s[5]))) return false;
    const double n3 = (double)(float)g.n[2];
    for (int32_t sc = 1; sc <= p.checks; ++sc) {
        const double a = (double)(float)sc / (double)(float)p.checks;
        const double gx = (s0[0] + dx * a - (double)g.lo[0]) / (double)g.step[0];
        const double gy = (s0[1] + dy * a - (double)g.lo[1]) / (double)g.step[1];
        if (gx < 0 || gy < 0 || gx > (double)(float)(g.n[0] - 1) || gy > (double)(float)(g.n[1] - 1)) return false;
        double gh = (s0[2] + dh * a - (double)g.lo[2]) / (double)g.step[2];
        gh = gh - std::floor(gh / n3) * n3;
        if (!(gh >= 0 && gh <= n3)) return false;
        const int32_t ci = (int32_t)jround(gx), cj = (int32_t)jround(gy), ck = (int32_t)jround(gh) % g.n[2];
        if (occ[((int64_t)ci * g.n[1] + cj) * g.n[2] + ck]) return false;
    }
    double snh, cnh;
    sincos_t(s[2], snh, cnh);
    const double sv[6] = {s[0], s[1], s[2], cnh * s[3] - snh * s[4], snh * s[3] + cnh * s[4], s[5]};
    for (int a = 0; a < 6; ++a) f[a] = (sv[a] - (double)g.lo[a]) / (double)g.step[a];
    f[2] = f[2] - std::floor(f[2] / n3) * n3;
    if (!(f[2] >= 0 && f[2] <= n3)) return false;
    for (int a = 0; a < 6; ++a) {
        if (a == 2) continue;
        if (!(f[a] >= 0 && f[a] <= (double)(float)(g.n[a] - 1))) return false;
    }
    return true;
}

void row_of(const pg_grid& g, const double f[6], int64_t idx, Row& r) {
    int32_t own[6];
    unflatten(g, idx, own);
    r.self = 1.0;
    for (int a = 0; a < 6; ++a) {
        const double b = std::floor(f[a]);
        r.t[a] = f[a] - b;
        r.base[a] = (int64_t)b;
        int32_t d = own[a] - (int32_t)b;
        if (a == 2) d = ((d % g.n[2]) + g.n[2]) % g.n[2];
        r.self *= d == 0 ? 1.0 - r.t[a] : (d == 1 ? r.t[a] : 0.0);
    }
}

template <class Fn>
void for_corners(const pg_grid& g, const Row& r, Fn fn) {
    for (int d = 0; d < 64; ++d) {
        int32_t j[6];
        double w = 1.0;
        for (int k = 0; k < 6; ++k) {
            const int c = (d >> k) & 1;
            const int32_t b = (int32_t)r.base[k];
            j[k] = (k == 2) ? (b + c) % g.n[2] : jmin<int32_t>(b + c, g.n[k] - 1);
            w *= c ? r.t[k] : 1.0 - r.t[k];
        }
        fn(flatten(g, j[0], j[1], j[2], j[3], j[4], j[5]), w);
    }
}

}  // namespace

extern "C" {

int orc_abi(void) { return PG_ABI_VERSION; }

// sizeof of the C-ABI structs as the compiler lays them out from the header
// (0 pg_grid, 1 pg_model, 2 pg_params), to check the ctypes mirrors.
int64_t orc_sizeof(int32_t which) {
    switch (which) {
        case 0: return (int64_t)sizeof(pg_grid);
        case 1: return (int64_t)sizeof(pg_model);
        case 2: return (int64_t)sizeof(pg_params);
        default: return -1;
    }
}

void orc_sincosf(float x, float* s, float* c) { sincos_f32(x, *s, *c); }

// Batch form for accuracy tests.
void orc_sincosf_batch(const float* x, float* s, float* c, int64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) sincos_f32(x[i], s[i], c[i]);
}

float orc_interp_f32(const float* V, const pg_grid* g, const float* st, int32_t nearest, float cap,
                     int32_t vclamp) {
    PlainV<float> pv{V};
    return interp<float>(pv, *g, st[0], st[1], st[2], st[3], st[4], st[5], nearest != 0, cap,
                         vclamp != 0);
}

double orc_interp_f64(const double* V, const pg_grid* g, const double* st, int32_t nearest,
                      double cap, int32_t vclamp) {
    PlainV<double> pv{V};
    return interp<double>(pv, *g, st[0], st[1], st[2], st[3], st[4], st[5], nearest != 0, cap,
                          vclamp != 0);
}

float orc_interp_kuhn_f32(const float* V, const pg_grid* g, const float* st, float cap,
                          int32_t vclamp, float* grad) {
    PlainV<float> pv{V};
    return interp_kuhn<float>(pv, *g, st[0], st[1], st[2], st[3], st[4], st[5], cap, vclamp != 0,
                              grad);
}

void orc_step_state_f32(const pg_model* m, const float* s0, const float* u, float dt, int32_t nsub,
                        int32_t rk2, float* out) {
    for (int j = 0; j < 6; ++j) out[j] = s0[j];
    if (rk2) step_state_rk2<float>(*m, out, u[0], u[1], u[2], dt, nsub);
    else step_state<float>(*m, out, u[0], u[1], u[2], dt, nsub);
}

void orc_step_state_f64(const pg_model* m, const double* s0, const double* u, double dt,
                        int32_t nsub, int32_t rk2, double* out) {
    for (int j = 0; j < 6; ++j) out[j] = s0[j];
    if (rk2) step_state_rk2<double>(*m, out, u[0], u[1], u[2], dt, nsub);
    else step_state<double>(*m, out, u[0], u[1], u[2], dt, nsub);
}

float orc_traction_gain(float u1, float u2, float u3, float knee) {
    return traction_gain<float>(u1, u2, u3, knee);
}

// Body-frame acceleration of the model at raw command u and body velocity v
// (saturate + body_accel_sat, PeregrineSolver.jl:139-162); a[3] receives the
// traction gain.
void orc_body_accel_f64(const pg_model* m, const double* u, const double* v, double* a) {
    double u1, u2, u3;
    saturate<double>(*m, u[0], u[1], u[2], u1, u2, u3);
    body_accel_sat<double>(*m, u1, u2, u3, v[0], v[1], v[2], a[0], a[1], a[2]);
    a[3] = traction_gain<double>(u[0], u[1], u[2], m->knee);
}

void orc_body_accel_f32(const pg_model* m, const float* u, const float* v, float* a) {
    float u1, u2, u3;
    saturate<float>(*m, u[0], u[1], u[2], u1, u2, u3);
    body_accel_sat<float>(*m, u1, u2, u3, v[0], v[1], v[2], a[0], a[1], a[2]);
    a[3] = traction_gain<float>(u[0], u[1], u[2], m->knee);
}

float orc_lookahead_f32(const float* V, const uint8_t* occ, const pg_grid* g, const pg_model* m,
                        const float* s, const float* u, float tau, int32_t nsub,
                        const pg_params* p) {
    PlainV<float> pv{V};
    return lookahead<float>(pv, occ, false, *g, *m, s[0], s[1], s[2], s[3], s[4], s[5], u[0], u[1],
                            u[2], tau, nsub, *p);
}

int32_t orc_substeps_for(float tau, const pg_params* p) { return substeps_for<float>(tau, *p); }

float orc_cfl_tau(const pg_grid* g, const pg_model* m, float vfx, float vfy, float w,
                  const pg_params* p) {
    return cfl_tau<float>(*g, *m, vfx, vfy, w, *p);
}

// One backup of one cell: value and winning control index.
float orc_cell_update_f32(int64_t idx, const float* V, const uint8_t* occ, const pg_grid* g,
                          const pg_model* m, const float* ctl, int32_t nctl, const pg_params* p,
                          int32_t phase, int32_t esc, int32_t* best_k) {
    return cell_update<float>(idx, V, occ, *g, *m, ctl, nctl, *p, phase, esc != 0, best_k);
}

double orc_cell_update_f64(int64_t idx, const double* V, const uint8_t* occ, const pg_grid* g,
                           const pg_model* m, const float* ctl, int32_t nctl, const pg_params* p,
                           int32_t phase, int32_t esc, int32_t* best_k) {
    return cell_update<double>(idx, V, occ, *g, *m, ctl, nctl, *p, phase, esc != 0, best_k);
}

// Backups of a list of cells against a fixed V (no write), for subsampled
// error measurement at sizes the full sweep cannot reach on the CPU.
int orc_backup_cells_f64(const double* V, const uint8_t* occ, const pg_grid* g, const pg_model* m,
                         const float* ctl, int32_t nctl, const pg_params* p, const int64_t* cells,
                         int64_t ncell_list, double* out, int32_t* kout, int32_t nthreads) {
    if (!params_supported(p) || nctl < 2) return -3;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t j = 0; j < ncell_list; ++j) {
        int32_t k = 0;
        out[j] = cell_update<double>(cells[j], V, occ, *g, *m, ctl, nctl, *p, 0, false, &k);
        if (kout) kout[j] = k;
    }
    return 0;
}

// Backups of a list of cells against V in float (no write): the timed CPU
// baseline on a bounded sample of a grid too large to sweep on the host.
int orc_backup_cells_f32(const float* V, const uint8_t* occ, const pg_grid* g, const pg_model* m,
                         const float* ctl, int32_t nctl, const pg_params* p, const int64_t* cells,
                         int64_t ncell_list, float* out, int32_t nthreads) {
    if (!params_supported(p) || nctl < 2) return -3;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t j = 0; j < ncell_list; ++j) {
        int32_t k = 0;
        out[j] = cell_update<float>(cells[j], V, occ, *g, *m, ctl, nctl, *p, 0, false, &k);
    }
    return 0;
}

int orc_sweep_f32(const float* Vin, float* Vout, const uint8_t* occ, const pg_grid* g,
                  const pg_model* m, const float* ctl, int32_t nctl, const pg_params* p,
                  int32_t mode, int32_t rev, int32_t phase, int32_t nthreads, uint8_t* pol,
                  double* sum, double* linf, int32_t esc) {
    return sweep_impl<float>(Vin, Vout, occ, g, m, ctl, nctl, p, mode, rev, phase, nthreads, pol,
                             sum, linf, esc);
}

// A sweep with the warm-start policy store (Int8, 3 per cell, interleaved):
// each cell reads its entry and, when its value improves, writes the
// winning command back (gpu_sweep_kernel!, Gpu.jl:5-25).
int orc_sweep_warm_f32(const float* Vin, float* Vout, const uint8_t* occ, const pg_grid* g,
                       const pg_model* m, const float* ctl, int32_t nctl, const pg_params* p,
                       int32_t mode, int32_t rev, int32_t phase, int32_t nthreads, int8_t* wpol,
                       double* sum, double* linf) {
    return sweep_impl<float>(Vin, Vout, occ, g, m, ctl, nctl, p, mode, rev, phase, nthreads, nullptr,
                             sum, linf, 0, wpol);
}

// An in-place, OpenMP-threaded sweep (mode 2) of the cells [lo, hi) only:
// the reference CPU path's update order over a contiguous slab, for timing
// it on a bounded part of a grid too large to sweep whole on the host.
int orc_sweep_range_f32(float* V, const uint8_t* occ, const pg_grid* g, const pg_model* m,
                        const float* ctl, int32_t nctl, const pg_params* p, int64_t lo, int64_t hi,
                        int32_t rev, int32_t phase, int32_t nthreads, double* sum, double* linf) {
    return sweep_impl<float>(V, V, occ, g, m, ctl, nctl, p, 2, rev, phase, nthreads, nullptr, sum, linf, 0,
                             nullptr, lo, hi);
}

// The oracle's float tanh of a non-negative argument (the traction gain's).
float orc_tanh_f32(float r) { return tanh_t(r); }

int orc_sweep_f64(const double* Vin, double* Vout, const uint8_t* occ, const pg_grid* g,
                  const pg_model* m, const float* ctl, int32_t nctl, const pg_params* p,
                  int32_t mode, int32_t rev, int32_t phase, int32_t nthreads, uint8_t* pol,
                  double* sum, double* linf, int32_t esc) {
    return sweep_impl<double>(Vin, Vout, occ, g, m, ctl, nctl, p, mode, rev, phase, nthreads, pol,
                              sum, linf, esc);
}

// target_cells (solver/src/Pipeline.jl:504-536): cells strictly within tol on
// every axis (heading wrapped), else the single nearest cell. Writes up to
// `cap_out` 0-based indices (sorted, unique) and returns the count.
int64_t orc_target_cells(const pg_grid* g, const float* s, const float* tol, int64_t* out,
                         int64_t cap_out) {
    int64_t lo[6], hi[6];
    for (int k = 0; k < 6; ++k) {
        float c = (s[k] - g->lo[k]) / g->step[k];
        if (k == 2) {
            float n3 = (float)g->n[2];
            c = c - std::floor(c / n3) * n3;
        }
        float r = tol[k] / g->step[k];
        int64_t l = (int64_t)std::ceil(c - r);
        int64_t h = (int64_t)std::floor(c + r);
        if (k != 2) {
            l = std::max<int64_t>(0, l);
            h = std::min<int64_t>((int64_t)g->n[k] - 1, h);
        }
        lo[k] = l; hi[k] = h;
    }
    std::vector<int64_t> idx;
    for (int64_t a = lo[0]; a <= hi[0]; ++a)
        for (int64_t b = lo[1]; b <= hi[1]; ++b)
            for (int64_t c = lo[2]; c <= hi[2]; ++c)
                for (int64_t d = lo[3]; d <= hi[3]; ++d)
                    for (int64_t e = lo[4]; e <= hi[4]; ++e)
                        for (int64_t f = lo[5]; f <= hi[5]; ++f) {
                            int64_t n3 = g->n[2];
                            int64_t j3 = ((c % n3) + n3) % n3;
                            idx.push_back(flatten(*g, (int32_t)a, (int32_t)b, (int32_t)j3,
                                                  (int32_t)d, (int32_t)e, (int32_t)f));
                        }
    if (idx.empty()) {
        int32_t nr[6];
        for (int k = 0; k < 6; ++k) {
            float c = (s[k] - g->lo[k]) / g->step[k];
            int64_t r = (int64_t)std::nearbyint(c);
            if (k == 2) {
                int64_t n3 = g->n[2];
                nr[k] = (int32_t)(((r % n3) + n3) % n3);
            } else {
                nr[k] = (int32_t)std::min<int64_t>(std::max<int64_t>(r, 0), g->n[k] - 1);
            }
        }
        idx.push_back(flatten(*g, nr[0], nr[1], nr[2], nr[3], nr[4], nr[5]));
    }
    std::sort(idx.begin(), idx.end());
    idx.erase(std::unique(idx.begin(), idx.end()), idx.end());
    int64_t nout = (int64_t)idx.size();
    for (int64_t i = 0; i < nout && i < cap_out; ++i) out[i] = idx[i];
    return nout;
}

// ---------------------------------------------------------------------------
// Adjoint of L = mean over starts of interp(V, start) with respect to the 30
// drivetrain parameters at a fixed point V of the plain backup (no
// counterpart in the reference). With the argmin control held fixed, every
// reached state's value is tau + P_i . V (P_i: its successor's 64 corner
// weights); targets, blocked and capped states have no row. Then
//   dL/dtheta = sum_i lambda_i  d(P_i . V)/dtheta,   (I - P)^T lambda = g,
// g = mean of the starts' corner weights, and d(P_i . V)/dtheta is the
// interpolant's gradient at the successor times the successor's tangent.
// lambda is iterated as lambda_j <- (g_j + sum_{i != j} P_ij lambda_i) / (1 - P_jj).
// Returns the iterations used (or -1 if max_iters ran out).
// ---------------------------------------------------------------------------

int32_t orc_adjoint_f64(const double* V, const uint8_t* occ, const pg_grid* g, const pg_model* m,
                        const float* ctl, int32_t nctl, const pg_params* p, const int64_t* seeds,
                        int64_t nseeds, const double* starts, int64_t nstarts, int32_t max_iters, double tol,
                        int32_t nthreads, double* grad, double* L, int32_t* pol_out) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
    const int64_t n = ncells(*g);
    const double cap = (double)p->cap;
    std::vector<uint8_t> seed(n, 0);
    for (int64_t i = 0; i < nseeds; ++i) seed[seeds[i]] = 1;
    std::vector<Row> rows(n);
    std::vector<int32_t> pol(n, 0);
    const int64_t blk = (int64_t)g->n[3] * g->n[4] * g->n[5];
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
        Row& r = rows[i];
        r.live = false;
        if (seed[i] || occ[i / blk] || !(V[i] < cap)) continue;
        int32_t k = 0;
        const double best = cell_update<double>(i, V, occ, *g, *m, ctl, nctl, *p, 0, false, &k);
        pol[i] = k;
        if (!(best < cap)) continue;
        double s0[6], f[6];
        if (!successor_f64(i, k, occ, *g, *m, ctl, nctl, *p, s0, f)) continue;
        row_of(*g, f, i, r);
        r.live = true;
    }
    // g and L from the starts
    std::vector<double> gv(n, 0.0), lam(n, 0.0), nxt(n, 0.0);
    double Ls = 0.0;
    PlainV<double> pv{V};
    for (int64_t k = 0; k < nstarts; ++k) {
        const double* st = starts + 6 * k;
        double f[6];
        for (int a = 0; a < 6; ++a) f[a] = (st[a] - (double)g->lo[a]) / (double)g->step[a];
        const double n3 = (double)(float)g->n[2];
        f[2] = f[2] - std::floor(f[2] / n3) * n3;
        const double v = interp<double>(pv, *g, st[0], st[1], st[2], st[3], st[4], st[5], false, cap, false);
        Ls += v;
        if (!(v < cap)) continue;
        Row r;
        row_of(*g, f, 0, r);
        for_corners(*g, r, [&](int64_t j, double w) { gv[j] += w / (double)nstarts; });
    }
    *L = Ls / (double)nstarts;
    for (int64_t j = 0; j < n; ++j) lam[j] = gv[j];
    int32_t it = 0;
    bool done = false;
    for (it = 1; it <= max_iters; ++it) {
        std::fill(nxt.begin(), nxt.end(), 0.0);
        for (int64_t i = 0; i < n; ++i) {
            if (!rows[i].live || lam[i] == 0.0) continue;
            const double li = lam[i];
            for_corners(*g, rows[i], [&](int64_t j, double w) { nxt[j] += w * li; });
        }
        double dmax = 0.0, lmax = 0.0;
        for (int64_t j = 0; j < n; ++j) {
            const double self = rows[j].live ? rows[j].self : 0.0;
            const double v = (gv[j] + nxt[j] - self * lam[j]) / (1.0 - self);
            dmax = std::max(dmax, std::fabs(v - lam[j]));
            lmax = std::max(lmax, std::fabs(v));
            lam[j] = v;
        }
        if (dmax <= tol * std::max(lmax, 1e-300)) { done = true; break; }
    }
    for (int j = 0; j < 30; ++j) grad[j] = 0.0;
    std::vector<double> acc(30 * (size_t)std::max(1, omp_get_max_threads()), 0.0);
#pragma omp parallel
    {
        double* my = acc.data() + 30 * (size_t)omp_get_thread_num();
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < n; ++i) {
            if (!rows[i].live || lam[i] == 0.0) continue;
            double s0[6], f[6], gi[6];
            successor_f64(i, pol[i], occ, *g, *m, ctl, nctl, *p, s0, f);
            interp_grad(pv, *g, f, gi);
            const int32_t k = pol[i];
            for (int j = 0; j < 30; ++j) {
                double s[6], ds[6];
                for (int q = 0; q < 6; ++q) s[q] = s0[q];
                const double tau = (double)p->dt;
                step_state_tangent(*m, s, ds, (double)ctl[k], (double)ctl[k + nctl], (double)ctl[k + 2 * nctl],
                                   tau, substeps_for<double>(tau, *p), j);
                double snh, cnh;
                sincos_t(s[2], snh, cnh);
                const double a1 = ds[3] - s[4] * ds[2], a2 = ds[4] + s[3] * ds[2];
                const double dsv[6] = {ds[0], ds[1], ds[2], cnh * a1 - snh * a2, snh * a1 + cnh * a2, ds[5]};
                double r = 0.0;
                for (int a = 0; a < 6; ++a) r += gi[a] * dsv[a] / (double)g->step[a];
                my[j] += lam[i] * r;
            }
        }
    }
    for (size_t t = 0; t < acc.size() / 30; ++t)
        for (int j = 0; j < 30; ++j) grad[j] += acc[30 * t + j];
    if (pol_out)
        for (int64_t i = 0; i < n; ++i) pol_out[i] = pol[i];
    return done ? it : -1;
}

// Property check of the quotient the HIP kernels use in place of a float
// division by a grid constant b:  q = RN32(RN64(a) * RN64(1/b)).  Compared
// with IEEE a / b over every non-NaN float a; returns the number of values
// that differ (NaN results compare equal to each other).
int64_t orc_check_quotient(float b, int32_t nthreads) {
    const double rb = 1.0 / (double)b;
    int64_t bad = 0;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static) reduction(+ : bad)
    for (int64_t u = 0; u < (int64_t)1 << 32; ++u) {
        uint32_t bits = (uint32_t)u;
        float a;
        std::memcpy(&a, &bits, 4);
        if (std::isnan(a)) continue;
        float ref = a / b;
        float q = (float)((double)a * rb);
        uint32_t rq, rr;
        std::memcpy(&rq, &q, 4);
        std::memcpy(&rr, &ref, 4);
        if (rq != rr && !(std::isnan(q) && std::isnan(ref))) ++bad;
    }
    return bad;
}

}  // extern "C"
