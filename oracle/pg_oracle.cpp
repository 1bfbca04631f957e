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
    return p->rounds <= 0 && p->ntau <= 1;
}

template <class T>
int sweep_impl(const T* Vin, T* Vout, const uint8_t* occ, const pg_grid* g, const pg_model* m,
               const float* ctl, int32_t nctl, const pg_params* p, int32_t mode, int32_t rev,
               int32_t phase, int32_t nthreads, uint8_t* pol, double* sum_out, double* linf_out,
               int32_t esc) {
    if (!params_supported(p) || nctl < 2) return -3;
    const int64_t n = ncells(*g);
    double sum = 0.0, linf = 0.0;
    if (mode == 0) {
        if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 512) reduction(+ : sum) reduction(max : linf)
        for (int64_t i = 0; i < n; ++i) {
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
            T c = cell_update<T>(i, Vin, occ, *g, *m, ctl, nctl, *p, phase, false, &k);
            if (c < old) {
                Vout[i] = c;
                sum += (double)(old - c);
                linf = std::max(linf, (double)(old - c));
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
            int64_t idx = rev ? (n - 1 - ii) : ii;
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
            T c = cell_update<T>(idx, V, occ, *g, *m, ctl, nctl, *p, phase, false, &k);
            if (c < old) {
                V[idx] = c;
                s += (double)(old - c);
                l = std::max(l, (double)(old - c));
                if (pol) pol[idx] = (uint8_t)k;
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
