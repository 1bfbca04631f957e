// Launch wrappers shared by the kernels and the C-ABI implementation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "pg_device.hpp"

hipError_t pgk_launch_sweep(const pgk::Consts& K, const float* Vin, float* Vout,
                            const uint8_t* occ, const float* bu, uint8_t* pol, double* part,
                            unsigned* linf_bits, int write_v, hipStream_t st, int64_t* nblocks);
hipError_t pgk_launch_reduce(const double* part, int64_t n, double* out, hipStream_t st);
hipError_t pgk_launch_fill(float* V, int64_t n, float v, hipStream_t st);
hipError_t pgk_launch_seed(float* V, const int64_t* seeds, int64_t n, hipStream_t st);

// LDS-staged path (pg_sweep_lds.hip).
int pgk_window_shapes(const int (**shapes)[6]);
hipError_t pgk_launch_plan(const pgk::Consts& K, const pgk::BoxGeom& G, const uint8_t* occ,
                           const float* bu, int16_t* worig, uint8_t* wext, int diag,
                           hipStream_t st);
hipError_t pgk_launch_sweep_lds(int shape, int diag, const pgk::Consts& K, const pgk::BoxGeom& G,
                                const float* Vin, float* Vout, const uint8_t* occ, const float* bu,
                                const int16_t* worig, uint8_t* pol, double* part, unsigned* linf,
                                int write_v, hipStream_t st);
