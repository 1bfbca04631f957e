// The device scratch of pg_sweep_noalloc: one statically sized array per
// device, so that entry point allocates nothing.
//
// Layout (byte offsets, each 256-aligned):
//   0        double sum + the two-level reduction's kReduceBlocks partials
//   8448     unsigned ||dV||inf bits
//   8704     float B * u_sat, 3 per control (at most 255 controls)
//   11776    float raw commands, 3 per control
//   16384    double per-workgroup partial sums of the sweep, to the end
// Every call on a device shares its copy and leaves its contents behind;
// each call rewrites what it reads before reading it (the tables by copy, the
// sup-norm word by zero_u32_kernel, the partials by the sweep kernel).
#include <hip/hip_runtime.h>
#include "pg_kernels.h"
#include "../../include/peregrine_hip.h"

namespace {

__device__ __align__(256) unsigned char scratch[PG_NOALLOC_SCRATCH_BYTES];

__global__ void zero_u32_kernel(unsigned* p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0u;
}

}  // namespace

// The device address of the current device's scratch (the array's name is
// only a host-side stand-in). Called once per device, with it current.
hipError_t pgk_noalloc_scratch(void** ptr) {
    *ptr = nullptr;
    return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch));
}

hipError_t pgk_launch_zero_u32(unsigned* p, hipStream_t st) {
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, st, p);
    return hipGetLastError();
}
