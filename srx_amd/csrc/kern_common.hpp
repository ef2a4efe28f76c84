// kern_common.hpp -- helpers shared by the kernel files (fft.hip, newton.hip, increment.hip):
// the launch macro with optional per-dispatch events, launch error reporting, wavefront
// reductions and the B-basis constants.
#pragma once
#include "srx_dev.hpp"

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <cmath>
#include <cstdio>

namespace srx {

// Optional start/stop events for the next launch (set by set_launch_events, defined in fft.hip):
// when set, the launch goes through hipExtLaunchKernelGGL, which stamps the events on the
// kernel's own dispatch packet.
extern thread_local hipEvent_t g_ev_start, g_ev_stop;
#define SRX_LAUNCH(kernel, grid, block, shmem, stream, ...)                                              \
    do {                                                                                                 \
        if (g_ev_start) {                                                                                \
            hipExtLaunchKernelGGL(kernel, grid, block, shmem, stream, g_ev_start, g_ev_stop, 0, __VA_ARGS__); \
            g_ev_start = g_ev_stop = nullptr;                                                            \
        } else {                                                                                         \
            hipLaunchKernelGGL(kernel, grid, block, shmem, stream, __VA_ARGS__);                         \
        }                                                                                                \
    } while (0)

inline void check_launch(const char *name) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) fprintf(stderr, "srx: launch of %s failed: %s\n", name, hipGetErrorString(e));
}

// B-basis constants: 1/sqrt(2), 1/sqrt(3), 1/sqrt(6) (chg_basis, kinematics.c:55-60)
#define SRX_R2 0.70710678118654744
#define SRX_R3 0.57735026918962584
#define SRX_R6 0.40824829046386304

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
    return v;
}

}  // namespace srx
