// noalloc.hip -- srx_fft_forward_noalloc: the forward FFT of srx_fft_forward without any allocation,
// for callers that cannot allocate device or host memory (code under HIP graph capture, serving
// loops, harnesses that forbid hipMalloc).  Input and output are the caller's device buffers; the
// two half-spectrum buffers and the twiddle tables the passes need live in one static device array
// per GPU (`scratch`, SRX_NOALLOC_SCRATCH_BYTES), held until the process ends.
//
// Host code never passes `scratch` itself (in host code that name is a host-side stand-in, not a
// device address): the first call on each device looks its address up with hipGetSymbolAddress,
// with that device current, and keeps it per device.  Every call on a GPU shares that GPU's array,
// so calls that overlap on one GPU race; nothing in it needs to start at zero (every pass writes
// each element it later reads).
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "srx_dev.hpp"
#include "../../include/srx.h"

namespace srx {

static __device__ __align__(256) unsigned char scratch[SRX_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
void *g_scratch_ptr[MAX_DEVICES] = {};
std::mutex g_scratch_mu;

// Device address of this device's `scratch` (looked up once per device, device current).
int scratch_ptr(void **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return SRX_E_HIP;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return SRX_E_HIP;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// e^{-2 pi i m / n}, m < n, as (re, im) in T
template <typename T>
__global__ void k_twiddles(T *tw, int n) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    double s, c;
    sincospi(-2.0 * (double)m / (double)n, &s, &c);
    tw[2 * m] = (T)c;
    tw[2 * m + 1] = (T)s;
}

// layout C [c][kz][y][x] (complex T) -> out [c][x][y][kz] (complex double), the order of
// srx_fft_forward's output
template <typename T>
__global__ void k_spec_out(const T *__restrict__ specC, double *__restrict__ out, int ncomp, int n0, int n1, int H) {
    const size_t total = (size_t)ncomp * n0 * n1 * H;
    for (size_t d = blockIdx.x * (size_t)blockDim.x + threadIdx.x; d < total; d += (size_t)gridDim.x * blockDim.x) {
        const size_t kz = d % H;
        const size_t y = (d / H) % n1;
        const size_t x = (d / ((size_t)H * n1)) % n0;
        const size_t c = d / ((size_t)H * n1 * n0);
        const size_t src = ((c * H + kz) * n1 + y) * n0 + x;
        out[2 * d] = (double)specC[2 * src];
        out[2 * d + 1] = (double)specC[2 * src + 1];
    }
}

}  // namespace

}  // namespace srx

extern "C" int srx_fft_forward_noalloc(const double *d_in, int ncomp, int n0, int n1, int n2, int precision,
                                       double *d_out) {
    using namespace srx;
    if (!d_in || !d_out || ncomp < 1 || ncomp > 9) return SRX_E_ARG;
    if (precision != SRX_PREC_FP64 && precision != SRX_PREC_MIXED) return SRX_E_ARG;
    const int n[3] = {n0, n1, n2};
    FftPlan pl[3];
    for (int i = 0; i < 3; ++i)
        if (!make_plan(n[i], pl[i])) return SRX_E_GRID;
    if (!srx_grid_supported(pl, n, precision)) return SRX_E_GRID;
    const size_t csz = precision == SRX_PREC_MIXED ? 8 : 16;
    const size_t H = (size_t)n2 / 2 + 1;
    const size_t spec = align256((size_t)ncomp * H * n0 * n1 * csz);
    size_t off[3], need = 0;
    for (int ax = 0; ax < 3; ++ax) {
        off[ax] = need;
        need += align256((size_t)n[ax] * csz);
    }
    const size_t offB = need, offC = need + spec;
    need += 2 * spec;
    if (need > (size_t)SRX_NOALLOC_SCRATCH_BYTES) return SRX_E_SCRATCH;
    void *base = nullptr;
    if (int rc = scratch_ptr(&base)) return rc;
    unsigned char *b = static_cast<unsigned char *>(base);
    void *tw[3];
    for (int ax = 0; ax < 3; ++ax) {
        tw[ax] = b + off[ax];
        const dim3 grid((n[ax] + 255) / 256);
        if (precision == SRX_PREC_MIXED) hipLaunchKernelGGL(k_twiddles<float>, grid, dim3(256), 0, nullptr, (float *)tw[ax], n[ax]);
        else hipLaunchKernelGGL(k_twiddles<double>, grid, dim3(256), 0, nullptr, (double *)tw[ax], n[ax]);
    }
    void *sB = b + offB, *sC = b + offC;
    launch_fft_real_forward_test(nullptr, precision, d_in, ncomp, sB, sC, n, pl, tw);
    const size_t total = (size_t)ncomp * n0 * n1 * H;
    const dim3 grid((unsigned)std::min<size_t>((total + 255) / 256, 4096));
    if (precision == SRX_PREC_MIXED)
        hipLaunchKernelGGL(k_spec_out<float>, grid, dim3(256), 0, nullptr, (const float *)sC, d_out, ncomp, n0, n1, (int)H);
    else
        hipLaunchKernelGGL(k_spec_out<double>, grid, dim3(256), 0, nullptr, (const double *)sC, d_out, ncomp, n0, n1, (int)H);
    if (hipGetLastError() != hipSuccess) return SRX_E_HIP;
    if (hipStreamSynchronize(nullptr) != hipSuccess) return SRX_E_HIP;
    return 0;
}
