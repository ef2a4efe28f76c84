// mlp.hip -- the learned Newton start: packing of the bf16 operands, inference for tests, and
// training on the GPU (forward on the matrix cores as in the Newton kernel, backward in FP32,
// Adam on FP32 master weights).  The network has no counterpart in the reference; its inputs
// and outputs are defined in mlp_dev.hpp.
#include "mlp_dev.hpp"
#include "srx_dev.hpp"
#include "../../include/srx.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

namespace srx {

__device__ __forceinline__ float bf16r(float v) { return (float)(__bf16)v; }
__device__ __forceinline__ bool mlp_is_weight(int p) {
    return p < MLP_OFF_B1 || (p >= MLP_OFF_W2 && p < MLP_OFF_B2) || (p >= MLP_OFF_W3 && p < MLP_OFF_B3);
}

// Builds the MFMA operand layout of mlp_dev.hpp from an FP32 blob.
__global__ void k_mlp_pack(const float *__restrict__ w, MlpPacked *__restrict__ M) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 2048) {   // a1 [rb][s][lane][e]
        const int e = t & 7, lane = (t >> 3) & 63, s = (t >> 9) & 1, rb = t >> 10;
        const int m = 32 * rb + (lane & 31), k = 16 * s + 8 * (lane >> 5) + e;
        M->a1[rb][s][lane][e] = (__bf16)(k < MLP_IN ? w[MLP_OFF_W1 + k * MLP_H + m] : 0.0f);
    } else if (t < 2048 + 4096) {   // a2 [rb2][ks][lane][e]
        const int u0 = t - 2048;
        const int e = u0 & 7, lane = (u0 >> 3) & 63, ks = (u0 >> 9) & 3, rb2 = u0 >> 11;
        const int m = 32 * rb2 + (lane & 31), u = mlp_k_unit(ks >> 1, ks & 1, lane >> 5, e);
        M->a2[rb2][ks][lane][e] = (__bf16)w[MLP_OFF_W2 + u * MLP_H + m];
    } else if (t < 2048 + 4096 + 2048) {   // a3 [ks][lane][e]
        const int u0 = t - 6144;
        const int e = u0 & 7, lane = (u0 >> 3) & 63, ks = u0 >> 9;
        const int m = lane & 31, u = mlp_k_unit(ks >> 1, ks & 1, lane >> 5, e);
        M->a3[ks][lane][e] = (__bf16)(m < MLP_OUT ? w[MLP_OFF_W3 + u * MLP_OUT + m] : 0.0f);
    } else if (t < 8192 + 64) {   // b1, b2 [rb][g][r]
        const int u0 = t - 8192;
        const int r = u0 & 15, g = (u0 >> 4) & 1, rb = u0 >> 5;
        M->b1[rb][g][r] = w[MLP_OFF_B1 + mlp_acc_unit(rb, g, r)];
        M->b2[rb][g][r] = w[MLP_OFF_B2 + mlp_acc_unit(rb, g, r)];
    } else if (t < 8256 + MLP_OUT) {
        M->b3[t - 8256] = w[MLP_OFF_B3 + t - 8256];
    } else if (t < 8262 + MLP_IN) {
        M->xm[t - 8262] = w[MLP_OFF_XM + t - 8262];
        M->xs[t - 8262] = w[MLP_OFF_XS + t - 8262];
    } else if (t < 8286 + MLP_OUT) {
        M->tm[t - 8286] = w[MLP_OFF_TM + t - 8286];
        M->ts[t - 8286] = w[MLP_OFF_TS + t - 8286];
    }
    if (t < MLP_NPARAM) M->wq[t] = mlp_is_weight(t) ? bf16r(w[t]) : w[t];
}
constexpr int MLP_PACK_THREADS = 8292 > MLP_NPARAM ? 8292 : MLP_NPARAM;

void mlp_pack(hipStream_t s, const float *dblob, MlpPacked *M) {
    hipLaunchKernelGGL(k_mlp_pack, dim3((MLP_PACK_THREADS + 255) / 256), dim3(256), 0, s, dblob, M);
}
size_t mlp_packed_bytes() { return sizeof(MlpPacked); }

// y = network(normalise(x)) * spread + mean for n samples (features [n][24], outputs [n][6]).
__global__ void __launch_bounds__(64) k_mlp_infer(const MlpPacked *__restrict__ M, const float *__restrict__ x,
                                                  int n, float *__restrict__ y) {
    const int j = blockIdx.x * 64 + threadIdx.x;
    const bool valid = j < n;
    float xn[MLP_IN], o[MLP_OUT];
#pragma unroll
    for (int k = 0; k < MLP_IN; ++k) xn[k] = valid ? x[(size_t)j * MLP_IN + k] : 0.0f;
    mlp_normalise(M, xn);
    MlpAcc unused;
    mlp_forward_wave<false>(M, xn, o, unused);
    if (valid)
#pragma unroll
        for (int k = 0; k < MLP_OUT; ++k) y[(size_t)j * MLP_OUT + k] = o[k] * M->ts[k] + M->tm[k];
}

// One training wavefront: forward for its 64 samples (indices idx[base + ...], or base + ... when
// idx is null), loss sum((y - t')^2) over normalised targets t' = (t - mean) / spread, and the
// gradient of scale * loss, written to part[block][0, MLP_NPARAM) with the loss at MLP_NPARAM.
// Straight-through bf16: the backward pass uses the same rounded weights, inputs and hidden
// values as the forward pass, in FP32 arithmetic.
constexpr int MLP_P = MLP_H + 1;   // LDS pitch of the [sample][unit] arrays
__global__ void __launch_bounds__(64) k_mlp_fwdbwd(const MlpPacked *__restrict__ M, const float *__restrict__ x,
                                                   const float *__restrict__ t, const int *__restrict__ idx,
                                                   int base, int nb, float scale, float *__restrict__ part) {
    __shared__ float h1l[64 * MLP_P], h2l[64 * MLP_P], d2l[64 * MLP_P];
    __shared__ float xq[64 * (MLP_IN + 1)], dyl[64 * (MLP_OUT + 1)];
    float *d1l = h2l;   // dz1 replaces h2 once the W3 gradient has been taken
    const int lane = threadIdx.x;
    const int j = blockIdx.x * 64 + lane;
    const bool valid = j < nb;
    const size_t si = valid ? (size_t)(idx ? idx[base + j] : base + j) : 0;
    float xn[MLP_IN], y[MLP_OUT];
#pragma unroll
    for (int k = 0; k < MLP_IN; ++k) xn[k] = valid ? x[si * MLP_IN + k] : 0.0f;
    mlp_normalise(M, xn);
    MlpAcc z;
    mlp_forward_wave<true>(M, xn, y, z);
    const int g = lane >> 5;
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int cb = 0; cb < 2; ++cb)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int s = 32 * cb + (lane & 31), u = mlp_acc_unit(rb, g, r);
                h1l[s * MLP_P + u] = bf16r(fmaxf(z.z1[rb][cb][r], 0.0f));
                h2l[s * MLP_P + u] = bf16r(fmaxf(z.z2[rb][cb][r], 0.0f));
            }
#pragma unroll
    for (int k = 0; k < MLP_IN; ++k) xq[lane * (MLP_IN + 1) + k] = bf16r(xn[k]);
    float dy[MLP_OUT], loss = 0.0f;
#pragma unroll
    for (int o = 0; o < MLP_OUT; ++o) {
        const float tn = valid ? (t[si * MLP_OUT + o] - M->tm[o]) / M->ts[o] : 0.0f;
        const float e = y[o] - tn;
        dy[o] = valid ? scale * e : 0.0f;
        loss += valid ? e * e : 0.0f;
        dyl[lane * (MLP_OUT + 1) + o] = dy[o];
    }
    __syncthreads();
    float *out = part + (size_t)blockIdx.x * (MLP_NPARAM + 1);
    const float *wq = M->wq;
    {   // W3 [u][o] and b3, lane = hidden unit u of layer 2
        float acc[MLP_OUT] = {0, 0, 0, 0, 0, 0};
        for (int s = 0; s < 64; ++s) {
            const float h = h2l[s * MLP_P + lane];
#pragma unroll
            for (int o = 0; o < MLP_OUT; ++o) acc[o] += h * dyl[s * (MLP_OUT + 1) + o];
        }
#pragma unroll
        for (int o = 0; o < MLP_OUT; ++o) out[MLP_OFF_W3 + lane * MLP_OUT + o] = acc[o];
        if (lane < MLP_OUT) {
            float b = 0.0f;
            for (int s = 0; s < 64; ++s) b += dyl[s * (MLP_OUT + 1) + lane];
            out[MLP_OFF_B3 + lane] = b;
        }
    }
    // lane = sample: dz2 = (W3 dy) o [z2 > 0], dz1 = (W2 dz2) o [z1 > 0]
    float dz2[MLP_H];
#pragma unroll
    for (int u = 0; u < MLP_H; ++u) {
        float a = 0.0f;
#pragma unroll
        for (int o = 0; o < MLP_OUT; ++o) a += wq[MLP_OFF_W3 + u * MLP_OUT + o] * dy[o];
        dz2[u] = h2l[lane * MLP_P + u] > 0.0f ? a : 0.0f;
        d2l[lane * MLP_P + u] = dz2[u];
    }
    __syncthreads();   // every lane has read h2l before it is overwritten by dz1
    for (int i = 0; i < MLP_H; ++i) {
        float a = 0.0f;
#pragma unroll
        for (int u = 0; u < MLP_H; ++u) a += wq[MLP_OFF_W2 + i * MLP_H + u] * dz2[u];
        d1l[lane * MLP_P + i] = h1l[lane * MLP_P + i] > 0.0f ? a : 0.0f;
    }
    __syncthreads();
    {   // W2 [i][u] and b2, lane = unit u of layer 2
        float acc[MLP_H];
#pragma unroll
        for (int i = 0; i < MLP_H; ++i) acc[i] = 0.0f;
        float b = 0.0f;
        for (int s = 0; s < 64; ++s) {
            const float d = d2l[s * MLP_P + lane];
            b += d;
#pragma unroll
            for (int i = 0; i < MLP_H; ++i) acc[i] += h1l[s * MLP_P + i] * d;
        }
#pragma unroll
        for (int i = 0; i < MLP_H; ++i) out[MLP_OFF_W2 + i * MLP_H + lane] = acc[i];
        out[MLP_OFF_B2 + lane] = b;
    }
    {   // W1 [k][u] and b1, lane = unit u of layer 1
        float acc[MLP_IN];
#pragma unroll
        for (int k = 0; k < MLP_IN; ++k) acc[k] = 0.0f;
        float b = 0.0f;
        for (int s = 0; s < 64; ++s) {
            const float d = d1l[s * MLP_P + lane];
            b += d;
#pragma unroll
            for (int k = 0; k < MLP_IN; ++k) acc[k] += xq[s * (MLP_IN + 1) + k] * d;
        }
#pragma unroll
        for (int k = 0; k < MLP_IN; ++k) out[MLP_OFF_W1 + k * MLP_H + lane] = acc[k];
        out[MLP_OFF_B1 + lane] = b;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o, 64);
    if (lane == 0) out[MLP_NPARAM] = loss;
}

// Sum of the block partials in a fixed order (gradient and loss), then optionally one Adam step
// on the FP32 master weights w (lr > 0).
__global__ void k_mlp_adam(const float *__restrict__ part, int nparts, float *__restrict__ w, float *__restrict__ m,
                           float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2,
                           float *__restrict__ loss_out, float *__restrict__ grad_out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > MLP_NPARAM) return;
    float gsum = 0.0f;
    for (int b = 0; b < nparts; ++b) gsum += part[(size_t)b * (MLP_NPARAM + 1) + p];
    if (p == MLP_NPARAM) {
        if (loss_out) *loss_out = gsum;
        return;
    }
    if (grad_out) grad_out[p] = gsum;
    if (lr > 0.0f) {
        const float mm = b1 * m[p] + (1.0f - b1) * gsum;
        const float vv = b2 * v[p] + (1.0f - b2) * gsum * gsum;
        m[p] = mm;
        v[p] = vv;
        w[p] -= lr * (mm * c1) / (sqrtf(vv * c2) + eps);
    }
}

}  // namespace srx

using namespace srx;

namespace {

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    template <typename T>
    T *as() const { return static_cast<T *>(p); }
};

#define MLP_CHK(x)                                                                                   \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) {                                                                      \
            fprintf(stderr, "srx: %s failed: %s\n", #x, hipGetErrorString(e_));                      \
            return SRX_E_HIP;                                                                        \
        }                                                                                            \
    } while (0)

int alloc(DevBuf &b, size_t bytes) {
    if (hipMalloc(&b.p, bytes ? bytes : 4) != hipSuccess) {
        b.p = nullptr;
        return SRX_E_NOMEM;
    }
    return 0;
}

int check_blob(const float *blob) {
    for (int k = 0; k < MLP_OUT; ++k)
        if (!(blob[MLP_OFF_TS + k] != 0.0f)) return SRX_E_ARG;
    return 0;
}

// one forward/backward over samples [base, base + nb) (or idx[base ...]), summed into grad/loss
int grad_pass(hipStream_t s, const MlpPacked *M, const float *x, const float *t, const int *idx, int base, int nb,
              float scale, float *part, float *w, float *m, float *v, float lr, float c1, float c2, float *loss,
              float *grad) {
    const int nblk = (nb + 63) / 64;
    hipLaunchKernelGGL(k_mlp_fwdbwd, dim3(nblk), dim3(64), 0, s, M, x, t, idx, base, nb, scale, part);
    hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + 1 + 255) / 256), dim3(256), 0, s, part, nblk, w, m, v, lr,
                       0.9f, 0.999f, 1e-8f, c1, c2, loss, grad);
    return hipGetLastError() == hipSuccess ? 0 : SRX_E_HIP;
}

}  // namespace

extern "C" {

int srx_mlp_forward(const float *blob, const float *x, size_t n, float *y) {
    if (!blob || (!x && n) || (!y && n) || n > (size_t)1 << 30) return SRX_E_ARG;
    if (int rc = check_blob(blob)) return rc;
    if (n == 0) return 0;
    DevBuf dw, dM, dx, dy;
    if (alloc(dw, MLP_BLOB * 4) || alloc(dM, sizeof(MlpPacked)) || alloc(dx, n * MLP_IN * 4) ||
        alloc(dy, n * MLP_OUT * 4))
        return SRX_E_NOMEM;
    MLP_CHK(hipMemcpy(dw.p, blob, MLP_BLOB * 4, hipMemcpyHostToDevice));
    MLP_CHK(hipMemcpy(dx.p, x, n * MLP_IN * 4, hipMemcpyHostToDevice));
    mlp_pack(0, dw.as<float>(), dM.as<MlpPacked>());
    hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, 0, dM.as<MlpPacked>(),
                       dx.as<float>(), (int)n, dy.as<float>());
    MLP_CHK(hipGetLastError());
    MLP_CHK(hipMemcpy(y, dy.p, n * MLP_OUT * 4, hipMemcpyDeviceToHost));
    return 0;
}

int srx_mlp_grad(const float *blob, const float *x, const float *t, size_t n, float *grad, double *loss) {
    if (!blob || !x || !t || !grad || !loss || n == 0 || n > (size_t)1 << 24) return SRX_E_ARG;
    if (int rc = check_blob(blob)) return rc;
    const int nblk = (int)((n + 63) / 64);
    DevBuf dw, dM, dx, dt, dpart, dg, dl;
    if (alloc(dw, MLP_BLOB * 4) || alloc(dM, sizeof(MlpPacked)) || alloc(dx, n * MLP_IN * 4) ||
        alloc(dt, n * MLP_OUT * 4) || alloc(dpart, (size_t)nblk * (MLP_NPARAM + 1) * 4) ||
        alloc(dg, MLP_NPARAM * 4) || alloc(dl, 4))
        return SRX_E_NOMEM;
    MLP_CHK(hipMemcpy(dw.p, blob, MLP_BLOB * 4, hipMemcpyHostToDevice));
    MLP_CHK(hipMemcpy(dx.p, x, n * MLP_IN * 4, hipMemcpyHostToDevice));
    MLP_CHK(hipMemcpy(dt.p, t, n * MLP_OUT * 4, hipMemcpyHostToDevice));
    mlp_pack(0, dw.as<float>(), dM.as<MlpPacked>());
    const float scale = 1.0f / (float)(n * MLP_OUT);
    if (grad_pass(0, dM.as<MlpPacked>(), dx.as<float>(), dt.as<float>(), nullptr, 0, (int)n, scale,
                  dpart.as<float>(), dw.as<float>(), nullptr, nullptr, 0.0f, 1.0f, 1.0f, dl.as<float>(),
                  dg.as<float>()))
        return SRX_E_HIP;
    // d/dw of mean (y - t')^2 is 2 * scale * sum (y - t') dy/dw: the kernel applies `scale` once
    std::vector<float> g(MLP_NPARAM);
    float l = 0.0f;
    MLP_CHK(hipMemcpy(g.data(), dg.p, MLP_NPARAM * 4, hipMemcpyDeviceToHost));
    MLP_CHK(hipMemcpy(&l, dl.p, 4, hipMemcpyDeviceToHost));
    for (int p = 0; p < MLP_NPARAM; ++p) grad[p] = 2.0f * g[p];
    *loss = (double)l / (double)(n * MLP_OUT);
    return 0;
}

void srx_mlp_default_train_options(srx_mlp_train_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->epochs = 20;
    o->batch = 4096;
    o->lr = 2e-3f;
    o->lr_final = 1e-4f;
    o->seed = 1234;
    o->init = 1;
}

int srx_mlp_train(const float *x, const float *t, size_t n, const srx_mlp_train_opts *opt, float *blob,
                  float *loss_hist) {
    srx_mlp_train_opts o;
    if (opt) o = *opt;
    else srx_mlp_default_train_options(&o);
    if (!x || !t || !blob || n < 64 || n > (size_t)1 << 27 || o.epochs < 1 || o.batch < 64 || o.batch % 64 ||
        !(o.lr > 0.0f))
        return SRX_E_ARG;
    std::mt19937_64 rng(o.seed);
    if (o.init) {
        // feature and target statistics, then He-uniform weights and zero biases
        for (int k = 0; k < MLP_IN; ++k) {
            double s = 0.0, s2 = 0.0;
            for (size_t i = 0; i < n; ++i) s += x[i * MLP_IN + k];
            const double mu = s / (double)n;
            for (size_t i = 0; i < n; ++i) s2 += (x[i * MLP_IN + k] - mu) * (x[i * MLP_IN + k] - mu);
            const double sd = std::sqrt(s2 / (double)n);
            blob[MLP_OFF_XM + k] = (float)mu;
            blob[MLP_OFF_XS + k] = sd > 1e-6 ? (float)(1.0 / sd) : 1.0f;
        }
        for (int k = 0; k < MLP_OUT; ++k) {
            double s = 0.0, s2 = 0.0;
            for (size_t i = 0; i < n; ++i) s += t[i * MLP_OUT + k];
            const double mu = s / (double)n;
            for (size_t i = 0; i < n; ++i) s2 += (t[i * MLP_OUT + k] - mu) * (t[i * MLP_OUT + k] - mu);
            const double sd = std::sqrt(s2 / (double)n);
            blob[MLP_OFF_TM + k] = (float)mu;
            blob[MLP_OFF_TS + k] = sd > 1e-12 ? (float)sd : 1.0f;
        }
        auto fill = [&](int off, int fan_in, int count) {
            const double lim = std::sqrt(6.0 / fan_in);
            std::uniform_real_distribution<double> U(-lim, lim);
            for (int i = 0; i < count; ++i) blob[off + i] = (float)U(rng);
        };
        fill(MLP_OFF_W1, MLP_IN, MLP_IN * MLP_H);
        fill(MLP_OFF_W2, MLP_H, MLP_H * MLP_H);
        fill(MLP_OFF_W3, MLP_H, MLP_H * MLP_OUT);
        for (int i = 0; i < MLP_H; ++i) blob[MLP_OFF_B1 + i] = blob[MLP_OFF_B2 + i] = 0.0f;
        for (int i = 0; i < MLP_OUT; ++i) blob[MLP_OFF_B3 + i] = 0.0f;
    }
    if (int rc = check_blob(blob)) return rc;
    const int B = o.batch;
    const size_t steps_per_epoch = (n + B - 1) / B;
    const int nblk = B / 64;
    DevBuf dw, dM, dx, dt, didx, dpart, dm, dv, dl;
    if (alloc(dw, MLP_BLOB * 4) || alloc(dM, sizeof(MlpPacked)) || alloc(dx, n * MLP_IN * 4) ||
        alloc(dt, n * MLP_OUT * 4) || alloc(didx, n * 4) || alloc(dpart, (size_t)nblk * (MLP_NPARAM + 1) * 4) ||
        alloc(dm, MLP_NPARAM * 4) || alloc(dv, MLP_NPARAM * 4) ||
        alloc(dl, (size_t)o.epochs * steps_per_epoch * 4))
        return SRX_E_NOMEM;
    hipStream_t s;
    MLP_CHK(hipStreamCreate(&s));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{s};
    MLP_CHK(hipMemcpyAsync(dw.p, blob, MLP_BLOB * 4, hipMemcpyHostToDevice, s));
    MLP_CHK(hipMemcpyAsync(dx.p, x, n * MLP_IN * 4, hipMemcpyHostToDevice, s));
    MLP_CHK(hipMemcpyAsync(dt.p, t, n * MLP_OUT * 4, hipMemcpyHostToDevice, s));
    MLP_CHK(hipMemsetAsync(dm.p, 0, MLP_NPARAM * 4, s));
    MLP_CHK(hipMemsetAsync(dv.p, 0, MLP_NPARAM * 4, s));
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    const size_t total = (size_t)o.epochs * steps_per_epoch;
    size_t step = 0;
    double b1t = 1.0, b2t = 1.0;
    const double lr0 = o.lr, lr1 = o.lr_final > 0.0f ? o.lr_final : o.lr;
    for (int ep = 0; ep < o.epochs; ++ep) {
        std::shuffle(idx.begin(), idx.end(), rng);
        // the index upload must complete before the host reshuffles for the next epoch
        MLP_CHK(hipMemcpyAsync(didx.p, idx.data(), n * 4, hipMemcpyHostToDevice, s));
        for (size_t b0 = 0; b0 < n; b0 += B, ++step) {
            const int nb = (int)std::min((size_t)B, n - b0);
            if (step == 0) mlp_pack(s, dw.as<float>(), dM.as<MlpPacked>());
            b1t *= 0.9;
            b2t *= 0.999;
            // cosine decay from lr to lr_final
            const double frac = total > 1 ? (double)step / (double)(total - 1) : 1.0;
            const float lr = (float)(lr1 + 0.5 * (lr0 - lr1) * (1.0 + std::cos(M_PI * frac)));
            if (grad_pass(s, dM.as<MlpPacked>(), dx.as<float>(), dt.as<float>(), didx.as<int>(), (int)b0, nb,
                          1.0f / (float)(nb * MLP_OUT), dpart.as<float>(), dw.as<float>(), dm.as<float>(),
                          dv.as<float>(), lr, (float)(1.0 / (1.0 - b1t)), (float)(1.0 / (1.0 - b2t)),
                          dl.as<float>() + step, nullptr))
                return SRX_E_HIP;
            mlp_pack(s, dw.as<float>(), dM.as<MlpPacked>());
        }
        MLP_CHK(hipStreamSynchronize(s));
    }
    MLP_CHK(hipMemcpyAsync(blob, dw.p, MLP_NPARAM * 4, hipMemcpyDeviceToHost, s));
    if (loss_hist) {
        std::vector<float> lh(total);
        MLP_CHK(hipMemcpyAsync(lh.data(), dl.p, total * 4, hipMemcpyDeviceToHost, s));
        MLP_CHK(hipStreamSynchronize(s));
        // per epoch: mean squared error of the normalised outputs over the epoch's samples
        for (int ep = 0; ep < o.epochs; ++ep) {
            double acc = 0.0;
            for (size_t k = 0; k < steps_per_epoch; ++k) acc += lh[ep * steps_per_epoch + k];
            loss_hist[ep] = (float)(acc / ((double)n * MLP_OUT));
        }
    }
    MLP_CHK(hipStreamSynchronize(s));
    return 0;
}

}  // extern "C"
