// Learned Newton start: a 24-64-64-6 perceptron (ReLU) evaluated in bf16 on the matrix cores
// with FP32 accumulation, one wavefront for 64 voxels (or training samples) at a time.
//
// Every layer is computed transposed, Z^T (units x voxels) = W^T (units x K) . X^T (K x voxels),
// with v_mfma_f32_32x32x16_bf16.  In that form the accumulator layout of one layer (lane l holds
// column l % 32 and rows 8 (r / 4) + 4 (l / 32) + r % 4 of its 32x32 tile) is, register for
// register, a valid B operand of the next layer (lane l holds column l % 32 and eight K rows):
// the K order of each 16-row step is a permutation of the hidden units, and the weight operand is
// packed in the same permuted order, so no data moves between layers.  Only the first layer's
// input (one voxel per lane) and the last layer's output need one exchange between lane halves.
#pragma once
#include <hip/hip_runtime.h>

namespace srx {

constexpr int MLP_IN = 24, MLP_H = 64, MLP_OUT = 6;
// FP32 parameter blob: W1 [24][64], b1 [64], W2 [64][64], b2 [64], W3 [64][6], b3 [6]
// (W[k][i]: from input/unit k to unit i), then feature mean [24] and inverse spread [24],
// target mean [6] and spread [6].
constexpr int MLP_OFF_W1 = 0, MLP_OFF_B1 = MLP_OFF_W1 + MLP_IN * MLP_H;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H, MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H, MLP_OFF_B3 = MLP_OFF_W3 + MLP_H * MLP_OUT;
constexpr int MLP_NPARAM = MLP_OFF_B3 + MLP_OUT;   // 6150 trained values
constexpr int MLP_OFF_XM = MLP_NPARAM, MLP_OFF_XS = MLP_OFF_XM + MLP_IN;
constexpr int MLP_OFF_TM = MLP_OFF_XS + MLP_IN, MLP_OFF_TS = MLP_OFF_TM + MLP_OUT;
constexpr int MLP_BLOB = MLP_OFF_TS + MLP_OUT;     // 6210 floats

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Weights in the per-lane operand order of the MFMA tiles (built on the device from the blob).
struct MlpPacked {
    bf16x8 a1[2][2][64];   // W1^T: [row block][k step][lane]
    bf16x8 a2[2][4][64];   // W2^T: [row block][k step = 2 * (layer-1 row block) + half][lane]
    bf16x8 a3[4][64];      // W3^T: rows 0..5 of one row block, [k step][lane]
    float b1[2][2][16];    // bias per [row block][lane half][accumulator register]
    float b2[2][2][16];
    float b3[MLP_OUT];
    float xm[MLP_IN], xs[MLP_IN], tm[MLP_OUT], ts[MLP_OUT];
    float wq[MLP_NPARAM];  // the blob's weights rounded to bf16 (as FP32), for the backward pass
};

// hidden unit held in accumulator register r by lane half g of a 32-row block rb
__host__ __device__ constexpr int mlp_acc_unit(int rb, int g, int r) { return 32 * rb + 8 * (r / 4) + 4 * g + r % 4; }
// hidden unit behind element e of the B operand of k step (rb, h) in lane half g
__host__ __device__ constexpr int mlp_k_unit(int rb, int h, int g, int e) {
    return 32 * rb + 16 * h + 8 * (e / 4) + 4 * g + e % 4;
}

__device__ __forceinline__ bf16x8 mlp_swap_halves(bf16x8 v) {
    union {
        bf16x8 b;
        int i[4];
    } u;
    u.b = v;
#pragma unroll
    for (int k = 0; k < 4; ++k) u.i[k] = __shfl_xor(u.i[k], 32, 64);
    return u.b;
}

__device__ __forceinline__ f32x16 mlp_mfma(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// Pre-activations of both hidden layers, kept by the training pass: [row block][column block].
struct MlpAcc {
    f32x16 z1[2][2];
    f32x16 z2[2][2];
};

// Forward pass of the wavefront: lane l supplies the normalised features xn of its voxel and
// receives the normalised outputs y.  Every lane of the wavefront must execute this call.
template <bool KEEP>
__device__ __forceinline__ void mlp_forward_wave(const MlpPacked *__restrict__ M, const float xn[MLP_IN],
                                                 float y[MLP_OUT], MlpAcc &keep) {
    const int lane = threadIdx.x & 63;
    const bool lo = lane < 32;
    const int g = lane >> 5;
    bf16x8 G[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int e = 0; e < 8; ++e) G[q][e] = (__bf16)(8 * q + e < MLP_IN ? xn[8 * q + e] : 0.0f);
    // B operand of layer 1, column block cb (voxels 32 cb ..), k step s (features 16 s ..)
    bf16x8 B1[2][2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const bf16x8 R = mlp_swap_halves(lo ? G[2 * s + 1] : G[2 * s]);
        B1[0][s] = lo ? G[2 * s] : R;
        B1[1][s] = lo ? R : G[2 * s + 1];
    }
    bf16x8 H1[2][2][2];   // [row block][column block][half]
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) {
            f32x16 acc = {};
#pragma unroll
            for (int s = 0; s < 2; ++s) acc = mlp_mfma(M->a1[rb][s][lane], B1[cb][s], acc);
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[r] += M->b1[rb][g][r];
            if (KEEP) keep.z1[rb][cb] = acc;
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int e = 0; e < 8; ++e) H1[rb][cb][h][e] = (__bf16)fmaxf(acc[8 * h + e], 0.0f);
        }
    bf16x8 H2[2][2][2];
#pragma unroll
    for (int rb2 = 0; rb2 < 2; ++rb2)
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) {
            f32x16 acc = {};
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                for (int h = 0; h < 2; ++h) acc = mlp_mfma(M->a2[rb2][2 * rb + h][lane], H1[rb][cb][h], acc);
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[r] += M->b2[rb2][g][r];
            if (KEEP) keep.z2[rb2][cb] = acc;
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int e = 0; e < 8; ++e) H2[rb2][cb][h][e] = (__bf16)fmaxf(acc[8 * h + e], 0.0f);
        }
    f32x16 o[2];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
        f32x16 acc = {};
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int h = 0; h < 2; ++h) acc = mlp_mfma(M->a3[2 * rb + h][lane], H2[rb][cb][h], acc);
        o[cb] = acc;
    }
    // output rows 0..3 sit in the low lane half, rows 4..5 in the high half, of the voxel's
    // column block: each lane sends its partner the part the partner's voxel needs
    float snd[4] = {lo ? o[1][0] : o[0][0], lo ? o[1][1] : o[0][1], lo ? o[1][2] : 0.0f, lo ? o[1][3] : 0.0f};
    float rcv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) rcv[k] = __shfl_xor(snd[k], 32, 64);
    if (lo) {
        y[0] = o[0][0]; y[1] = o[0][1]; y[2] = o[0][2]; y[3] = o[0][3];
        y[4] = rcv[0]; y[5] = rcv[1];
    } else {
        y[0] = rcv[0]; y[1] = rcv[1]; y[2] = rcv[2]; y[3] = rcv[3];
        y[4] = o[1][0]; y[5] = o[1][1];
    }
#pragma unroll
    for (int k = 0; k < MLP_OUT; ++k) y[k] += M->b3[k];
}

// Features of one voxel (raw, before normalisation):
//   [0, 4)   unit quaternion (w, x, y, z), w >= 0, of the orientation matrix a (crystal -> sample)
//   [4, 10)  direction of the previous Newton call's stress change d (crystal frame, B basis:
//            5 deviatoric components and the hydrostatic one)
//   [10, 16) the Newton call's starting stress lam in the crystal frame, in units of 100 MPa
//   [16, 22) the forcing of the coming call: e = MLP_E_SCALE * (eps_c - eps_c of the previous call)
//            (crystal frame, B basis; eps_c = sym(DisGrad)) as e / (|d| + |e|), in [-1, 1]
//   [22]     log10(|d| / |lam|)
//   [23]     log10 of the mean dislocation density rho_s over the slip systems
// nd = |d| is the scale the outputs are given in.  The training target of a sample is the change
// t of the coming Newton call as the correction to the extrapolated one: t / nd - u.
constexpr double MLP_E_SCALE = 1.0e5;   // MPa: an elastic modulus scale turning strain into stress
__device__ __forceinline__ void mlp_features(const double a[9], const double lam[6], const float d[6],
                                             const double dec[6], double rho_mean, float x[MLP_IN], float &nd) {
    const double a00 = a[0], a11 = a[4], a22 = a[8];
    const double qw = 0.5 * sqrt(fmax(0.0, 1.0 + a00 + a11 + a22));
    const double qx = copysign(0.5 * sqrt(fmax(0.0, 1.0 + a00 - a11 - a22)), a[7] - a[5]);
    const double qy = copysign(0.5 * sqrt(fmax(0.0, 1.0 - a00 + a11 - a22)), a[2] - a[6]);
    const double qz = copysign(0.5 * sqrt(fmax(0.0, 1.0 - a00 - a11 + a22)), a[3] - a[1]);
    x[0] = (float)qw; x[1] = (float)qx; x[2] = (float)qy; x[3] = (float)qz;
    double n2 = 0.0, l2 = 0.0, e2 = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        n2 += (double)d[k] * (double)d[k];
        l2 += lam[k] * lam[k];
        e2 += dec[k] * dec[k];
    }
    const double n = sqrt(n2);
    const double rn = n > 0.0 ? 1.0 / n : 0.0;
    const double en = MLP_E_SCALE * sqrt(e2);
    const double re = n + en > 0.0 ? MLP_E_SCALE / (n + en) : 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        x[4 + k] = (float)((double)d[k] * rn);
        x[10 + k] = (float)(lam[k] * 0.01);
        x[16 + k] = (float)(dec[k] * re);
    }
    x[22] = (float)log10(fmax(n, 1e-300) / fmax(sqrt(l2), 1e-300));
    x[23] = (float)log10(fmax(rho_mean, 1e-30));
    nd = (float)n;
}

// Normalised features of a voxel (x - mean) * inverse spread
__device__ __forceinline__ void mlp_normalise(const MlpPacked *__restrict__ M, float x[MLP_IN]) {
#pragma unroll
    for (int k = 0; k < MLP_IN; ++k) x[k] = (x[k] - M->xm[k]) * M->xs[k];
}

// Stress change predicted for the Newton start: the previous change d = nd u extrapolated with a
// learned step length, alpha d, alpha = 1 + c . u the network's de-normalised correction
// c = y * spread + mean projected on u, limited to 0 <= alpha <= 1 (a non-finite alpha gives 1);
// zero where the previous change is zero.  The correction's component across d is not applied,
// and the step never goes beyond the extrapolated one: moving the start off the line through the
// last two iterates, or past the extrapolation, lengthened the Newton solves (a start past the
// solution on an active slip system costs about one iteration per unit of overshoot in tau / tau_0,
// a start short of it far less).  A zero correction is the extrapolated start.
__device__ __forceinline__ void mlp_start(const MlpPacked *__restrict__ M, const float y[MLP_OUT],
                                          const float u[MLP_OUT], float nd, double ws[6]) {
    double cu = 0.0;
#pragma unroll
    for (int k = 0; k < MLP_OUT; ++k) cu += ((double)y[k] * M->ts[k] + M->tm[k]) * (double)u[k];
    double alpha = 1.0 + cu;
    alpha = alpha == alpha ? fmin(fmax(alpha, 0.0), 1.0) : 1.0;
    const bool ok = nd > 0.0f && nd < 1e30f;
#pragma unroll
    for (int k = 0; k < MLP_OUT; ++k) ws[k] = ok ? (double)nd * alpha * (double)u[k] : 0.0;
}

}  // namespace srx
