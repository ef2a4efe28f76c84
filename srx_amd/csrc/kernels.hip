// kernels.hip -- CDNA4 (gfx950) kernels of the EVP-FFT increment loop.
//
// FFT passes (evolution.c:1157-1201 replaced): LDS-resident Stockham
// radix-2/3/4/5/8 transforms over batches of pencils, 256 threads per block, their global
// loads issued in chunks before the first wait.
//   fwd_z   real [c][x][y][z]      -> half spectrum B [c][kz][x][y]   (two real comps per complex FFT)
//   fft_y   B <-> C [c][kz][y][x]  (the y transform doubles as the pencil transpose)
//   green   C(6 stress comps) -> forward x -> Green operator -> inverse x -> C(9 comps), in place
//           (k_green_wave: one wavefront per row, cross-lane butterflies, for x extents 64/128/256;
//           k_green_x: LDS transforms otherwise)
//   inv_z   B(9 comps) -> c2r along z, fused with the FP64 DisGrad update (Eq. 15)
// Per-voxel kernels (one voxel per lane, FP64; fields read and written as raw buffer accesses,
// one 32-bit lane offset from a scalar base):
//   nr      the constitutive Newton solve of UpdateStress (evolution.c:157-427)
//   end     Edot/Eps update, update_orient, DislocationEvolution (evolution.c:1268-1362)
// Block partial sums are reduced in a fixed order (k_reduce, k_reduce_seg, k_control), so every
// sum is bitwise repeatable from run to run.
#include "srx_dev.hpp"
#include "mlp_dev.hpp"

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <cmath>
#include <cstdio>

namespace srx {

// Optional start/stop events for the next launch: when set, the launch goes
// through hipExtLaunchKernelGGL, which stamps the events on the kernel's own
// dispatch packet.
static thread_local hipEvent_t g_ev_start = nullptr, g_ev_stop = nullptr;
void set_launch_events(hipEvent_t a, hipEvent_t b) {
    g_ev_start = a;
    g_ev_stop = b;
}
#define SRX_LAUNCH(kernel, grid, block, shmem, stream, ...)                                              \
    do {                                                                                                 \
        if (g_ev_start) {                                                                                \
            hipExtLaunchKernelGGL(kernel, grid, block, shmem, stream, g_ev_start, g_ev_stop, 0, __VA_ARGS__); \
            g_ev_start = g_ev_stop = nullptr;                                                            \
        } else {                                                                                         \
            hipLaunchKernelGGL(kernel, grid, block, shmem, stream, __VA_ARGS__);                         \
        }                                                                                                \
    } while (0)

// ---------------------------------------------------------------------------
// complex helpers
// ---------------------------------------------------------------------------
template <typename T>
struct cx {
    T x, y;
};
template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
template <typename T>
__device__ __forceinline__ cx<T> cscale(cx<T> a, T s) { return {a.x * s, a.y * s}; }
// multiply by -i*sign ... helpers: i*a = (-a.y, a.x)
template <typename T>
__device__ __forceinline__ cx<T> mul_i(cx<T> a) { return {-a.y, a.x}; }
// two complex values (a pair of components loaded together)
template <typename T>
struct cx2 {
    cx<T> a, b;
};

// ---------------------------------------------------------------------------
// radix butterflies; sign = -1 forward (e^{-2 pi i/R}), +1 inverse
// ---------------------------------------------------------------------------
template <typename T, int R>
struct Dft;

template <typename T>
struct Dft<T, 2> {
    __device__ __forceinline__ static void run(cx<T> *v, int) {
        cx<T> a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
    }
};
template <typename T>
struct Dft<T, 3> {
    __device__ __forceinline__ static void run(cx<T> *v, int sign) {
        const T c = T(-0.5), s = T(0.86602540378443864676) * T(sign);
        cx<T> t1 = v[1] + v[2], t2 = v[1] - v[2];
        cx<T> a = v[0];
        cx<T> m = {a.x + c * t1.x, a.y + c * t1.y};
        cx<T> r = {-s * t2.y, s * t2.x};   // i*s*t2
        v[0] = a + t1;
        v[1] = m + r;
        v[2] = m - r;
    }
};
template <typename T>
struct Dft<T, 4> {
    __device__ __forceinline__ static void run(cx<T> *v, int sign) {
        cx<T> t0 = v[0] + v[2], t1 = v[0] - v[2], t2 = v[1] + v[3], d = v[1] - v[3];
        // t3 = sign*i*(b-d): forward -> -i(b-d)
        cx<T> t3 = sign < 0 ? cx<T>{d.y, -d.x} : cx<T>{-d.y, d.x};
        v[0] = t0 + t2;
        v[2] = t0 - t2;
        v[1] = t1 + t3;
        v[3] = t1 - t3;
    }
};
template <typename T>
struct Dft<T, 5> {
    __device__ __forceinline__ static void run(cx<T> *v, int sign) {
        const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
        const T s1 = T(0.95105651629515357212) * T(sign), s2 = T(0.58778525229247312917) * T(sign);
        cx<T> a = v[0];
        cx<T> p1 = v[1] + v[4], m1 = v[1] - v[4], p2 = v[2] + v[3], m2 = v[2] - v[3];
        cx<T> r1 = {a.x + c1 * p1.x + c2 * p2.x, a.y + c1 * p1.y + c2 * p2.y};
        cx<T> r2 = {a.x + c2 * p1.x + c1 * p2.x, a.y + c2 * p1.y + c1 * p2.y};
        // i*(s1*m1 + s2*m2) and i*(s2*m1 - s1*m2)
        cx<T> q1 = {s1 * m1.x + s2 * m2.x, s1 * m1.y + s2 * m2.y};
        cx<T> q2 = {s2 * m1.x - s1 * m2.x, s2 * m1.y - s1 * m2.y};
        v[0] = a + p1 + p2;
        v[1] = r1 + mul_i(q1);
        v[4] = r1 - mul_i(q1);
        v[2] = r2 + mul_i(q2);
        v[3] = r2 - mul_i(q2);
    }
};

template <typename T>
struct Dft<T, 8> {
    // even/odd split into two radix-4 transforms
    __device__ __forceinline__ static void run(cx<T> *v, int sign) {
        cx<T> e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
        Dft<T, 4>::run(e, sign);
        Dft<T, 4>::run(o, sign);
        const T h = T(0.70710678118654752440);
        const T sh = T(sign) * h;
        const cx<T> o1 = cmul(o[1], cx<T>{h, sh});
        const cx<T> o2 = sign < 0 ? cx<T>{o[2].y, -o[2].x} : cx<T>{-o[2].y, o[2].x};
        const cx<T> o3 = cmul(o[3], cx<T>{-h, sh});
        v[0] = e[0] + o[0];
        v[4] = e[0] - o[0];
        v[1] = e[1] + o1;
        v[5] = e[1] - o1;
        v[2] = e[2] + o2;
        v[6] = e[2] - o2;
        v[3] = e[3] + o3;
        v[7] = e[3] - o3;
    }
};

constexpr int FFT_BLOCK = 256;

// n / d for 0 <= n with n * d < 2^32, as one high multiply by ceil(2^32 / d) (exact in that
// range); replaces the long integer-division sequence in the index arithmetic of the passes.
struct FastDiv {
    unsigned m;
    int d;
    __device__ __forceinline__ explicit FastDiv(int dd) : m(dd > 1 ? 0xFFFFFFFFu / (unsigned)dd + 1u : 0u), d(dd) {}
    __device__ __forceinline__ int div(int n) const { return d > 1 ? (int)__umulhi((unsigned)n, m) : n; }
};

// Element loop of a pass's global-memory side: elements [0, tot) dealt to the block's threads,
// taken in chunks of CH per thread whose loads ld(i) are all issued before the first store
// st(i, value) waits on them (one round trip to HBM per chunk, not per element).  Threads past the
// end load element tot - 1 again and store nothing.
constexpr int FFT_CHUNK = 8;
template <typename V, typename Ld, typename St>
__device__ __forceinline__ void chunked_copy(int tot, Ld ld, St st) {
    for (int b = 0; b < tot; b += FFT_CHUNK * FFT_BLOCK) {
        V v[FFT_CHUNK];
#pragma unroll
        for (int e = 0; e < FFT_CHUNK; ++e) v[e] = ld(min(b + e * FFT_BLOCK + (int)threadIdx.x, tot - 1));
#pragma unroll
        for (int e = 0; e < FFT_CHUNK; ++e) {
            const int i = b + e * FFT_BLOCK + (int)threadIdx.x;
            if (i < tot) st(i, v[e]);
        }
    }
}

// One Stockham stage (radix R, sub-transform span L) over `npen` pencils of
// length N stored pencil-major in LDS.  Every butterfly of the stage is held
// in registers between the load and the store, so the stage runs in place.
template <typename T, int R, int MAXB>
__device__ __forceinline__ void stockham_stage(cx<T> *buf, int npen, int N, int L, const cx<T> *__restrict__ tw,
                                               int sign, int pitch) {
    const int nb = N / R;
    const int total = npen * nb;
    const int tstep = N / (L * R);
    const FastDiv dnb(nb), dL(L);
    cx<T> v[MAXB][R];
#pragma unroll
    for (int m = 0; m < MAXB; ++m) {
        const int t = threadIdx.x + m * FFT_BLOCK;
        if (t < total) {
            const int p = dnb.div(t), j = t - p * nb;
            const cx<T> *src = buf + p * pitch + j;
#pragma unroll
            for (int r = 0; r < R; ++r) v[m][r] = src[r * nb];
        }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < MAXB; ++m) {
        const int t = threadIdx.x + m * FFT_BLOCK;
        if (t < total) {
            const int p = dnb.div(t), j = t - p * nb;
            const int jl = dL.div(j), k = j - jl * L;
            if (k) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    cx<T> w = tw[r * k * tstep];
                    if (sign > 0) w.y = -w.y;
                    v[m][r] = cmul(v[m][r], w);
                }
            }
            Dft<T, R>::run(v[m], sign);
            cx<T> *dst = buf + p * pitch + jl * L * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) dst[r * L] = v[m][r];
        }
    }
    __syncthreads();
}

// The same stage with the pencil length and span known at compile time (powers of two):
// the index arithmetic reduces to shifts and masks.
template <typename T, int R, int MAXB, int N, int L>
__device__ __forceinline__ void stockham_stage_ct(cx<T> *buf, int npen, const cx<T> *__restrict__ tw, int sign,
                                                  int pitch) {
    constexpr int nb = N / R;
    constexpr int tstep = N / (L * R);
    const int total = npen * nb;
    cx<T> v[MAXB][R];
#pragma unroll
    for (int m = 0; m < MAXB; ++m) {
        const int t = threadIdx.x + m * FFT_BLOCK;
        if (t < total) {
            const int p = t / nb, j = t % nb;
            const cx<T> *src = buf + p * pitch + j;
#pragma unroll
            for (int r = 0; r < R; ++r) v[m][r] = src[r * nb];
        }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < MAXB; ++m) {
        const int t = threadIdx.x + m * FFT_BLOCK;
        if (t < total) {
            const int p = t / nb, j = t % nb;
            const int k = j % L;
            if (k) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    cx<T> w = tw[r * k * tstep];
                    if (sign > 0) w.y = -w.y;
                    v[m][r] = cmul(v[m][r], w);
                }
            }
            Dft<T, R>::run(v[m], sign);
            cx<T> *dst = buf + p * pitch + (j / L) * L * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) dst[r * L] = v[m][r];
        }
    }
    __syncthreads();
}

// radix of stage st of a power-of-two length N, in the order make_plan chooses
// (radix 8 as often as possible without a radix-2 stage, then 4s, then a 2)
__host__ __device__ constexpr int pow2_radix(int N, int st) {
    int k2 = 0;
    while ((1 << k2) < N) ++k2;
    int n8 = k2 / 3, n4 = 0, n2 = 0;
    if (k2 % 3 == 1) { if (n8 > 0) { --n8; n4 = 2; } else { n2 = 1; } }
    else if (k2 % 3 == 2) n4 = 1;
    if (st < n8) return 8;
    if (st < n8 + n4) return 4;
    if (st < n8 + n4 + n2) return 2;
    return 0;
}

template <typename T, int MAXB, int N, int ST, int L>
__device__ __forceinline__ void lds_fft_pow2(cx<T> *buf, int npen, const cx<T> *__restrict__ tw, int sign,
                                             int pitch) {
    constexpr int R = pow2_radix(N, ST);
    if constexpr (R != 0) {
        stockham_stage_ct<T, R, MAXB, N, L>(buf, npen, tw, sign, pitch);
        lds_fft_pow2<T, MAXB, N, ST + 1, L * R>(buf, npen, tw, sign, pitch);
    }
}

// Full 1-D transform of npen LDS-resident pencils.  Caller guarantees
// npen*N/min_radix <= MAXB*FFT_BLOCK and a __syncthreads() before the call.
// POW2: lengths 32..256 use the compile-time stages (measured faster in the x and y passes;
// in the z passes, which interleave the transform with the real/complex packing, the extra
// registers cost an occupancy step and the runtime stages are faster).
template <typename T, int MAXB, bool POW2 = true>
__device__ void lds_fft(cx<T> *buf, int npen, const FftPlan &pl, const cx<T> *__restrict__ tw, int sign,
                        int pitch) {
    if constexpr (POW2) {
        switch (pl.n) {
        case 32: lds_fft_pow2<T, MAXB, 32, 0, 1>(buf, npen, tw, sign, pitch); return;
        case 64: lds_fft_pow2<T, MAXB, 64, 0, 1>(buf, npen, tw, sign, pitch); return;
        case 128: lds_fft_pow2<T, MAXB, 128, 0, 1>(buf, npen, tw, sign, pitch); return;
        case 256: lds_fft_pow2<T, MAXB, 256, 0, 1>(buf, npen, tw, sign, pitch); return;
        default: break;
        }
    }
    int L = 1;
    for (int st = 0; st < pl.nst; ++st) {
        const int R = pl.rad[st];
        switch (R) {
        case 8: stockham_stage<T, 8, MAXB>(buf, npen, pl.n, L, tw, sign, pitch); break;
        case 4: stockham_stage<T, 4, MAXB>(buf, npen, pl.n, L, tw, sign, pitch); break;
        case 2: stockham_stage<T, 2, MAXB>(buf, npen, pl.n, L, tw, sign, pitch); break;
        case 3: stockham_stage<T, 3, MAXB>(buf, npen, pl.n, L, tw, sign, pitch); break;
        default: stockham_stage<T, 5, MAXB>(buf, npen, pl.n, L, tw, sign, pitch); break;
        }
        L *= R;
    }
}

// Waves per SIMD targeted for the z passes (0: the compiler's choice).
#ifndef SRX_ZPASS_WPE
#define SRX_ZPASS_WPE 0
#endif
#if SRX_ZPASS_WPE > 0
#define SRX_ZPASS_ATTR __attribute__((amdgpu_waves_per_eu(SRX_ZPASS_WPE, 8)))
#else
#define SRX_ZPASS_ATTR
#endif

// ---------------------------------------------------------------------------
// forward z: real comps -> half spectrum, layout B [c][kz][x][y]
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(FFT_BLOCK) SRX_ZPASS_ATTR k_fwd_z(const double *__restrict__ in, cx<T> *__restrict__ specB,
                                                     int ncomp, int n0, int n1, int n2, FftPlan pz,
                                                     const cx<T> *__restrict__ tw, int TY, double m0, double m1,
                                                     double m2, double m3, double m4, double m5,
                                                     const IterCtl *__restrict__ ctl) {
    if (ctl && ctl->done) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
    const int P = n2 + 1;           // padded pencil pitch: rows fall on different LDS banks
    cx<T> *stw = buf + TY * P;      // twiddle table staged in LDS
    const int x = blockIdx.x;
    const int y0 = blockIdx.y * TY;
    const int ty = min(TY, n1 - y0);
    const int H = n2 / 2 + 1;
    const size_t N = (size_t)n0 * n1 * n2;
    const size_t HS = (size_t)H * n0 * n1;
    double mean[6] = {m0, m1, m2, m3, m4, m5};
    if (ctl)
        for (int k = 0; k < 6; ++k) mean[k] = ctl->mean6[k];
    const FastDiv dn2(n2), dty(ty);
    for (int i = threadIdx.x; i < n2; i += FFT_BLOCK) stw[i] = tw[i];
    // one pair of components per block along grid z
    for (int ca = 2 * blockIdx.z; ca < ncomp && ca < 2 * blockIdx.z + 2; ca += 2) {
        const int cb = ca + 1;
        const bool two = cb < ncomp;
        const double ma = ca < 6 ? mean[ca] : 0.0, mb = (two && cb < 6) ? mean[cb] : 0.0;
        const double *ia = in + ca * N, *ib = in + (two ? cb : ca) * N;
        const size_t g0 = ((size_t)x * n1 + y0) * n2;   // the block's rows are contiguous
        chunked_copy<cx<double>>(
            ty * n2, [&](int i) { return cx<double>{ia[g0 + i], ib[g0 + i]}; },
            [&](int i, cx<double> v) {
                const int yy = dn2.div(i), z = i - yy * n2;
                buf[yy * P + z] = cx<T>{T(v.x - ma), T(two ? v.y - mb : 0.0)};
            });
        __syncthreads();
        lds_fft<T, 2, false>(buf, ty, pz, stw, -1, P);
        for (int i = threadIdx.x; i < H * ty; i += FFT_BLOCK) {
            const int kz = dty.div(i), yy = i - kz * ty;
            const cx<T> z1 = buf[yy * P + kz];
            const cx<T> z2 = buf[yy * P + (kz == 0 ? 0 : n2 - kz)];
            const size_t o = ((size_t)kz * n0 + x) * n1 + y0 + yy;
            specB[ca * HS + o] = cx<T>{T(0.5) * (z1.x + z2.x), T(0.5) * (z1.y - z2.y)};
            if (two) specB[cb * HS + o] = cx<T>{T(0.5) * (z1.y + z2.y), T(-0.5) * (z1.x - z2.x)};
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// y transform with transpose.  b_to_c: B [c][kz][x][y] -> C [c][kz][y][x]
// (forward, sign -1); otherwise C -> B (inverse, sign +1).
// grid: (ceil(n0/TX), H, ncomp)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(FFT_BLOCK) k_fft_y(const cx<T> *__restrict__ in, cx<T> *__restrict__ out,
                                                     int n0, int n1, int n2, int b_to_c, FftPlan py,
                                                     const cx<T> *__restrict__ tw, int TX,
                                                     const IterCtl *__restrict__ ctl) {
    if (ctl && ctl->done) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
    const int P = n1 + 1;   // padded pencil pitch
    cx<T> *stw = buf + TX * P;
    for (int i = threadIdx.x; i < n1; i += FFT_BLOCK) stw[i] = tw[i];
    const int H = n2 / 2 + 1;
    const int x0 = blockIdx.x * TX;
    const int tx = min(TX, n0 - x0);
    const int kz = blockIdx.y;
    const int c = blockIdx.z;
    const size_t plane = (size_t)n0 * n1;
    const size_t base = ((size_t)c * H + kz) * plane;
    const FastDiv dn1(n1), dtx(tx);
    if (b_to_c) {
        // the block's x rows are contiguous in B
        const cx<T> *src = in + base + (size_t)x0 * n1;
        chunked_copy<cx<T>>(tx * n1, [&](int i) { return src[i]; },
                            [&](int i, cx<T> v) {
                                const int xx = dn1.div(i), y = i - xx * n1;
                                buf[xx * P + y] = v;
                            });
    } else {
        chunked_copy<cx<T>>(
            tx * n1,
            [&](int i) {
                const int y = dtx.div(i), xx = i - y * tx;
                return in[base + (size_t)y * n0 + x0 + xx];
            },
            [&](int i, cx<T> v) {
                const int y = dtx.div(i), xx = i - y * tx;
                buf[xx * P + y] = v;
            });
    }
    __syncthreads();
    lds_fft<T, 2>(buf, tx, py, stw, b_to_c ? -1 : +1, P);
    if (b_to_c) {
        for (int i = threadIdx.x; i < tx * n1; i += FFT_BLOCK) {
            const int y = dtx.div(i), xx = i - y * tx;
            out[base + (size_t)y * n0 + x0 + xx] = buf[xx * P + y];
        }
    } else {
        for (int i = threadIdx.x; i < tx * n1; i += FFT_BLOCK) {
            const int xx = dn1.div(i), y = i - xx * n1;
            out[base + (size_t)(x0 + xx) * n1 + y] = buf[xx * P + y];
        }
    }
}

// ---------------------------------------------------------------------------
// Green operator at one k point (init.c:542-592 with evolution.c:1169-1183):
//   Omega = (C0_ijkl n_j n_l)^-1,  du_ij = -Omega_ik (sig_kl n_l) n_j
// sig: 6 B-basis comps (complex); du: 9 comps (row-major tensor).
// ---------------------------------------------------------------------------
template <typename A>
__device__ __forceinline__ void unit_k(int idx, int n, A &t) {
    A v = (A)idx * (A)(1.0 / n);
    v -= round(v);
    t = v * (A)6.28318530718;
}

__device__ void green_w(const double *__restrict__ C0, const double nv[3], const double sr[9], const double si[9],
                        double wr[3], double wi[3]) {
    double A[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double acc = 0.0;
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int l = 0; l < 3; ++l) acc += C0[((i * 3 + j) * 3 + k) * 3 + l] * nv[j] * nv[l];
            A[i][k] = acc;
        }
    // M3Inverse (init.c:518-540), cofactors over the determinant
    const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
                       A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                       A[0][2] * (A[1][0] * A[2][1] - A[2][0] * A[1][1]);
    double om[3][3];
    om[0][0] = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    om[0][1] = A[0][2] * A[2][1] - A[0][1] * A[2][2];
    om[0][2] = A[0][1] * A[1][2] - A[0][2] * A[1][1];
    om[1][0] = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    om[1][1] = A[0][0] * A[2][2] - A[0][2] * A[2][0];
    om[1][2] = A[0][2] * A[1][0] - A[0][0] * A[1][2];
    om[2][0] = A[1][0] * A[2][1] - A[1][1] * A[2][0];
    om[2][1] = A[0][1] * A[2][0] - A[0][0] * A[2][1];
    om[2][2] = A[0][0] * A[1][1] - A[0][1] * A[1][0];
    const double rdet = 1.0 / det;
    double tr[3], ti[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        tr[k] = sr[3 * k] * nv[0] + sr[3 * k + 1] * nv[1] + sr[3 * k + 2] * nv[2];
        ti[k] = si[3 * k] * nv[0] + si[3 * k + 1] * nv[1] + si[3 * k + 2] * nv[2];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        wr[i] = (om[i][0] * tr[0] + om[i][1] * tr[1] + om[i][2] * tr[2]) * rdet;
        wi[i] = (om[i][0] * ti[0] + om[i][1] * ti[1] + om[i][2] * ti[2]) * rdet;
    }
}

// B-basis 6-vector -> symmetric tensor (chg_basis opt 1, kinematics.c:55-60)
#define SRX_R2 0.70710678118654744
#define SRX_R3 0.57735026918962584
#define SRX_R6 0.40824829046386304

template <typename V>
__device__ __forceinline__ void b6_to_t9(const V v[6], V t[9]) {
    t[0] = v[0] * V(-SRX_R2) + v[1] * V(-SRX_R6) + v[5] * V(SRX_R3);
    t[4] = v[0] * V(SRX_R2) + v[1] * V(-SRX_R6) + v[5] * V(SRX_R3);
    t[8] = v[1] * V(2. * SRX_R6) + v[5] * V(SRX_R3);
    t[5] = t[7] = v[2] * V(SRX_R2);
    t[2] = t[6] = v[3] * V(SRX_R2);
    t[1] = t[3] = v[4] * V(SRX_R2);
}

// B-basis 6-vector -> symmetric tensor entries (00, 11, 22, 12, 02, 01)
template <typename A>
__device__ __forceinline__ void b6_to_sym(const A v[6], A t[6]) {
    t[0] = v[0] * A(-SRX_R2) + v[1] * A(-SRX_R6) + v[5] * A(SRX_R3);
    t[1] = v[0] * A(SRX_R2) + v[1] * A(-SRX_R6) + v[5] * A(SRX_R3);
    t[2] = v[1] * A(2. * SRX_R6) + v[5] * A(SRX_R3);
    t[3] = v[2] * A(SRX_R2);
    t[4] = v[3] * A(SRX_R2);
    t[5] = v[4] * A(SRX_R2);
}
__device__ __forceinline__ void sym_to_t9(const double t[6], double m[9]) {
    m[0] = t[0]; m[4] = t[1]; m[8] = t[2];
    m[5] = m[7] = t[3];
    m[2] = m[6] = t[4];
    m[1] = m[3] = t[5];
}

// Green operator at one k point from the symmetric contraction table G6:
//   A_ik = sum_jl C0_ijkl n_j n_l = sum_b G6[(ik)][b] m_b,  m = (n0^2, n1^2, n2^2, n1n2, n0n2, n0n1)
//   w = A^-1 (sig n)  (cofactors of the symmetric A, as M3Inverse, init.c:518-540)
// and du_ij = -w_i n_j  (init.c:585 with evolution.c:1169-1183).
template <typename A>
__device__ __forceinline__ void green_w6(const A *__restrict__ G6, const A n[3], const A sr[6], const A si[6],
                                         A wr[3], A wi[3]) {
    const A m[6] = {n[0] * n[0], n[1] * n[1], n[2] * n[2], n[1] * n[2], n[0] * n[2], n[0] * n[1]};
    A Am[6];
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        A acc = A(0);
#pragma unroll
        for (int b = 0; b < 6; ++b) acc += G6[6 * a + b] * m[b];
        Am[a] = acc;
    }
    // A = [[A0, A5, A4], [A5, A1, A3], [A4, A3, A2]]
    const A c00 = Am[1] * Am[2] - Am[3] * Am[3], c11 = Am[0] * Am[2] - Am[4] * Am[4], c22 = Am[0] * Am[1] - Am[5] * Am[5];
    const A c01 = Am[4] * Am[3] - Am[5] * Am[2], c02 = Am[5] * Am[3] - Am[4] * Am[1], c12 = Am[5] * Am[4] - Am[0] * Am[3];
    const A rdet = A(1) / (Am[0] * c00 + Am[5] * c01 + Am[4] * c02);
    // t = sig n with sig = [[s0, s5, s4], [s5, s1, s3], [s4, s3, s2]]
    const A t0r = sr[0] * n[0] + sr[5] * n[1] + sr[4] * n[2], t0i = si[0] * n[0] + si[5] * n[1] + si[4] * n[2];
    const A t1r = sr[5] * n[0] + sr[1] * n[1] + sr[3] * n[2], t1i = si[5] * n[0] + si[1] * n[1] + si[3] * n[2];
    const A t2r = sr[4] * n[0] + sr[3] * n[1] + sr[2] * n[2], t2i = si[4] * n[0] + si[3] * n[1] + si[2] * n[2];
    wr[0] = (c00 * t0r + c01 * t1r + c02 * t2r) * rdet;
    wi[0] = (c00 * t0i + c01 * t1i + c02 * t2i) * rdet;
    wr[1] = (c01 * t0r + c11 * t1r + c12 * t2r) * rdet;
    wi[1] = (c01 * t0i + c11 * t1i + c12 * t2i) * rdet;
    wr[2] = (c02 * t0r + c12 * t1r + c22 * t2r) * rdet;
    wi[2] = (c02 * t0i + c12 * t1i + c22 * t2i) * rdet;
}

// Green operator at one k point (kx, ky, kz): B-basis spectrum q[6] of the stress -> du[9]
// (row-major), computed in the spectrum's precision T (FP32 in the mixed-precision path, whose
// convolution is FP32 throughout).  k = 0: -S0 (sig + dc6) for ElastBC 0, else 0 (init.c:558-569);
// dc6 is the grid sum of sig that a mean-subtracted (mixed precision) spectrum lacks, B basis.
template <typename T>
__device__ __forceinline__ void green_point(const DevConst *__restrict__ cst, const double dc6[6], int kx, int ky,
                                            int kz, int n0, int n1, int n2, const cx<T> q[6], cx<T> out[9]) {
    using A = T;
    const A *G6;
    if constexpr (sizeof(T) == 4) G6 = cst->G6f;
    else G6 = cst->G6;
    if (kx == 0 && ky == 0 && kz == 0) {
        double sr[6], si[6], vr[6], vi[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            vr[c] = (double)q[c].x;
            vi[c] = (double)q[c].y;
        }
        b6_to_sym(vr, sr);
        b6_to_sym(vi, si);
        double dr[9], di[9];
        if (cst->elast_bc == 0) {
            double dcs[6], t9r[9], t9i[9];
            b6_to_sym(dc6, dcs);
            sym_to_t9(sr, t9r);
            sym_to_t9(si, t9i);
            double d9[9];
            sym_to_t9(dcs, d9);
            for (int qq = 0; qq < 9; ++qq) {
                double accr = 0.0, acci = 0.0;
                for (int r = 0; r < 9; ++r) {
                    accr += -cst->S0[qq * 9 + r] * (t9r[r] + d9[r]);
                    acci += -cst->S0[qq * 9 + r] * t9i[r];
                }
                dr[qq] = accr;
                di[qq] = acci;
            }
        } else {
#pragma unroll
            for (int qq = 0; qq < 9; ++qq) dr[qq] = di[qq] = 0.0;
        }
#pragma unroll
        for (int qq = 0; qq < 9; ++qq) out[qq] = cx<T>{T(dr[qq]), T(di[qq])};
        return;
    }
    A tzv, tyv, txv;
    unit_k(kz, n2, tzv);
    unit_k(ky, n1, tyv);
    unit_k(kx, n0, txv);
    const bool nyz = (n2 % 2 == 0) && (kz == n2 / 2);
    const bool nyx = (n0 % 2 == 0) && (kx == n0 / 2);
    const bool nyy = (n1 % 2 == 0) && (ky == n1 / 2);
    A sr[6], si[6];
    {
        A vr[6], vi[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            vr[c] = (A)q[c].x;
            vi[c] = (A)q[c].y;
        }
        b6_to_sym(vr, sr);
        b6_to_sym(vi, si);
    }
    A nv[3] = {txv, tyv, tzv};
    const A rg = A(1) / sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
    nv[0] *= rg; nv[1] *= rg; nv[2] *= rg;
    A wr[3], wi[3];
    green_w6(G6, nv, sr, si, wr, wi);
    if (!(nyx || nyy || nyz)) {
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) out[3 * a + b] = cx<T>{T(-wr[a] * nv[b]), T(-wi[a] * nv[b])};
    } else {
        // Taking Re() of the inverse transform (evolution.c:1197) keeps the Hermitian part of du; at
        // a Nyquist index that part averages Gamma(n) with Gamma(n') where n' has the Nyquist
        // components negated.
        const A np[3] = {nyx ? -nv[0] : nv[0], nyy ? -nv[1] : nv[1], nyz ? -nv[2] : nv[2]};
        A wr2[3], wi2[3];
        green_w6(G6, np, sr, si, wr2, wi2);
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b)
                out[3 * a + b] = cx<T>{T(A(0.5) * (-wr[a] * nv[b] - wr2[a] * np[b])),
                                       T(A(0.5) * (-wi[a] * nv[b] - wi2[a] * np[b]))};
    }
}

// x pass with the Green operator.  grid: (ceil(n1/TY), H).  LDS: 9 slots of TY*n0.
// mode 1: forward x of 6 comps, Green, inverse x of 9 comps (the solver pass)
// mode 0: forward x only (ncomp comps, testing)   mode 2: inverse x only (testing)
template <typename T, int MAXB>
__global__ void __launch_bounds__(FFT_BLOCK) k_green_x(cx<T> *__restrict__ spec, int n0, int n1, int n2, FftPlan px,
                                                       const cx<T> *__restrict__ tw, int TY, int mode, int ncomp,
                                                       const DevConst *__restrict__ cst, double dc0, double dc1,
                                                       double dc2, double dc3, double dc4, double dc5,
                                                       const IterCtl *__restrict__ ctl) {
    if (ctl && ctl->done) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
    const int H = n2 / 2 + 1;
    const int y0 = blockIdx.x * TY;
    const int ty = min(TY, n1 - y0);
    const int kz = blockIdx.y;
    const size_t plane = (size_t)n0 * n1;
    const int slot = ty * n0;
    const int cin = mode == 1 ? 6 : ncomp;
    const int cout = mode == 1 ? 9 : ncomp;
    const FastDiv dslot(slot), dn0(n0);
    cx<T> *stw = buf + (mode == 1 ? 9 : ncomp) * slot;
    for (int i = threadIdx.x; i < n0; i += FFT_BLOCK) stw[i] = tw[i];
    // the block's rows of one component are contiguous: y0 * n0 .. (y0 + ty) * n0 of plane kz
    chunked_copy<cx<T>>(
        cin * slot,
        [&](int i) {
            const int c = dslot.div(i), r = i - c * slot;
            return spec[((size_t)c * H + kz) * plane + (size_t)y0 * n0 + r];
        },
        [&](int i, cx<T> v) { buf[i] = v; });
    __syncthreads();
    if (mode != 2) lds_fft<T, MAXB>(buf, cin * ty, px, stw, -1, n0);
    if (mode == 1) {
        double dc6[6] = {dc0, dc1, dc2, dc3, dc4, dc5};
        if (ctl)
            for (int k = 0; k < 6; ++k) dc6[k] = ctl->dcs6[k];
        for (int i = threadIdx.x; i < slot; i += FFT_BLOCK) {
            const int yy = dn0.div(i), kx = i - yy * n0, ky = y0 + yy;
            cx<T> q[6], o[9];
#pragma unroll
            for (int c = 0; c < 6; ++c) q[c] = buf[c * slot + i];
            green_point(cst, dc6, kx, ky, kz, n0, n1, n2, q, o);
#pragma unroll
            for (int c = 0; c < 9; ++c) buf[c * slot + i] = o[c];
        }
        __syncthreads();
    }
    if (mode != 0) lds_fft<T, MAXB>(buf, cout * ty, px, stw, +1, n0);
    for (int i = threadIdx.x; i < cout * slot; i += FFT_BLOCK) {
        const int c = dslot.div(i), r = i - c * slot, yy = dn0.div(r), x = r - yy * n0;
        spec[((size_t)c * H + kz) * plane + (size_t)(y0 + yy) * n0 + x] = buf[c * slot + yy * n0 + x];
    }
}

// ---------------------------------------------------------------------------
// The solver's x pass with one wavefront per (ky, kz) row and no LDS or barriers, for
// n0 = 64 NPL.  Lane l holds points l + 64 r (r < NPL) of each component.  Forward transform as
// four steps: an in-lane NPL-point DFT, the twiddle w_N^(l k1), then a 64-point radix-2
// decimation-in-frequency transform across lanes (partners l ^ h, h = 32 .. 1), which leaves
// X[k1 + NPL bitrev6(l)] in lane l, register k1.  The Green operator is applied in that order and
// the inverse runs the conjugate-transposed stages in reverse (decimation in time), which returns
// natural order: an unnormalised inverse DFT, as lds_fft with sign +1.
// grid: (ceil(n1 / 4), H), 256 threads (four rows per block).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int brev6(int l) { return (int)(__brev((unsigned)l) >> 26); }

// value of lane ^ H: DPP quad permutations for H = 1, 2 (no LDS traffic), ds_swizzle's xor mode
// within 32 lanes for H = 4, 8, 16, ds_bpermute for H = 32
template <int H>
__device__ __forceinline__ int xor_lane(int v) {
    if constexpr (H == 1) return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false);   // quad_perm 1,0,3,2
    else if constexpr (H == 2) return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false);   // quad_perm 2,3,0,1
    else if constexpr (H < 32) return __builtin_amdgcn_ds_swizzle(v, (H << 10) | 0x1F);
    else return __shfl_xor(v, 32, 64);
}
template <int H>
__device__ __forceinline__ float xor_lane(float v) { return __int_as_float(xor_lane<H>(__float_as_int(v))); }
template <int H>
__device__ __forceinline__ double xor_lane(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = xor_lane<H>((int)(b & 0xffffffffll)), hi = xor_lane<H>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <int H, typename T>
__device__ __forceinline__ cx<T> shfl_xor_cx(cx<T> v) {
    return cx<T>{xor_lane<H>(v.x), xor_lane<H>(v.y)};
}
template <typename T>
__device__ __forceinline__ cx<T> conj_cx(cx<T> a) { return cx<T>{a.x, -a.y}; }

// Cross-lane stage ST (partner lane ^ h, h = 32 >> ST) and the ones after it, for C components at
// once: all partner values of a stage are requested before any is used, so one wait covers them.
// Branch free: with s = +1 in the lower lane of a pair and -1 in the upper one, and the upper
// lane's twiddle w (the lower lane's is 1),
//   decimation in frequency: lower x_lo + x_hi, upper (x_lo - x_hi) w  =  (u + s v) w'
//   decimation in time:      lower x_lo + w x_hi, upper x_lo - w x_hi  =  m_partner + s m,
//                            m = w' v (each upper lane multiplies its own value before the exchange)
// (u the partner's value, v the lane's own, w' = w in the upper lane and 1 in the lower one; a
// product by exactly 1 returns its operand, so the results equal the two-sided forms').  The last
// DIF and first DIT stage (h = 1) has w = 1 in every lane and no product.
template <typename T>
__device__ __forceinline__ cx<T> pair_twiddle(const cx<T> w, bool hi) {
    return cx<T>{hi ? w.x : T(1), hi ? w.y : T(0)};
}
template <typename T, int NPL, int C, int ST>
__device__ __forceinline__ void wave_dif_stage(cx<T> (&v)[C][NPL], const cx<T> ws[6], int lane) {
    if constexpr (ST < 6) {
        constexpr int h = 32 >> ST;
        const bool hi = lane & h;
        const T s = hi ? T(-1) : T(1);
        const cx<T> w = pair_twiddle(ws[ST], hi);
        cx<T> u[C][NPL];
#pragma unroll
        for (int c = 0; c < C; ++c)
#pragma unroll
            for (int k = 0; k < NPL; ++k) u[c][k] = shfl_xor_cx<h>(v[c][k]);
#pragma unroll
        for (int c = 0; c < C; ++c)
#pragma unroll
            for (int k = 0; k < NPL; ++k) {
                const cx<T> t{fma(s, v[c][k].x, u[c][k].x), fma(s, v[c][k].y, u[c][k].y)};
                v[c][k] = h == 1 ? t : cmul(t, w);
            }
        wave_dif_stage<T, NPL, C, ST + 1>(v, ws, lane);
    }
}
template <typename T, int NPL, int C, int ST>
__device__ __forceinline__ void wave_dit_stage(cx<T> (&v)[C][NPL], const cx<T> ws[6], int lane) {
    if constexpr (ST >= 0) {
        constexpr int h = 32 >> ST;
        const bool hi = lane & h;
        const T s = hi ? T(-1) : T(1);
        if constexpr (h != 1) {
            const cx<T> w = pair_twiddle(conj_cx(ws[ST]), hi);
#pragma unroll
            for (int c = 0; c < C; ++c)
#pragma unroll
                for (int k = 0; k < NPL; ++k) v[c][k] = cmul(v[c][k], w);
        }
        cx<T> u[C][NPL];
#pragma unroll
        for (int c = 0; c < C; ++c)
#pragma unroll
            for (int k = 0; k < NPL; ++k) u[c][k] = shfl_xor_cx<h>(v[c][k]);
#pragma unroll
        for (int c = 0; c < C; ++c)
#pragma unroll
            for (int k = 0; k < NPL; ++k) v[c][k] = cx<T>{fma(s, v[c][k].x, u[c][k].x), fma(s, v[c][k].y, u[c][k].y)};
        wave_dit_stage<T, NPL, C, ST - 1>(v, ws, lane);
    }
}

template <typename T, int NPL, int C>
__device__ __forceinline__ void wave_fft_fwd(cx<T> (&v)[C][NPL], const cx<T> wl[], const cx<T> ws[6], int lane) {
    if constexpr (NPL > 1) {
#pragma unroll
        for (int c = 0; c < C; ++c) {
            Dft<T, NPL>::run(v[c], -1);
#pragma unroll
            for (int k = 1; k < NPL; ++k) v[c][k] = cmul(v[c][k], wl[k - 1]);
        }
    }
    wave_dif_stage<T, NPL, C, 0>(v, ws, lane);
}

template <typename T, int NPL, int C>
__device__ __forceinline__ void wave_fft_inv(cx<T> (&v)[C][NPL], const cx<T> wl[], const cx<T> ws[6], int lane) {
    wave_dit_stage<T, NPL, C, 5>(v, ws, lane);
    if constexpr (NPL > 1) {
#pragma unroll
        for (int c = 0; c < C; ++c) {
#pragma unroll
            for (int k = 1; k < NPL; ++k) v[c][k] = cmul(v[c][k], conj_cx(wl[k - 1]));
            Dft<T, NPL>::run(v[c], +1);
        }
    }
}

// blocks of the wavefront Green kernel per compute unit the register allocator targets
#ifndef SRX_GREEN_MIN_BLOCKS
#define SRX_GREEN_MIN_BLOCKS 1
#endif
template <typename T, int NPL>
__global__ void __launch_bounds__(256, SRX_GREEN_MIN_BLOCKS) k_green_wave(cx<T> *__restrict__ spec, int n1, int n2,
                                                    const cx<T> *__restrict__ tw, const DevConst *__restrict__ cst,
                                                    double dc0, double dc1, double dc2, double dc3, double dc4,
                                                    double dc5, const IterCtl *__restrict__ ctl) {
    if (ctl && ctl->done) return;
    constexpr int N = 64 * NPL;
    const int lane = threadIdx.x & 63;
    const int ky = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int kz = blockIdx.y;
    if (ky >= n1) return;   // the whole wavefront: no barrier follows
    const int H = n2 / 2 + 1;
    const size_t plane = (size_t)N * n1;
    double dc6[6] = {dc0, dc1, dc2, dc3, dc4, dc5};
    if (ctl)
        for (int k = 0; k < 6; ++k) dc6[k] = ctl->dcs6[k];
    // twiddles of this lane: w_N^(l k1) for the four-step split, and W_2h^(l mod h) = w_N^(..) of
    // each cross-lane stage
    cx<T> wl[NPL > 1 ? NPL - 1 : 1];
#pragma unroll
    for (int k = 1; k < NPL; ++k) wl[k - 1] = tw[lane * k];
    cx<T> ws[6];
#pragma unroll
    for (int st = 0; st < 6; ++st) {
        const int h = 32 >> st;
        ws[st] = tw[(lane & (h - 1)) * (32 / h) * NPL];
    }
    const size_t row = (size_t)kz * plane + (size_t)ky * N + lane;
    const size_t cstride = (size_t)H * plane;
    cx<T> q[6][NPL];
#pragma unroll
    for (int c = 0; c < 6; ++c)
#pragma unroll
        for (int r = 0; r < NPL; ++r) q[c][r] = spec[c * cstride + row + 64 * r];
    wave_fft_fwd<T, NPL, 6>(q, wl, ws, lane);
    cx<T> o[9][NPL];
    const int kb = NPL * brev6(lane);
#pragma unroll
    for (int r = 0; r < NPL; ++r) {
        cx<T> qi[6], oi[9];
#pragma unroll
        for (int c = 0; c < 6; ++c) qi[c] = q[c][r];
        green_point(cst, dc6, r + kb, ky, kz, N, n1, n2, qi, oi);
#pragma unroll
        for (int c = 0; c < 9; ++c) o[c][r] = oi[c];
    }
    wave_fft_inv<T, NPL, 9>(o, wl, ws, lane);
#pragma unroll
    for (int c = 0; c < 9; ++c)
#pragma unroll
        for (int r = 0; r < NPL; ++r) spec[c * cstride + row + 64 * r] = o[c][r];
}

// ---------------------------------------------------------------------------
// inverse z (c2r) fused with Eq. 15: dg[c] += ddg[c] + Re(IFFT)/Nxyz
// grid: (n0, ceil(n1/TY)); `dg` may be an output (scale mode) when accumulate=0
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(FFT_BLOCK) SRX_ZPASS_ATTR k_inv_z(const cx<T> *__restrict__ specB, double *__restrict__ dg,
                                                     int ncomp, int n0, int n1, int n2, FftPlan pz,
                                                     const cx<T> *__restrict__ tw, int TY, int accumulate,
                                                     double divisor, double a0, double a1, double a2, double a3,
                                                     double a4, double a5, double a6, double a7, double a8,
                                                     const IterCtl *__restrict__ ctl) {
    if (ctl && ctl->done) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
    const int P = n2 + 1;   // padded pencil pitch
    cx<T> *stw = buf + TY * P;
    for (int i = threadIdx.x; i < n2; i += FFT_BLOCK) stw[i] = tw[i];
    const int x = blockIdx.x;
    const int y0 = blockIdx.y * TY;
    const int ty = min(TY, n1 - y0);
    const int H = n2 / 2 + 1;
    const size_t N = (size_t)n0 * n1 * n2;
    const size_t HS = (size_t)H * n0 * n1;
    double add[9] = {a0, a1, a2, a3, a4, a5, a6, a7, a8};
    if (ctl)
        for (int k = 0; k < 9; ++k) add[k] = ctl->ddg[k];
    const FastDiv dn2(n2), dty(ty);
    // 1/Nxyz: a power of two for the usual grids, where the product equals the division
    const double rdiv = 1.0 / divisor;
    // one pair of components per block along grid z
    for (int ca = 2 * blockIdx.z; ca < ncomp && ca < 2 * blockIdx.z + 2; ca += 2) {
        const int cb = ca + 1;
        const bool two = cb < ncomp;
        const cx<T> *sa = specB + ca * HS, *sb = specB + (two ? cb : ca) * HS;
        chunked_copy<cx2<T>>(
            H * ty,
            [&](int i) {
                const int kz = dty.div(i), yy = i - kz * ty;
                const size_t o = ((size_t)kz * n0 + x) * n1 + y0 + yy;
                return cx2<T>{sa[o], sb[o]};
            },
            [&](int i, cx2<T> v) {
                const int kz = dty.div(i), yy = i - kz * ty;
                const cx<T> A = v.a;
                const cx<T> B = two ? v.b : cx<T>{T(0), T(0)};
                buf[yy * P + kz] = cx<T>{A.x - B.y, A.y + B.x};
                if (kz >= 1 && kz <= n2 - H) buf[yy * P + (n2 - kz)] = cx<T>{A.x + B.y, B.x - A.y};
            });
        // DisGrad of the block's rows (contiguous): the first chunk of each thread's values is
        // loaded before the transform, so its round trip to HBM overlaps the LDS work
        double *da = dg + ca * N, *db = dg + (two ? cb : ca) * N;
        const size_t g0 = ((size_t)x * n1 + y0) * n2;
        const int tot = ty * n2;
        cx<double> pre[FFT_CHUNK];
        if (accumulate) {
#pragma unroll
            for (int e = 0; e < FFT_CHUNK; ++e) {
                const int i = min(e * FFT_BLOCK + (int)threadIdx.x, tot - 1);
                pre[e] = cx<double>{da[g0 + i], db[g0 + i]};
            }
        }
        __syncthreads();
        lds_fft<T, 2, false>(buf, ty, pz, stw, +1, P);
        for (int b = 0; b < tot; b += FFT_CHUNK * FFT_BLOCK) {
            if (accumulate && b > 0) {
#pragma unroll
                for (int e = 0; e < FFT_CHUNK; ++e) {
                    const int i = min(b + e * FFT_BLOCK + (int)threadIdx.x, tot - 1);
                    pre[e] = cx<double>{da[g0 + i], db[g0 + i]};
                }
            }
#pragma unroll
            for (int e = 0; e < FFT_CHUNK; ++e) {
                const int i = b + e * FFT_BLOCK + (int)threadIdx.x;
                if (i < tot) {
                    const int yy = dn2.div(i), z = i - yy * n2;
                    const cx<T> v = buf[yy * P + z];
                    if (accumulate) {
                        da[g0 + i] = pre[e].x + (add[ca] + (double)v.x * rdiv);
                        if (two) db[g0 + i] = pre[e].y + (add[cb] + (double)v.y * rdiv);
                    } else {
                        da[g0 + i] = (double)v.x * rdiv;
                        if (two) db[g0 + i] = (double)v.y * rdiv;
                    }
                }
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// rotations in the B basis
// ---------------------------------------------------------------------------
// Q[al][be] = B_al : (a B_be a^T) for the 5 deviatoric basis tensors; maps
// crystal-frame B-vectors to sample-frame ones when a = TranMat_xt2sa.
__device__ __forceinline__ void sym_to_b5(double m00, double m11, double m22, double m12, double m02, double m01,
                                          double *q) {
    q[0] = SRX_R2 * (m11 - m00);
    q[1] = SRX_R6 * (2.0 * m22 - m00 - m11);
    q[2] = 2.0 * SRX_R2 * m12;
    q[3] = 2.0 * SRX_R2 * m02;
    q[4] = 2.0 * SRX_R2 * m01;
}

__device__ __forceinline__ void q5_from(const double a[9], double Q[5][5]) {
    // columns of a
    const double c0[3] = {a[0], a[3], a[6]}, c1[3] = {a[1], a[4], a[7]}, c2[3] = {a[2], a[5], a[8]};
    double col[5];
#define SRX_OUTER(u, v) u[0] * v[0], u[1] * v[1], u[2] * v[2], u[1] * v[2], u[0] * v[2], u[0] * v[1]
    {   // beta 0: r2 (-c0c0 + c1c1)
        const double m00 = SRX_R2 * (c1[0] * c1[0] - c0[0] * c0[0]), m11 = SRX_R2 * (c1[1] * c1[1] - c0[1] * c0[1]);
        const double m22 = SRX_R2 * (c1[2] * c1[2] - c0[2] * c0[2]), m12 = SRX_R2 * (c1[1] * c1[2] - c0[1] * c0[2]);
        const double m02 = SRX_R2 * (c1[0] * c1[2] - c0[0] * c0[2]), m01 = SRX_R2 * (c1[0] * c1[1] - c0[0] * c0[1]);
        sym_to_b5(m00, m11, m22, m12, m02, m01, col);
        for (int k = 0; k < 5; ++k) Q[k][0] = col[k];
    }
    {   // beta 1: r6 (-c0c0 - c1c1 + 2 c2c2)
        double m[6];
        for (int t = 0; t < 3; ++t) m[t] = SRX_R6 * (2.0 * c2[t] * c2[t] - c0[t] * c0[t] - c1[t] * c1[t]);
        m[3] = SRX_R6 * (2.0 * c2[1] * c2[2] - c0[1] * c0[2] - c1[1] * c1[2]);
        m[4] = SRX_R6 * (2.0 * c2[0] * c2[2] - c0[0] * c0[2] - c1[0] * c1[2]);
        m[5] = SRX_R6 * (2.0 * c2[0] * c2[1] - c0[0] * c0[1] - c1[0] * c1[1]);
        sym_to_b5(m[0], m[1], m[2], m[3], m[4], m[5], col);
        for (int k = 0; k < 5; ++k) Q[k][1] = col[k];
    }
    // beta 2,3,4: r2 (u v^T + v u^T)
    const double *us[3] = {c1, c0, c0}, *vs[3] = {c2, c2, c1};
    for (int bb = 0; bb < 3; ++bb) {
        const double *u = us[bb], *v = vs[bb];
        const double m00 = SRX_R2 * 2.0 * u[0] * v[0], m11 = SRX_R2 * 2.0 * u[1] * v[1], m22 = SRX_R2 * 2.0 * u[2] * v[2];
        const double m12 = SRX_R2 * (u[1] * v[2] + v[1] * u[2]), m02 = SRX_R2 * (u[0] * v[2] + v[0] * u[2]);
        const double m01 = SRX_R2 * (u[0] * v[1] + v[0] * u[1]);
        sym_to_b5(m00, m11, m22, m12, m02, m01, col);
        for (int k = 0; k < 5; ++k) Q[k][2 + bb] = col[k];
    }
#undef SRX_OUTER
}

// sample-frame stress (B basis) -> crystal frame of a: deviatoric part Q^T lam, hydrostatic unchanged
__device__ __forceinline__ void crystal_stress(const double a[9], const double lam6[6], double lamc[6]) {
    double Q[5][5];
    q5_from(a, Q);
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        double al = 0.0;
#pragma unroll
        for (int k = 0; k < 5; ++k) al += Q[k][i] * lam6[k];
        lamc[i] = al;
    }
    lamc[5] = lam6[5];
}

// packed symmetric 5x5: index of (i,j), i<=j
__device__ __forceinline__ constexpr int ps(int i, int j) {
    return i <= j ? (i * (9 - i)) / 2 + j : (j * (9 - j)) / 2 + i;
}

// S' = Q2 diag(d) Q2^T, packed
__device__ __forceinline__ void rot_diag(const double Q2[5][5], const double d[5], double S[15]) {
    for (int i = 0; i < 5; ++i)
        for (int j = i; j < 5; ++j) {
            double acc = 0.0;
            for (int k = 0; k < 5; ++k) acc += Q2[i][k] * d[k] * Q2[j][k];
            S[ps(i, j)] = acc;
        }
}

// rotation of the compliance: crystal frame of `a` vs crystal frame of `a0`
__device__ __forceinline__ void sprime(const double a[9], const double a0[9], const double d[5], double S[15]) {
    // R = a^T a0 maps crystal0 vectors to crystal(a) vectors
    double R[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) R[3 * i + j] = a[i] * a0[j] + a[3 + i] * a0[3 + j] + a[6 + i] * a0[6 + j];
    double Q2[5][5];
    q5_from(R, Q2);
    rot_diag(Q2, d, S);
}

// ---------------------------------------------------------------------------
// constitutive helpers (constitutive.c:556-684)
// ---------------------------------------------------------------------------
// 1/x: hardware reciprocal refined by two Newton steps (correctly rounded to a
// few ulp; the IEEE division sequence is several times longer).
__device__ __forceinline__ double rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    r = fma(r, fma(-x, r, 1.0), r);
    return r;
}
// one refinement step only (relative error about the square of the hardware estimate's): the
// pivots of the Newton solve's LDL^T, where it only perturbs the step direction at the 1e-14 level
__device__ __forceinline__ double rcp_step(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}

// forest (chi1) and parallel (chi) densities of system s (constitutive.c:562-577); the
// tables are zero beyond nsys, so the fixed-length sums equal the reference's.
__device__ __forceinline__ void rho_pf(const DevPhase &P, const double rs[NSYS], int s, double &rP, double &rF) {
    rP = 0.0;
    rF = 0.0;
#pragma unroll
    for (int j = 0; j < NSYS; ++j) {
        rF += P.chi1[s][j] * rs[j];
        rP += P.chi[s][j] * rs[j];
    }
}

// GND parts of rho_F and rho_P (constitutive.c:565-576 with DD_GND): sum_j chi1_sj (|rho_g1,j
// cos_g1,sj| + |rho_g2,j cos_g2,sj|) and the chi / sin counterpart, kept per voxel in `gnd_fp` and
// refreshed whenever the GND densities change (nothing added without the GND option).
__device__ __forceinline__ void gnd_add(const DevConst *__restrict__ cst, int s, size_t N, int v, double &rP,
                                        double &rF) {
    const double *fp = cst->gnd_fp;
    rF += fp[s * N + v];
    rP += fp[(NSYS + s) * N + v];
}

// mobile density (constitutive.c:632)
// sqrt(x) and 1/sqrt(x) for normal x > 0: hardware reciprocal square root refined by two
// coupled Newton steps and a final residual correction of the root (within an ulp); the
// dislocation densities here are O(1e-3..1e4).  x = 0 gives (0, inf) as sqrt and 1/sqrt do.
__device__ __forceinline__ void sqrt_rsqrt(double x, double &sq, double &rs) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    double e = fma(-g, h, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    e = fma(-g, h, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    const double d = fma(-g, g, x);
    sq = x == 0.0 ? 0.0 : fma(d, h, g);
    rs = x == 0.0 ? y : 2.0 * h;
}
__device__ __forceinline__ double fsqrt(double x) {
    double sq, rs;
    sqrt_rsqrt(x, sq, rs);
    return sq;
}

__device__ __forceinline__ double rho_mobile(const DevPhase &P, double rP, double rF) {
    return P.rhom_k * fsqrt(rP * rF) * 160.2177;
}

// tau_pass with a given mobile density (587), 1/(tau_0 + tau_cut) (630, 639) and gamma_0 (633)
__device__ __forceinline__ void sys_const(const DevPhase &P, double rP, double rF, double rm_pass, double rmn,
                                          double &tp, double &inv, double &g0) {
    double sF, rsF;
    sqrt_rsqrt(rF, sF, rsF);
    tp = P.c1Gb3 * fsqrt(fabs(rP + rm_pass));
    inv = rcp(P.tau0 + P.k_tcut * sF);
    g0 = P.k_g0 * rmn * rsF;
}

// sinh(x) and cosh(x) from one reduction x = k ln2 + r, |r| <= ln2/2 (Cody-Waite split of ln2):
// C = cosh r and S = sinh r by their Taylor series (to r^12 and r^13, truncation below 3e-18), and
// with e^x / 2 = 2^(k-1) (C + S), e^-x / 2 = 2^(-k-1) (C - S):
//   sinh x = a S + b C,  cosh x = a C + b S,  a = 2^(k-1) + 2^(-k-1),  b = 2^(k-1) - 2^(-k-1).
// For k = 0, a = 1 and b = 0 exactly, so small arguments are the series themselves (no
// cancellation); for |k| >= 1 the two terms of sinh have the sign of x.  Accurate to a few ulp for
// either sign of x; overflows to inf where e^|x| / 2 does.
// Series coefficients of cosh (1/12! .. 1/2!, 1) and sinh (1/13! .. 1/3!, 1) read from memory: the
// uniform loads land in scalar registers, which the three-operand FMA takes as its addend directly
// (a literal addend is first copied into a vector register at every use).
__constant__ double k_sinh_coef[14] = {2.08767569878680989792e-09, 2.75573192239858906526e-07,
                                       2.48015873015873015873e-05, 1.38888888888888888889e-03,
                                       4.16666666666666666667e-02, 0.5, 1.0,
                                       1.60590438368216145994e-10, 2.50521083854417187751e-08,
                                       2.75573192239858906526e-06, 1.98412698412698412698e-04,
                                       8.33333333333333333333e-03, 1.66666666666666666667e-01, 1.0};
__device__ __forceinline__ void sinh_cosh(double x, double &sh, double &ch) {
    const double kf = rint(x * 1.4426950408889634);
    double r = fma(-kf, 6.93147180369123816490e-01, x);
    r = fma(-kf, 1.90821492927058770002e-10, r);
    const double r2 = r * r;
    const double *K = k_sinh_coef;
    double c = K[0];
#pragma unroll
    for (int i = 1; i < 7; ++i) c = fma(c, r2, K[i]);
    double o = K[7];
#pragma unroll
    for (int i = 8; i < 14; ++i) o = fma(o, r2, K[i]);
    o *= r;
    const int k = (int)kf;
    const double hp = ldexp(1.0, k - 1), hm = ldexp(1.0, -k - 1);
    const double a = hp + hm, b = hp - hm;
    sh = fma(a, o, b * c);
    ch = fma(a, c, b * o);
}

// slip rate and its derivative, FORWARD_PLUS_BACKWARD (constitutive.c:634-684):
//   g  = G sinh((|rss|-tp)/(tau0+tau_cut)) sign(rss),  g2 = G cosh(..)/(tau0+tau_cut), for |rss| > tp
// with inv = 1/(tau0+tau_cut) and e = |rss| inv - tp inv (one FMA: its sign is the activity test and
// it is the argument).  The caller folds any factor common to g and g2 into G (the Newton solve
// folds in the time step).  Branch free: inactive systems evaluate at 0 (sinh 0 = 0) and g2 is masked.
__device__ __forceinline__ void slip_rate(double rss, double e, double inv, double G, double &g, double &g2) {
    double sh, ch;
    sinh_cosh(copysign(fmax(e, 0.0), rss), sh, ch);
    g = G * sh;
    g2 = (e > 0.0 ? G * inv : 0.0) * ch;
}

// ---------------------------------------------------------------------------
// NR kernel: UpdateStress (evolution.c:157-427) in the crystal frame of the
// current orientation.  With Q the B-basis rotation, Sc = Q^T S066 Q,
// S' the grain compliance and D = d(edot)/d(sig):
//   Rt(sig) = Sc (sig - lam) + eps_p + edot(sig) dt + S' sig - eps_c  ( = S066 R )
//   K = Sc + S' + D dt  ( = S066 J, symmetric positive definite)
// so the Newton step K^-1 Rt equals the reference's J^-1 R; K is factored
// LDL^T in place.  Stopping rule, iteration cap and the pre-update iterate
// used for Edot and the errors are the reference's.
// ---------------------------------------------------------------------------
constexpr int NR_BLOCK = 64;
// slip systems evaluated together per trip of the Newton kernel's system loop (divides 12)
#ifndef SRX_NR_GROUP
#define SRX_NR_GROUP 2
#endif

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

struct NrResult {
    double cs, ce;
    double sig[6];
    int it;     // iterations of the solve that produced sig
    int itot;   // iterations executed (both attempts of a repeated solve)
    int bad;
};

// deviatoric stress block of A = Q^T S066 Q plus the compliance, packed
__device__ __forceinline__ void sc_rotate(const double *__restrict__ S, const double Q[5][5], double A[15]) {
    double T1[5][5];
#pragma unroll
    for (int i = 0; i < 5; ++i)
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < 5; ++k) acc += S[6 * i + k] * Q[k][j];
            T1[i][j] = acc;
        }
#pragma unroll
    for (int i = 0; i < 5; ++i)
#pragma unroll
        for (int j = i; j < 5; ++j) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < 5; ++k) acc += Q[k][i] * T1[k][j];
            A[ps(i, j)] = acc;
        }
}

// LDS slots of the per-system constants (tau_pass * inv, inv, gamma_0 dt of 12 systems), also the
// component order of Fields::nrc
constexpr int NRC_TP = 0, NRC_INV = NSYS, NRC_G0 = 2 * NSYS, NRC_COMPS = 3 * NSYS;

// eps_c = (DisGrad + DisGrad^T) / 2 in the B basis (chg_basis opt 2), from the 9 row-major components
__device__ __forceinline__ void ec_from_dg(const double d[9], double ec[6]) {
    const double s00 = d[0], s11 = d[4], s22 = d[8];
    const double s12 = (d[5] + d[7]) / 2.0, s02 = (d[2] + d[6]) / 2.0, s01 = (d[1] + d[3]) / 2.0;
    ec[0] = s00 * (-SRX_R2) + s11 * SRX_R2;
    ec[1] = s00 * (-SRX_R6) + s11 * (-SRX_R6) + s22 * (2. * SRX_R6);
    ec[2] = s12 * SRX_R2 + s12 * SRX_R2;
    ec[3] = s02 * SRX_R2 + s02 * SRX_R2;
    ec[4] = s01 * SRX_R2 + s01 * SRX_R2;
    ec[5] = s00 * SRX_R3 + s11 * SRX_R3 + s22 * SRX_R3;
}

// Per-voxel values every Newton call of an increment starts from, in the crystal frame of the
// current orientation: lam' = Q^T lam, q' = Q^T (eps_p - eps_c), the matrix A = Sc + S' (packed),
// r0 = q' - Sc lam' (= Q^T (eps_p - eps_c - S066 lam), Q being orthogonal), the hydrostatic parts
// and 1/||lam||^2.
struct NrPro {
    double lam[5], q[5], r0[5], A[15];
    double lamh, qh, rsig2;
};
// An empty statement that takes x in a register: x is computed before it, and the statement stays
// in program order with the memory waits around it.
__device__ __forceinline__ void pin(double &x) { asm volatile("" : "+v"(x)); }
// The prologue is formed while the copies issued before it are in flight, not after the wait for
// them (the compiler would otherwise sink the arithmetic to its first use).
__device__ __forceinline__ void pin_prologue(NrPro &p) {
#pragma unroll
    for (int i = 0; i < 5; ++i) { pin(p.lam[i]); pin(p.q[i]); pin(p.r0[i]); }
    pin(p.lamh); pin(p.qh); pin(p.rsig2);
}

// The inputs of a Newton call for one voxel: orientation (TranMat), stress, plastic strain,
// DisGrad and, after the first call of an increment, the increment's matrix A (packed).
struct NrIn {
    double a[9], lam6[6], ep[5], d[9], A[15];
};
template <bool LOAD_A>
__device__ __forceinline__ void nr_load(const Fields &f, size_t N, unsigned o8, NrIn &in) {
#pragma unroll
    for (int k = 0; k < 9; ++k) in.a[k] = ldc(f.tran + (size_t)k * N, N, o8);
#pragma unroll
    for (int k = 0; k < 6; ++k) in.lam6[k] = ldc(f.sig + (size_t)k * N, N, o8);
#pragma unroll
    for (int k = 0; k < 5; ++k) in.ep[k] = ldc(f.eps + (size_t)k * N, N, o8);
#pragma unroll
    for (int k = 0; k < 9; ++k) in.d[k] = ldc(f.dg + (size_t)k * N, N, o8);
    if (LOAD_A) {
#pragma unroll
        for (int k = 0; k < 15; ++k) in.A[k] = ldc(f.acr + (size_t)k * N, N, o8);
    }
}

// WS: Newton start 0 = lam (the reference), 1 = lam + the previous call's change (extrapolated),
// 2 = lam + the learned prediction `wsp` (crystal frame)
// Prologue of a Newton call from its inputs `in`.  FIRST (the increment's first call): A assembled
// from the orientations (Sc = Q^T S066 Q and the rotated grain compliance S') and stored in f.acr
// for the increment's later calls.  Otherwise A is in.A.
// KEEP_EC (warm starts): eps_c is also stored in f.ecp where `store` is set, for the next call's
// learned-start features.
template <bool FIRST, bool ROT, int PH, bool KEEP_EC>
__device__ __forceinline__ void nr_prologue(const DevConst *__restrict__ cst, const Fields &f, int v, size_t N,
                                            const NrIn &in, NrPro &pro, bool store = true) {
    const DevPhase &P = cst->ph[PH];
    const unsigned o8 = 8u * (unsigned)v;
    if (!FIRST) {
#pragma unroll
        for (int k = 0; k < 15; ++k) pro.A[k] = in.A[k];
    }
    const double *a = in.a;
    double Q[5][5];
    q5_from(a, Q);
    double lam6[6], w[5];
    {
        double ec[6];
        const double *ep = in.ep;
#pragma unroll
        for (int k = 0; k < 6; ++k) lam6[k] = in.lam6[k];
        // strain_aux = (DisGrad + DisGrad^T)/2 -> B basis (chg_basis opt 2)
        ec_from_dg(in.d, ec);
        if (KEEP_EC && store && f.ecp) {
#pragma unroll
            for (int k = 0; k < 6; ++k) stcf(f.ecp + (size_t)k * N, N, 4u * (unsigned)v, (float)ec[k]);
        }
#pragma unroll
        for (int k = 0; k < 5; ++k) w[k] = ep[k] - ec[k];
        pro.qh = -ec[5];
    }
    double sn = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) sn += lam6[k] * lam6[k];
    pro.rsig2 = 1.0 / sn;
    pro.lamh = lam6[5];
    // u = S066 lam (deviatoric block, sample frame)
    double u[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < 5; ++k) acc += cst->S066[i][k] * lam6[k];
        u[i] = acc;
    }
    // crystal frame: lam' = Q^T lam, q' = Q^T w, r0 = q' - Q^T u
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        double al = 0.0, aq = 0.0, au = 0.0;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            al += Q[k][i] * lam6[k];
            aq += Q[k][i] * w[k];
            au += Q[k][i] * u[k];
        }
        pro.lam[i] = al;
        pro.q[i] = aq;
        pro.r0[i] = aq - au;
    }
    if (FIRST) {
        sc_rotate(&cst->S066[0][0], Q, pro.A);
        if (ROT) {
            double a0[9], Sp[15];
#pragma unroll
            for (int k = 0; k < 9; ++k) a0[k] = ldc(f.tran0 + (size_t)(k) * N, N, o8);
            sprime(a, a0, P.sinv, Sp);
#pragma unroll
            for (int k = 0; k < 15; ++k) pro.A[k] += Sp[k];
        } else {
#pragma unroll
            for (int i = 0; i < 5; ++i) pro.A[ps(i, i)] += P.sinv[i];
        }
#pragma unroll
        for (int k = 0; k < 15; ++k) stc(f.acr + (size_t)(k) * N, N, o8, pro.A[k]);
    }
}

// WS: Newton start 0 = lam (the reference), 1 = lam + the previous call's change (extrapolated),
// 2 = lam + the learned prediction `wsp` (crystal frame).  FIRST: the call computes its prologue
// and per-system constants and stores them for the increment; otherwise the caller has put the
// prologue in `pro` and the constants in the LDS slots NRC_*.
template <bool FIRST, bool ROT, int PH, int WS, bool GND>
__device__ __forceinline__ void nr_voxel(const DevConst *__restrict__ cst, const Fields &f, int v, size_t N,
                                         double *lds, NrPro &pro, const float *wsp, NrResult &res) {
    // compile-time phase index: every material constant is a scalar load at a fixed offset
    const DevPhase &P = cst->ph[PH];
    const double dt = cst->dt;
    const unsigned o8 = 8u * (unsigned)v;
    if (FIRST) {
        NrIn in;
        nr_load<false>(f, N, o8, in);
        nr_prologue<true, ROT, PH, (WS != 0)>(cst, f, v, N, in, pro);
    }
    double (&lam)[5] = pro.lam;
    double (&q)[5] = pro.q;
    double (&A)[15] = pro.A;
    double (&r0)[5] = pro.r0;
    const double lamh = pro.lamh, qh = pro.qh, rsig2 = pro.rsig2;
    double ws[5] = {0, 0, 0, 0, 0}, wsh = 0.0;   // warm start: crystal-frame offset of the first iterate
    const double Shh = cst->S066[5][5];
    const double Ah = Shh + P.sinv[5];
    const double rAh = 1.0 / Ah;
    const double r0h = qh - Shh * lamh;

    // per-system constants, kept in LDS (lane-major, conflict free): tau_pass * inv, inv and
    // gamma_0 * dt (the slip rates enter the residual and the Jacobian times dt); the first call of
    // an increment computes them and stores them in f.nrc, from where the caller copies them into
    // LDS for the later calls.  Its own first iterate takes tau_pass with the lagged rho_m: those
    // values sit in the tau_pass slots for the first iteration, and the second reloads the
    // increment's from f.nrc (this lane's own stores).
    double *tp = lds + NRC_TP * NR_BLOCK + threadIdx.x, *inv = lds + NRC_INV * NR_BLOCK + threadIdx.x;
    double *g0 = lds + NRC_G0 * NR_BLOCK + threadIdx.x;
    if (FIRST) {
        const int nsys = P.nsys;
        double rs[NSYS];
#pragma unroll
        for (int s = 0; s < NSYS; ++s) rs[s] = ldc(f.rho_s + (size_t)(s) * N, N, o8);
#pragma unroll
        for (int s = 0; s < NSYS; ++s) {
            double ct, ci, cg;
            if (s < nsys) {
                double rP, rF;
                rho_pf(P, rs, s, rP, rF);
                if (GND) gnd_add(cst, s, N, v, rP, rF);
                const double rmn = rho_mobile(P, rP, rF);
                double t, iv, g;
                sys_const(P, rP, rF, rmn, rmn, t, iv, g);
                ct = t * iv;
                ci = iv;
                cg = g * dt;
                const double rmo = ldc(f.rho_m + (size_t)(s) * N, N, o8);
                tp[s * NR_BLOCK] = P.c1Gb3 * fsqrt(fabs(rP + rmo)) * iv;
                stc(f.rho_m_old + (size_t)(s) * N, N, o8, rmo);
                stc(f.rho_m + (size_t)(s) * N, N, o8, rmn);
            } else {
                ct = 1e300; ci = 0.0; cg = 0.0;
                tp[s * NR_BLOCK] = 1e300;
            }
            inv[s * NR_BLOCK] = ci;
            g0[s * NR_BLOCK] = cg;
            stc(f.nrc + (size_t)(NRC_TP + s) * nrc_pitch(N), nrc_pitch(N), o8, ct);
            stc(f.nrc + (size_t)(NRC_INV + s) * nrc_pitch(N), nrc_pitch(N), o8, ci);
            stc(f.nrc + (size_t)(NRC_G0 + s) * nrc_pitch(N), nrc_pitch(N), o8, cg);
        }
    }

    if (!FIRST && WS != 0) {
        // extrapolated start (WS 1): lam + (lam - lam of the previous outer iteration), the change
        // kept in the crystal frame of this increment's orientation, loaded by the caller with
        // the other inputs; learned start (WS 2): the network's offset
#pragma unroll
        for (int i = 0; i < 5; ++i) ws[i] = (double)wsp[i];
        wsh = (double)wsp[5];
    }
    double sg[5], sh = lamh + wsh, pre[5], preh = sh, ed[5];   // ed: edot dt of the iterate
#pragma unroll
    for (int i = 0; i < 5; ++i) { sg[i] = lam[i] + ws[i]; pre[i] = sg[i]; ed[i] = 0.0; }
    int it = 0, itot = 0;
    // stopping rule ||d sig|| / ||lam|| <= 1e-10, compared squared (no square root per iteration)
    double err2 = 1E-18;
    // With the learned start (WS 2), a solve that does not converge (non-finite iterate, or 500
    // iterations) is repeated from lam, the reference's start; `itot` counts both attempts.  (The
    // extrapolated start has not needed it; the second attempt costs it registers.  The first call
    // of an increment starts from lam already.)
    for (int att = 0; att < ((WS == 2 && !FIRST) ? 2 : 1); ++att) {
    if (att == 1) {
        if (err2 <= 1E-20) break;
#pragma unroll
        for (int i = 0; i < 5; ++i) { sg[i] = lam[i]; pre[i] = sg[i]; }
        sh = lamh;
        preh = sh;
        it = 0;
        err2 = 1E-18;
    }
    while (it < 500 && err2 > 1E-20) {
        ++it;
        ++itot;
        double K[15];
#pragma unroll
        for (int k = 0; k < 15; ++k) K[k] = A[k];
#pragma unroll
        for (int k = 0; k < 5; ++k) ed[k] = 0.0;
        if (FIRST && it == 2) {
#pragma unroll
            for (int s = 0; s < NSYS; ++s)
                tp[s * NR_BLOCK] = ldc(f.nrc + (size_t)(NRC_TP + s) * nrc_pitch(N), nrc_pitch(N), o8);
        }
        const double *tpu = tp;
        // SRX_NR_GROUP systems per trip, under one wavefront-uniform branch: their slip-rate
        // evaluations are independent dependency chains the scheduler interleaves (an inactive
        // system of the group evaluates at 0 and adds exact zeros)
        constexpr int NG = SRX_NR_GROUP;
        static_assert(NSYS % NG == 0, "systems are taken in groups of SRX_NR_GROUP");
#pragma unroll 1
        for (int s0 = 0; s0 < NSYS; s0 += NG) {
            double rss[NG], ivs[NG], e[NG];
            bool act = false;
#pragma unroll
            for (int u = 0; u < NG; ++u) {
                const double *p = P.p[s0 + u];
                rss[u] = p[0] * sg[0] + p[1] * sg[1] + p[2] * sg[2] + p[3] * sg[3] + p[4] * sg[4];
                ivs[u] = inv[(s0 + u) * NR_BLOCK];
                e[u] = fma(fabs(rss[u]), ivs[u], -tpu[(s0 + u) * NR_BLOCK]);
                act |= e[u] > 0.0;
            }
            if (__any(act)) {
                double g[NG], g2dt[NG];
#pragma unroll
                for (int u = 0; u < NG; ++u) slip_rate(rss[u], e[u], ivs[u], g0[(s0 + u) * NR_BLOCK], g[u], g2dt[u]);
#pragma unroll
                for (int u = 0; u < NG; ++u) {
                    const double *p = P.p[s0 + u];
#pragma unroll
                    for (int i = 0; i < 5; ++i) {
                        ed[i] += p[i] * g[u];
                        const double gp = g2dt[u] * p[i];
#pragma unroll
                        for (int j = i; j < 5; ++j) K[ps(i, j)] += gp * p[j];
                    }
                }
            }
        }
        double R[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            double acc = r0[i] + ed[i];
#pragma unroll
            for (int j = 0; j < 5; ++j) acc += A[ps(i, j)] * sg[j];
            R[i] = acc;
        }
        // LDL^T in place: K[ps(j,j)] <- 1/D_j, K[ps(i,j)] <- L_ij (i > j)
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            double dj = K[ps(j, j)];
#pragma unroll
            for (int k = 0; k < j; ++k) dj -= K[ps(j, k)] * K[ps(j, k)] * K[ps(k, k)];
            // K[ps(k,k)] holds D_k until it is inverted below
            const double rj = rcp_step(dj);
#pragma unroll
            for (int i = j + 1; i < 5; ++i) {
                double acc = K[ps(i, j)];
#pragma unroll
                for (int k = 0; k < j; ++k) acc -= K[ps(i, k)] * K[ps(j, k)] * K[ps(k, k)];
                K[ps(i, j)] = acc * rj;
            }
            K[ps(j, j)] = dj;
        }
        double y[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            double acc = R[i];
#pragma unroll
            for (int k = 0; k < i; ++k) acc -= K[ps(i, k)] * y[k];
            y[i] = acc;
        }
#pragma unroll
        for (int i = 0; i < 5; ++i) y[i] *= rcp_step(K[ps(i, i)]);
#pragma unroll
        for (int i = 4; i >= 0; --i) {
#pragma unroll
            for (int k = i + 1; k < 5; ++k) y[i] -= K[ps(k, i)] * y[k];
        }
        const double dh = (Ah * sh + r0h) * rAh;
        double cnr = dh * dh;
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            pre[i] = sg[i];
            sg[i] -= y[i];
            cnr += y[i] * y[i];
        }
        preh = sh;
        sh -= dh;
        err2 = cnr * rsig2;
    }
    }
    if (WS) {
        // this call's stress change (crystal frame) for the next call's extrapolated start
#pragma unroll
        for (int i = 0; i < 5; ++i) stcf(f.dsig + (size_t)i * N, N, 4u * (unsigned)v, (float)(sg[i] - lam[i]));
        stcf(f.dsig + 5 * N, N, 4u * (unsigned)v, (float)(sh - lamh));
    }
    // errors of evolution.c:381-391 (rotation invariant)
    double cs2 = (sh - lamh) * (sh - lamh);
#pragma unroll
    for (int i = 0; i < 5; ++i) cs2 += (sg[i] - lam[i]) * (sg[i] - lam[i]);
    res.cs = sqrt(cs2);
    {
        // the orientation again (the registers that held it were released for the Newton loop and
        // the LDS holds the per-system constants)
        double a[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) a[k] = ldc(f.tran + (size_t)(k) * N, N, o8);
        double Q[5][5];
        q5_from(a, Q);
        // eps_tot - eps_c at the pre-update iterate: q + edot dt + S' pre.  With the texture
        // rotating, S' pre = A pre - Sc pre (A = Sc + S' is still in registers, Sc = Q^T S066 Q), so
        // the initial orientation S' was built from is not read again.
        double spp[5];
        if (ROT) {
            double u[5], w[5];
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                double acc = 0.0;
#pragma unroll
                for (int k = 0; k < 5; ++k) acc += Q[i][k] * pre[k];
                u[i] = acc;
            }
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                double acc = 0.0;
#pragma unroll
                for (int k = 0; k < 5; ++k) acc += cst->S066[i][k] * u[k];
                w[i] = acc;
            }
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                double acc = 0.0;
#pragma unroll
                for (int j = 0; j < 5; ++j) acc += A[ps(i, j)] * pre[j];
#pragma unroll
                for (int k = 0; k < 5; ++k) acc -= Q[k][i] * w[k];
                spp[i] = acc;
            }
        } else {
#pragma unroll
            for (int i = 0; i < 5; ++i) spp[i] = P.sinv[i] * pre[i];
        }
        double ce2 = 0.0;
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            const double acc = q[i] + ed[i] + spp[i];
            ce2 += acc * acc;
        }
        const double eh = qh + P.sinv[5] * preh;
        ce2 += eh * eh;
        res.ce = sqrt(ce2);
        // stress back to the sample frame
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < 5; ++k) acc += Q[i][k] * sg[k];
            res.sig[i] = acc;
        }
    }
    res.sig[5] = sh;
#pragma unroll
    for (int k = 0; k < 5; ++k) stc(f.sig_pre + (size_t)(k) * N, N, o8, pre[k]);
    res.it = it;
    res.itot = itot;
    bool bad = false;
#pragma unroll
    for (int k = 0; k < 6; ++k) bad |= !isfinite(res.sig[k]);
    res.bad = bad;
}

// Waves per SIMD the register allocator targets for the Newton kernel: a second wavefront per
// SIMD to issue from (measured 31% faster than one), which the LDS of the per-system constants
// (18 KB per 64-lane block) also allows; both the first call of an increment and the later ones
// fit two without spills.
#ifndef SRX_NR_WAVES_PER_EU
#define SRX_NR_WAVES_PER_EU 2
#endif
#ifndef SRX_NR_FIRST_WAVES_PER_EU
#define SRX_NR_FIRST_WAVES_PER_EU 2
#endif

typedef __attribute__((address_space(3))) void lds_void;

// Asynchronous copies of the NC components of one field (component c at base + c P, the pitch P
// even, so that every component slice is 16-byte aligned) for the voxels [v0, v0 + 64) of this
// block into the LDS slots lds + 64 c, as raw buffer loads to LDS: the component's base in a buffer
// resource (scalar registers) and one 32-bit byte offset per lane.  One 16-byte load per lane moves
// two components (lanes 32-63 the second, at byte offset 8 P further; with an odd NC the slot after
// the last one receives a copy of it).  Lanes past the last voxel read the pitch's padding or
// re-read valid elements (the values are never used).  Every lane of the block must take part (a
// lane's copy fills other lanes' elements), and the caller waits for the copies (copy_wait).
template <int NC>
__device__ __forceinline__ void copy_field(const double *base, size_t P, double *lds, int blk) {
#if __HIP_DEVICE_COMPILE__
    const int lane = threadIdx.x;
    const int v0 = blk * NR_BLOCK;
    const unsigned pb = 8u * (unsigned)min((size_t)(v0 + 2 * (lane & 31)), P - 2);
    const unsigned hi = (lane >> 5) ? 8u * (unsigned)P : 0u;
#pragma unroll
    for (int c = 0; c < NC; c += 2) {
        const bool two = c + 1 < NC;
        __builtin_amdgcn_raw_ptr_buffer_load_lds(comp_rsrc(base + (size_t)c * P, two ? 2 * P : P),
                                                 (lds_void *)(lds + c * NR_BLOCK), 16, pb + (two ? hi : 0u), 0, 0, 0);
    }
#endif
}
__device__ __forceinline__ void copy_wait() {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
}

// Learned start of the wavefront's voxels (every lane calls; non-crystalline lanes get zero).
// Features from the Newton call's inputs (orientation, stress, DisGrad), the previous change and
// rho_s.
__device__ __forceinline__ void nr_learned_start(const DevConst *__restrict__ cst, const Fields &f, size_t N, int v,
                                                 bool xt, const NrIn &in, float wsp[6]) {
    float xf[MLP_IN], nd = 0.0f;
    if (xt) {
        double lamc[6], de[6], dec[6], rho = 0.0;
        const double *a = in.a;
        float d[6];
        crystal_stress(a, in.lam6, lamc);
        // forcing: eps_c now minus eps_c of the previous call, crystal frame
        ec_from_dg(in.d, de);
#pragma unroll
        for (int k = 0; k < 6; ++k) de[k] -= (double)f.ecp[k * N + v];
        crystal_stress(a, de, dec);
#pragma unroll
        for (int k = 0; k < 6; ++k) d[k] = f.dsig[k * N + v];
#pragma unroll
        for (int k = 0; k < NSYS; ++k) rho += ldc(f.rho_s + (size_t)(k) * N, N, 8u * (unsigned)v);
        mlp_features(a, lamc, d, dec, rho / NSYS, xf, nd);
    } else {
#pragma unroll
        for (int k = 0; k < MLP_IN; ++k) xf[k] = 0.0f;
    }
    float u[MLP_OUT];
#pragma unroll
    for (int k = 0; k < MLP_OUT; ++k) u[k] = xf[4 + k];
    mlp_normalise(f.mlp, xf);
    float y[MLP_OUT];
    MlpAcc unused;
    mlp_forward_wave<false>(f.mlp, xf, y, unused);
    double w6[6] = {0, 0, 0, 0, 0, 0};
    if (xt) mlp_start(f.mlp, y, u, nd, w6);
#pragma unroll
    for (int k = 0; k < 6; ++k) wsp[k] = (float)w6[k];
}

// One 64-voxel block of the Newton kernel: block `blk` of the `nblk` blocks of a grid of nvox
// voxels (partials: NR_NPART x nblk block partial sums).
// PH2: the second phase is crystalline (else its voxels are gas and its Newton path is not compiled).
template <bool FIRST, bool ROT, int WS, bool GND, bool PH2>
__device__ __forceinline__ void nr_block(const DevConst *__restrict__ cst, const Fields &f, int nvox,
                                         double *__restrict__ partials, const IterCtl *__restrict__ ctl, int blk,
                                         int nblk) {
    if (ctl && ctl->done) return;
    const int v = blk * NR_BLOCK + threadIdx.x;
    const size_t N = (size_t)nvox;
    const bool valid = v < nvox;
    const int ph = valid ? f.phase[v] : 0;
    // LDS: the 36 per-system constants of the block's 64 voxels (18 KB); the first call of an
    // increment computes them here, the later ones copy them from f.nrc.
    constexpr int NSLOT = NRC_COMPS;
    __shared__ __align__(16) double lds[NSLOT * NR_BLOCK];
    NrPro pro;
    float wsp[6] = {0, 0, 0, 0, 0, 0};
    if constexpr (!FIRST) {
        // After the first call of an increment: the call's inputs and A are loaded into registers,
        // then the increment's per-system constants are copied into LDS by asynchronous buffer
        // loads, so both arrive in one round trip to HBM and the prologue is formed while the
        // constants are in flight.  Every lane forms a prologue (lanes past the last voxel from
        // the last one's inputs).
        const int vv = valid ? v : nvox - 1;
        NrIn in;
        nr_load<true>(f, N, 8u * (unsigned)vv, in);
        if constexpr (WS == 1) {
#pragma unroll
            for (int k = 0; k < 6; ++k) wsp[k] = ldcf(f.dsig + (size_t)k * N, N, 4u * (unsigned)vv);
        }
        if constexpr (WS == 2)
            nr_learned_start(cst, f, N, v, valid && ph >= 1 && ph <= NPH && cst->ph[ph - 1].xtal, in, wsp);
        copy_field<NRC_COMPS>(f.nrc, nrc_pitch(N), lds, blk);
        nr_prologue<false, ROT, 0, (WS != 0)>(cst, f, vv, N, in, pro, valid);
        pin_prologue(pro);
        copy_wait();
    }
    NrResult r;
    r.cs = 0.0; r.ce = 0.0; r.it = 0; r.itot = 0; r.bad = 0;
#pragma unroll
    for (int k = 0; k < 6; ++k) r.sig[k] = 0.0;
    // one pass per distinct phase present in the wavefront, with the phase index uniform
    unsigned long long todo = __ballot(valid && ph >= 1 && ph <= NPH);
    while (todo) {
        const int first = __ffsll((unsigned long long)todo) - 1;
        const int pu = __builtin_amdgcn_readfirstlane(__shfl(ph, first, 64));
        const bool mine = valid && ph == pu;
        if (mine) {
            if (pu == 1) {
                if (cst->ph[0].xtal) nr_voxel<FIRST, ROT, 0, WS, GND>(cst, f, v, N, lds, pro, wsp, r);
            } else if constexpr (PH2) {
                if (cst->ph[1].xtal) nr_voxel<FIRST, ROT, 1, WS, GND>(cst, f, v, N, lds, pro, wsp, r);
            }
        }
        todo &= ~__ballot(mine);
    }
    if (valid) {
        const unsigned o8 = 8u * (unsigned)v;
        // gas voxels: Sig = Edot = 0 (evolution.c:418-423)
#pragma unroll
        for (int k = 0; k < 6; ++k) stc(f.sig + (size_t)(k) * N, N, o8, r.sig[k]);
        f.nr_it[v] = r.it;
    }
    // block partials: conv_s, conv_e, sig6, sig6 of phase 1, nr_sum, nr_max, nan, wavefront nr_max
    const int lane = threadIdx.x & 63;
    double red[NR_NPART];
    red[0] = wave_sum(r.cs);
    red[1] = wave_sum(r.ce);
#pragma unroll
    for (int k = 0; k < 6; ++k) red[2 + k] = wave_sum(r.sig[k]);
    // when every voxel of the wavefront is phase 1 the phase-1 sums are the full sums
    if (__any(valid && ph != 1)) {
#pragma unroll
        for (int k = 0; k < 6; ++k) red[8 + k] = wave_sum((valid && ph == 1) ? r.sig[k] : 0.0);
    } else {
#pragma unroll
        for (int k = 0; k < 6; ++k) red[8 + k] = red[2 + k];
    }
    red[14] = wave_sum((double)r.itot);
    // the wavefront's largest iteration count (the lane-iterations it executed / 64)
    red[15] = wave_max((double)r.itot);
    red[16] = __any(r.bad) ? 1.0 : 0.0;
    red[17] = red[15];
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NR_NPART; ++k) partials[(size_t)k * nblk + blk] = red[k];
    }
}

template <bool FIRST, bool ROT, int WS, bool GND, bool PH2>
__global__ void __launch_bounds__(NR_BLOCK)
__attribute__((amdgpu_waves_per_eu(FIRST ? SRX_NR_FIRST_WAVES_PER_EU : SRX_NR_WAVES_PER_EU, 8)))
k_nr(const DevConst *__restrict__ cst, Fields f, int nvox, double *__restrict__ partials,
     const IterCtl *__restrict__ ctl) {
    nr_block<FIRST, ROT, WS, GND, PH2>(cst, f, nvox, partials, ctl, blockIdx.x, gridDim.x);
}

// The Newton kernel over every member of an ensemble in one launch: member m owns the blocks
// [tab[m].blk0, tab[m].blk0 + tab[m].nblk) of the grid.
template <bool FIRST, bool ROT, int WS, bool GND, bool PH2>
__global__ void __launch_bounds__(NR_BLOCK)
__attribute__((amdgpu_waves_per_eu(FIRST ? SRX_NR_FIRST_WAVES_PER_EU : SRX_NR_WAVES_PER_EU, 8)))
k_nr_ens(const EnsMember *__restrict__ tab, int nmem) {
    int m = 0;
    while (m + 1 < nmem && (int)blockIdx.x >= tab[m + 1].blk0) ++m;
    const EnsMember &e = tab[m];
    nr_block<FIRST, ROT, WS, GND, PH2>(e.cst, e.f, e.nvox, e.partials, e.ctl, blockIdx.x - e.blk0, e.nblk);
}

// fixed-order reduction of block partials stored component-major:
// out[k] = sum over blocks of part[k][b] (max for k in maxmask); one block per k.  Each thread
// keeps eight independent accumulators over a fixed stride, combined in a fixed order, so the
// result is the same on every run.
constexpr int RED_BLOCK = 1024;
__global__ void __launch_bounds__(RED_BLOCK) k_reduce(const double *__restrict__ part, int nblocks, int nvals,
                                                      unsigned long long maxmask, double *__restrict__ out,
                                                      const IterCtl *__restrict__ ctl) {
    if (ctl && ctl->done) return;
    __shared__ double sh[RED_BLOCK];
    const int k = blockIdx.x;
    const bool ismax = (maxmask >> k) & 1ull;
    const double *row = part + (size_t)k * nblocks;
    const double init = ismax ? -1e308 : 0.0;
    double acc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] = init;
    for (int b0 = threadIdx.x; b0 < nblocks; b0 += 8 * RED_BLOCK) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int b = b0 + u * RED_BLOCK;
            if (b < nblocks) acc[u] = ismax ? fmax(acc[u], row[b]) : acc[u] + row[b];
        }
    }
    double a = acc[0];
#pragma unroll
    for (int u = 1; u < 8; ++u) a = ismax ? fmax(a, acc[u]) : a + acc[u];
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int st = RED_BLOCK / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st)
            sh[threadIdx.x] = ismax ? fmax(sh[threadIdx.x], sh[threadIdx.x + st]) : sh[threadIdx.x] + sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = sh[0];
    (void)nvals;
}

// First stage of the per-iteration reduction: block (k, sg) reduces value k over the sg-th of
// RED_SEGS contiguous ranges of block partials, in a fixed order (strided accumulators, then the
// wavefronts, then the four wavefront results), into seg[k * RED_SEGS + sg].  k_control finishes.
constexpr int RED_SEGS = 16, RED_SEG_BLOCK = 256;
static_assert(NR_RED_SEG_DOUBLES == NR_NPART * RED_SEGS, "segment buffer size");
__device__ __forceinline__ void reduce_seg_block(const double *__restrict__ part, int nblocks,
                                                 unsigned long long maxmask, double *__restrict__ seg,
                                                 const IterCtl *__restrict__ ctl, int k, int sg) {
    if (ctl && ctl->done) return;
    const bool ismax = (maxmask >> k) & 1ull;
    const int chunk = (nblocks + RED_SEGS - 1) / RED_SEGS;
    const int b1 = min(nblocks, (sg + 1) * chunk);
    const double *row = part + (size_t)k * nblocks;
    const double init = ismax ? -1e308 : 0.0;
    double acc[4] = {init, init, init, init};
    for (int b0 = sg * chunk + threadIdx.x; b0 < b1; b0 += 4 * RED_SEG_BLOCK) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int b = b0 + u * RED_SEG_BLOCK;
            if (b < b1) acc[u] = ismax ? fmax(acc[u], row[b]) : acc[u] + row[b];
        }
    }
    double a = ismax ? fmax(fmax(acc[0], acc[1]), fmax(acc[2], acc[3])) : (acc[0] + acc[1]) + (acc[2] + acc[3]);
    a = ismax ? wave_max(a) : wave_sum(a);
    __shared__ double sh[RED_SEG_BLOCK / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = sh[0];
        for (int w = 1; w < RED_SEG_BLOCK / 64; ++w) r = ismax ? fmax(r, sh[w]) : r + sh[w];
        seg[k * RED_SEGS + sg] = r;
    }
}

__global__ void __launch_bounds__(RED_SEG_BLOCK) k_reduce_seg(const double *__restrict__ part, int nblocks,
                                                              unsigned long long maxmask, double *__restrict__ seg,
                                                              const IterCtl *__restrict__ ctl) {
    reduce_seg_block(part, nblocks, maxmask, seg, ctl, blockIdx.x, blockIdx.y);
}

// the same for every member of an ensemble: grid (NR_NPART, RED_SEGS, members)
__global__ void __launch_bounds__(RED_SEG_BLOCK) k_reduce_seg_ens(const EnsMember *__restrict__ tab,
                                                                  unsigned long long maxmask) {
    const EnsMember &e = tab[blockIdx.z];
    reduce_seg_block(e.partials, e.nblk, maxmask, e.redseg, e.ctl, blockIdx.x, blockIdx.y);
}

// One outer iteration's bookkeeping on the device (the host loop of evolution.c:1150-1256):
// the dDisGradAvg this iteration's inverse-z pass applied joins the uniform part, the reduced
// sums give the errors and get_smacro's averages and correction, and `done` decides whether the
// next iteration runs.  One thread; the iteration's other kernels have finished (stream order).
__device__ __forceinline__ void control_step(const DevConst *__restrict__ cst, IterCtl &c, const double *red);

__device__ __forceinline__ void control_block(const DevConst *__restrict__ cst, const double *__restrict__ seg,
                                              IterCtl *ctl) {
    if (ctl->done) return;
    // The control block is copied to LDS by the whole block, worked on there by one thread (no
    // dependent HBM round trips) and copied back.
    static_assert(sizeof(IterCtl) % 8 == 0, "IterCtl is copied as 64-bit words");
    constexpr int NW = sizeof(IterCtl) / 8;
    __shared__ IterCtl lc;
    for (int i = threadIdx.x; i < NW; i += blockDim.x)
        reinterpret_cast<unsigned long long *>(&lc)[i] = reinterpret_cast<const unsigned long long *>(ctl)[i];
    // second stage of the reduction: the RED_SEGS partial results of each value, in order
    __shared__ double red[NR_NPART];
    if (threadIdx.x < NR_NPART) {
        const int k = threadIdx.x;
        const bool ismax = k == 15;
        double r = seg[k * RED_SEGS];
        for (int sg = 1; sg < RED_SEGS; ++sg) r = ismax ? fmax(r, seg[k * RED_SEGS + sg]) : r + seg[k * RED_SEGS + sg];
        red[k] = r;
    }
    __syncthreads();
    if (threadIdx.x == 0) control_step(cst, lc, red);
    __syncthreads();
    for (int i = threadIdx.x; i < NW; i += blockDim.x)
        reinterpret_cast<unsigned long long *>(ctl)[i] = reinterpret_cast<const unsigned long long *>(&lc)[i];
}

__global__ void k_control(const DevConst *__restrict__ cst, const double *__restrict__ seg, IterCtl *ctl) {
    control_block(cst, seg, ctl);
}

// the same for every member of an ensemble (one block each); done[m] receives the member's stop
// flag, the only value the host reads between batches of iterations
__global__ void k_control_ens(const EnsMember *__restrict__ tab, int *__restrict__ done) {
    const EnsMember &e = tab[blockIdx.x];
    control_block(e.cst, e.redseg, e.ctl);
    __syncthreads();
    if (threadIdx.x == 0) done[blockIdx.x] = e.ctl->done;
}

// one outer iteration's bookkeeping on the reduced values red[] (thread 0 of k_control)
__device__ __forceinline__ void control_step(const DevConst *__restrict__ cst, IterCtl &c, const double *red) {
    for (int k = 0; k < 9; ++k) c.uniform[k] += c.ddg[k];
    c.iter += 1;
    if (red[16] > 0) {   // a non-finite stress (the reference asserts, constitutive.c:683)
        c.nan = 1;
        c.done = 1;
        return;
    }
    for (int k = 0; k < 6; ++k) c.sig_sum6[k] = red[2 + k];
    c.nrtot += (long long)red[14];
    c.nrmax = max(c.nrmax, (int)red[15]);
    c.nrwave += red[17];
    c.Err_e = red[1] * c.wgt;
    c.Err_s = red[0] * c.wgt;
    smacro_update(c, cst->S0, red + 2, red + 8);
    if (fabs(c.e_vm_prev) < 1E-5) c.Err_e = 0.0;
    else c.Err_e /= c.e_vm_prev;
    c.Err_s /= c.s_vm;
    for (int k = 0; k < 6; ++k) {
        c.mean6[k] = c.mixed ? c.sig_sum6[k] / c.nvox : 0.0;
        c.dcs6[k] = c.mixed ? c.mean6[k] * c.nvox : 0.0;
    }
    c.done = ctl_stop(c);
}

// increment prologue (evolution.c:1115-1128): dg0 = dg; dg += Udot dt (or = set9 under creep)
struct Vec9 {
    double v[9];
};

// evolution.c:1115-1128.  In a creep run DisGrad is reset to the macroscopic value; the
// fluctuation it held (DisGrad minus the uniform part uold) is carried in fl0 so that the
// accumulated fluctuation of evolution.c:1198 survives the reset.
__global__ void k_prologue(Fields f, int nvox, Vec9 add, int creep, Vec9 set, Vec9 uold) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    const size_t N = (size_t)nvox;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const double d = f.dg[k * N + v];
        f.dg0[k * N + v] = d;
        f.dg[k * N + v] = creep ? set.v[k] : d + add.v[k];
        if (creep && f.fl0) f.fl0[k * N + v] += d - uold.v[k];
    }
}

// ---------------------------------------------------------------------------
// end of increment: Edot/Eps (evolution.c:1297-1301), update_orient
// (constitutive.c:142-273), DislocationEvolution (constitutive.c:1353-1529)
// and the partial sums of AvgDislocation and the Eps/Edot averages.
// ---------------------------------------------------------------------------
constexpr int END_BLOCK = 128;
// slip systems per trip of the end-of-increment kernel's slip-rate loop
#ifndef SRX_END_UNROLL
#define SRX_END_UNROLL 1
#endif

template <int PH, bool GND>
__device__ __forceinline__ void end_voxel(const DevConst *__restrict__ cst, const Fields &f, int v, size_t N,
                                          int lagged_possible, int do_orient, int istep, double *vals, double *rsl) {
    const DevPhase &P = cst->ph[PH];
    const double dt = cst->dt;
    const unsigned o8 = 8u * (unsigned)v;   // byte offset of this voxel's element of a component
    double a[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) a[k] = ldc(f.tran + (size_t)(k) * N, N, o8);
    const bool lagged = lagged_possible && f.nr_it[v] == 1;
    double pre[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) pre[k] = ldc(f.sig_pre + (size_t)(k) * N, N, o8);
    // final stress in the crystal frame (Schmid tensors of this increment)
    double sf[5];
    {
        double Q[5][5];
        q5_from(a, Q);
        double s6[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) s6[k] = ldc(f.sig + (size_t)(k) * N, N, o8);
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < 5; ++k) acc += Q[k][i] * s6[k];
            sf[i] = acc;
        }
    }
    const int nsys = P.nsys;
    // rho_s at the start of the increment into this lane's LDS column (rsl, stride END_BLOCK): every
    // system's rho_F and rho_P sum all of them, and the system loop below updates one at a time;
    // rho_tot of evolution.c:1040-1046, from the densities at the start of the increment
    double rho_tot = 0.0;
#pragma unroll
    for (int s = 0; s < NSYS; ++s) {
        const double rs = ldc(f.rho_s + (size_t)(s) * N, N, o8);
        rsl[s * END_BLOCK] = rs;
        if (s < nsys) {
            double g1 = 0.0, g2 = 0.0;
            if (GND) {
                g1 = ldc(cst->rho_g1 + (size_t)(s) * N, N, o8);
                g2 = ldc(cst->rho_g2 + (size_t)(s) * N, N, o8);
            }
            rho_tot += rs + ldc(f.rho_m_old + (size_t)(s) * N, N, o8) + sqrt(g1 * g1 + g2 * g2 + 0.0 * 0.0);
        }
    }
    stc(f.rho_tot, N, o8, rho_tot);
    double edc[5] = {0, 0, 0, 0, 0};
    double sumF = 0, sumP = 0, sumM = 0;
    double sR = 0, sD = 0, sFo = 0, sI = 0;
    // plastic spin sum_s b_s (x) n_s gamdot_s of update_orient (constitutive.c:215-250), with the
    // pre-update TranMat, accumulated with the slip rates so they need not stay in registers
    double lp[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    bool bad = false;
    const double G = P.G, nu = cst->nu, T = cst->T, D0 = cst->D0;
    // One trip per system: slip rate (StrainRate_Orowan at the pre-update iterate, constitutive.c
    // 507-728) and the hardening update of DislocationEvolution (istep <= 2e6 branch,
    // constitutive.c:1400-1426) with that slip rate and rho_F.  The per-system arrays of the next
    // system are loaded one trip ahead.
    double nx_m = ldc(f.rho_m, N, o8), nx_dip = ldc(f.rho_dipole, N, o8), nx_for = ldc(f.rho_forest, N, o8),
           nx_inp = ldc(f.rho_inplane, N, o8), nx_acc = ldc(f.gam_acum, N, o8);
#pragma unroll 1
    for (int s = 0; s < nsys; ++s) {
        const double rmn = nx_m, rdip0 = nx_dip, rfor0 = nx_for, rinp0 = nx_inp, gacc = nx_acc;
        {
            const size_t c = (size_t)min(s + 1, nsys - 1) * N;
            nx_m = ldc(f.rho_m + c, N, o8);
            nx_dip = ldc(f.rho_dipole + c, N, o8);
            nx_for = ldc(f.rho_forest + c, N, o8);
            nx_inp = ldc(f.rho_inplane + c, N, o8);
            nx_acc = ldc(f.gam_acum + c, N, o8);
        }
        double rP = 0.0, rF = 0.0;
#pragma unroll
        for (int j = 0; j < NSYS; ++j) {
            const double rj = rsl[j * END_BLOCK];
            rF += P.chi1[s][j] * rj;
            rP += P.chi[s][j] * rj;
        }
        if (GND) gnd_add(cst, s, N, v, rP, rF);
        const double rmo = lagged ? ldc(f.rho_m_old + (size_t)(s) * N, N, o8) : rmn;
        double tp, inv, g0;
        sys_const(P, rP, rF, rmo, rmn, tp, inv, g0);
        const double *p = P.p[s];
        const double rss = p[0] * pre[0] + p[1] * pre[1] + p[2] * pre[2] + p[3] * pre[3] + p[4] * pre[4];
        double g, g2;
        slip_rate(rss, fma(fabs(rss), inv, -(tp * inv)), inv, g0, g, g2);
        bad |= !isfinite(g);
#pragma unroll
        for (int i = 0; i < 5; ++i) edc[i] += p[i] * g;
        if (do_orient) {
            double bs[3], ns[3];
            for (int i = 0; i < 3; ++i) {
                bs[i] = a[3 * i] * P.bb[s][0] + a[3 * i + 1] * P.bb[s][1] + a[3 * i + 2] * P.bb[s][2];
                ns[i] = a[3 * i] * P.bn[s][0] + a[3 * i + 1] * P.bn[s][1] + a[3 * i + 2] * P.bn[s][2];
            }
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) lp[i][j] += bs[i] * ns[j] * g;
        }
        stc(f.rho_F + (size_t)(s) * N, N, o8, rF);
        stc(f.rho_P + (size_t)(s) * N, N, o8, rP);
        stc(f.gamdot + (size_t)(s) * N, N, o8, g);
        sumF += rF;
        sumP += rP;
        sumM += rmn;
        // hardening, with the final stress in the crystal frame
        const double rssf = p[0] * sf[0] + p[1] * sf[1] + p[2] * sf[2] + p[3] * sf[3] + p[4] * sf[4];
        const double bb = P.b;
        const double d_dip = sqrt(3.0) * G * bb / 16.0 / 3.14159265359 / (1 - nu) / (fabs(rssf) + 1.0E-4);
        const double hydro = G * bb / 3.14159265359 / (1 - nu) / d_dip;
        double r = rsl[s * END_BLOCK];
        double rdip = rdip0, rfor = rfor0, rinp = rinp0;
        const double gd = g;
        const double rFs = rF;
        if (istep <= 2000000) {
            const double climb = 1E-17 * rho_tot * 0.5 * 1E24 * (d_dip) * (D0 / bb) * pow(r, 2.0) *
                                 (exp(fabs(hydro) * bb * bb * bb * 100 / (1.38 * T)) - 1.0) * P.expmqb;
            rdip += climb * dt;
            rfor += fabs(gd) * (P.c4 * sqrt(rFs)) * dt;
            rinp += fabs(gd) * (P.c5 * r) * dt;
            const double dot = fabs(gd) * (P.c4 * sqrt(rFs) + P.c6 * (d_dip) * rmn - P.c5 * r) - climb;
            r += dot * dt;
        }
        stc(f.rho_s + (size_t)(s) * N, N, o8, r);
        stc(f.rho_dipole + (size_t)(s) * N, N, o8, rdip);
        stc(f.rho_forest + (size_t)(s) * N, N, o8, rfor);
        stc(f.rho_inplane + (size_t)(s) * N, N, o8, rinp);
        stc(f.gam_acum + (size_t)(s) * N, N, o8, gacc + gd * dt);
        sR += r;
        sD += rdip;
        sFo += rfor;
        sI += rinp;
    }
    // Edot in the sample frame and Eps += Edot dt (the rotation is rebuilt rather than kept in
    // registers through the system loop)
    double Q[5][5];
    q5_from(a, Q);
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < 5; ++k) acc += Q[i][k] * edc[k];
        stc(f.edot + (size_t)(i) * N, N, o8, acc);
        const double e = ldc(f.eps + (size_t)(i) * N, N, o8) + acc * dt;
        stc(f.eps + (size_t)(i) * N, N, o8, e);
        vals[(i) * END_BLOCK] = e;
        vals[(6 + i) * END_BLOCK] = acc;
    }
    vals[(5) * END_BLOCK] = 0.0;
    vals[(11) * END_BLOCK] = 0.0;
    // reorientation (update_orient), with the pre-update TranMat for b and n
    if (do_orient) {
        double rl[3][3];
        double vg[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) vg[k] = (ldc(f.dg + (size_t)(k) * N, N, o8) - ldc(f.dg0 + (size_t)(k) * N, N, o8)) / dt;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) rl[i][j] = (vg[3 * i + j] - vg[3 * j + i]) / 2.;
        double rot[3][3], rsw[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) rsw[i][j] = (lp[i][j] - lp[j][i]) / 2.;
        vals[(18) * END_BLOCK] = sqrt(rsw[2][1] * rsw[2][1] + rsw[0][2] * rsw[0][2] + rsw[1][0] * rsw[1][0]);
        vals[(19) * END_BLOCK] = sqrt(rl[2][1] * rl[2][1] + rl[0][2] * rl[0][2] + rl[1][0] * rl[1][0]);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) rot[i][j] = (cst->Udot_a[3 * i + j] + rl[i][j] - rsw[i][j]) * dt;
        // Orientation (constitutive.c:142-194)
        double w[3] = {rot[2][1], rot[0][2], rot[1][0]};
        double wn = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        double wt = tan(wn / 2.0);
        if (fabs(wn) < 1E-6) wn = 1.0;
        for (int i = 0; i < 3; ++i) w[i] *= wt / wn;
        const double om[3][3] = {{0., -w[2], w[1]}, {w[2], 0., -w[0]}, {-w[1], w[0], 0.}};
        double om2[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) om2[i][j] = om[i][0] * om[0][j] + om[i][1] * om[1][j] + om[i][2] * om[2][j];
        wt = wt * wt;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) rot[i][j] = (double)(i == j) + 2. * (om[i][j] + om2[i][j]) / (1. + wt);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                stc(f.tran + (size_t)((3 * i + j)) * N, N, o8, rot[i][0] * a[j] + rot[i][1] * a[3 + j] + rot[i][2] * a[6 + j]);
    }
    vals[(12) * END_BLOCK] = sumF;
    vals[(13) * END_BLOCK] = sumP;
    vals[(14) * END_BLOCK] = sR;
    vals[(15) * END_BLOCK] = sumM;
    vals[(16) * END_BLOCK] = sD;
    vals[(17) * END_BLOCK] = sFo;
    vals[(20) * END_BLOCK] = sI;
    vals[(21) * END_BLOCK] = bad ? 1.0 : 0.0;
}

#ifndef SRX_END_WAVES_PER_EU
#define SRX_END_WAVES_PER_EU 2
#endif
template <bool GND>
__global__ void __launch_bounds__(END_BLOCK) __attribute__((amdgpu_waves_per_eu(SRX_END_WAVES_PER_EU, 8))) k_end(const DevConst *__restrict__ cst, Fields f, int nvox,
                                                   int lagged_possible, int do_orient, int istep,
                                                   double *__restrict__ partials) {
    const int v = blockIdx.x * END_BLOCK + threadIdx.x;
    const size_t N = (size_t)nvox;
    const bool valid = v < nvox;
    const int ph = valid ? f.phase[v] : 0;
    // this lane's partial values, in an LDS column (stride END_BLOCK) rather than registers
    __shared__ double lv[END_NPART * END_BLOCK];
    __shared__ double lrs[NSYS * END_BLOCK];
    double *vals = lv + threadIdx.x;
#pragma unroll
    for (int k = 0; k < END_NPART; ++k) vals[k * END_BLOCK] = 0.0;
    unsigned long long todo = __ballot(valid && ph >= 1 && ph <= NPH);
    while (todo) {
        const int first = __ffsll((unsigned long long)todo) - 1;
        const int pu = __builtin_amdgcn_readfirstlane(__shfl(ph, first, 64));
        const bool mine = valid && ph == pu;
        if (mine) {
            if (pu == 1) {
                if (cst->ph[0].xtal) end_voxel<0, GND>(cst, f, v, N, lagged_possible, do_orient, istep, vals, lrs + threadIdx.x);
            } else {
                if (cst->ph[1].xtal) end_voxel<1, GND>(cst, f, v, N, lagged_possible, do_orient, istep, vals, lrs + threadIdx.x);
            }
        }
        todo &= ~__ballot(mine);
    }
    const bool xtal = valid && ph >= 1 && ph <= NPH && cst->ph[ph - 1].xtal;
    if (valid && !xtal) {
        // gas: Edot = 0, Eps unchanged (it is zero), gamdot untouched
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            stc(f.edot + (size_t)(i) * N, N, 8u * (unsigned)v, 0.0);
            vals[i * END_BLOCK] = ldc(f.eps + (size_t)(i) * N, N, 8u * (unsigned)v);
        }
    }
    __shared__ double red[END_BLOCK / 64][END_NPART];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < END_NPART; ++k) {
        const double s = wave_sum(vals[k * END_BLOCK]);
        if (lane == 0) red[wid][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < END_NPART) {
        const int k = threadIdx.x;
        double acc = red[0][k];
        for (int w = 1; w < END_BLOCK / 64; ++w) acc += red[w][k];
        partials[(size_t)k * gridDim.x + blockIdx.x] = acc;
    }
}

// ---------------------------------------------------------------------------
// GND option (a DD_GND build, first order): after the hardening update of an increment,
//   Gradient_ShearRate (constitutive.c:882-1115): per system, the one-sided or central difference
//   of gamdot along x, y, z inside the grain (the same-grain flags of Gradient_ExchangeGrainID,
//   constitutive.c:731-880, periodic neighbours, with its z- flag taken from the y- neighbour);
//   computed when Hard_Flag = 1 and istep > 1, zero before (init.c:1083);
//   DislocationEvolution's GND branch (constitutive.c:1428-1519 with DD_GND_FIRST_ORDER):
//   rho_g1 += (grad . bp) / b / L0 dt (screw), rho_g2 += -(grad . bb) / b / L0 dt (edge-I),
//   edge-II += 0;
// then the GND parts of rho_F and rho_P of the new densities, and per block the sum of
// sqrt(rho_g1^2 + rho_g2^2) for AvgDislocation's GND mean (constitutive.c:1745-1757).
// One thread per voxel; the neighbours' gamdot rows are read from HBM (L2 serves the reuse).
// ---------------------------------------------------------------------------
constexpr int GND_BLOCK = 256;

__device__ __forceinline__ double gb_diff(bool c0, bool c1, double v, double vp, double vm) {
    return (!c0 && c1) ? v - vm : (c0 && !c1) ? vp - v : (c0 && c1) ? (vp - vm) / 2.0 : 0.0;
}

__global__ void __launch_bounds__(GND_BLOCK) k_gnd(const DevConst *__restrict__ cst, const int *__restrict__ phase,
                                                   const double *__restrict__ gamdot, const unsigned char *__restrict__ gbc,
                                                   double *__restrict__ rho_g1, double *__restrict__ rho_g2,
                                                   double *__restrict__ gnd_fp, int nvox, int grad,
                                                   double *__restrict__ partials) {
    const int v = blockIdx.x * GND_BLOCK + threadIdx.x;
    const size_t N = (size_t)nvox;
    double acc = 0.0;
    const int ph = v < nvox ? phase[v] : 0;
    if (v < nvox && ph >= 1 && ph <= NPH && cst->ph[ph - 1].xtal) {
        const DevPhase &P = cst->ph[ph - 1];
        const int n0 = cst->n[0], n1 = cst->n[1], n2 = cst->n[2];
        const int x = v / (n1 * n2), y = (v / n2) % n1, z = v % n2;
        const int xp = x + 1 >= n0 ? 0 : x + 1, xm = x - 1 < 0 ? n0 - 1 : x - 1;
        const int yp = y + 1 >= n1 ? 0 : y + 1, ym = y - 1 < 0 ? n1 - 1 : y - 1;
        const int zp = z + 1 >= n2 ? 0 : z + 1, zm = z - 1 < 0 ? n2 - 1 : z - 1;
        const size_t pxp = ((size_t)xp * n1 + y) * n2 + z, pxm = ((size_t)xm * n1 + y) * n2 + z;
        const size_t pyp = ((size_t)x * n1 + yp) * n2 + z, pym = ((size_t)x * n1 + ym) * n2 + z;
        const size_t pzp = ((size_t)x * n1 + y) * n2 + zp, pzm = ((size_t)x * n1 + y) * n2 + zm;
        const unsigned b = gbc[v];
        const double dt = cst->dt, L0 = cst->L0;
        const int nsys = P.nsys;
        if (grad) {
#pragma unroll
            for (int i = 0; i < NSYS; ++i) {
                if (i >= nsys) continue;
                const double *gd = gamdot + (size_t)i * N;
                const double g0 = gd[v];
                const double gx = gb_diff(b & 1, b & 2, g0, gd[pxp], gd[pxm]);
                const double gy = gb_diff(b & 4, b & 8, g0, gd[pyp], gd[pym]);
                const double gz = gb_diff(b & 16, b & 32, g0, gd[pzp], gd[pzm]);
                const double s1 = gx * P.bp[i][0] + gy * P.bp[i][1] + gz * P.bp[i][2];
                rho_g1[(size_t)i * N + v] += s1 / P.b / L0 * dt;
                const double s2 = gx * P.bb[i][0] + gy * P.bb[i][1] + gz * P.bb[i][2];
                rho_g2[(size_t)i * N + v] += -1.0 * s2 / P.b / L0 * dt;
            }
        }
        // |rho_g| of the phase's systems (zero beyond nsys, where the tables are zero too); the
        // projection tables are non-negative, so |rho_g cos| = |rho_g| cos exactly
        double g1[NSYS], g2[NSYS];
#pragma unroll
        for (int j = 0; j < NSYS; ++j) {
            g1[j] = j < nsys ? fabs(rho_g1[(size_t)j * N + v]) : 0.0;
            g2[j] = j < nsys ? fabs(rho_g2[(size_t)j * N + v]) : 0.0;
        }
#pragma unroll
        for (int i = 0; i < NSYS; ++i) {
            double dF = 0.0, dP = 0.0;
#pragma unroll
            for (int j = 0; j < NSYS; ++j) {
                dF += P.chi1[i][j] * (g1[j] * P.cos_g1[i][j] + g2[j] * P.cos_g2[i][j]);
                dP += P.chi[i][j] * (g1[j] * P.sin_g1[i][j] + g2[j] * P.sin_g2[i][j]);
            }
            acc += sqrt(g1[i] * g1[i] + g2[i] * g2[i] + 0.0 * 0.0);
            gnd_fp[(size_t)i * N + v] = dF;
            gnd_fp[(size_t)(NSYS + i) * N + v] = dP;
        }
    }
    __shared__ double red[GND_BLOCK / 64];
    const double w = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = red[0];
        for (int k = 1; k < GND_BLOCK / 64; ++k) t += red[k];
        partials[blockIdx.x] = t;
    }
}

// ---------------------------------------------------------------------------
// Training samples for the learned Newton start: voxel j * stride (j < nsamp).  Before a Newton
// call: features and the scale |d| of the previous change, at slot *count + j.  After it: the
// target, this call's change over that scale minus the previous change's direction (the
// correction to the extrapolated start), and the count advanced by nsamp (recording stops when
// the buffer is full).
// ---------------------------------------------------------------------------
__global__ void k_mlp_rec_x(const DevConst *__restrict__ cst, Fields f, int nvox, int stride, int nsamp,
                            float *__restrict__ X, float *__restrict__ ND, const int *__restrict__ count, int cap,
                            const IterCtl *__restrict__ ctl) {
    if (ctl && ctl->done) return;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int base = *count;
    if (j >= nsamp || base + nsamp > cap) return;
    const int v = j * stride;
    const size_t N = (size_t)nvox;
    const int ph = f.phase[v];
    float x[MLP_IN], nd = 0.0f;
    if (ph >= 1 && ph <= NPH && cst->ph[ph - 1].xtal) {
        double a[9], lam6[6], lamc[6], dg[9], de[6], dec[6], rho = 0.0;
        float d[6];
        for (int k = 0; k < 9; ++k) a[k] = f.tran[k * N + v];
        for (int k = 0; k < 6; ++k) lam6[k] = f.sig[k * N + v];
        crystal_stress(a, lam6, lamc);
        for (int k = 0; k < 9; ++k) dg[k] = f.dg[k * N + v];
        ec_from_dg(dg, de);
        for (int k = 0; k < 6; ++k) de[k] -= (double)f.ecp[k * N + v];
        crystal_stress(a, de, dec);
        for (int k = 0; k < 6; ++k) d[k] = f.dsig[k * N + v];
        for (int k = 0; k < NSYS; ++k) rho += f.rho_s[k * N + v];
        mlp_features(a, lamc, d, dec, rho / NSYS, x, nd);
    } else {
        for (int k = 0; k < MLP_IN; ++k) x[k] = 0.0f;
    }
    for (int k = 0; k < MLP_IN; ++k) X[(size_t)(base + j) * MLP_IN + k] = x[k];
    ND[base + j] = nd;
}

__global__ void k_mlp_rec_t(Fields f, int nvox, int stride, int nsamp, const float *__restrict__ X,
                            float *__restrict__ T, const float *__restrict__ ND, const int *__restrict__ count, int cap,
                            const IterCtl *__restrict__ ctl) {
    if (ctl && ctl->done) return;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int base = *count;
    if (j >= nsamp || base + nsamp > cap) return;
    const int v = j * stride;
    const size_t N = (size_t)nvox;
    const float nd = ND[base + j];
    const float *u = X + (size_t)(base + j) * MLP_IN + 4;
    for (int k = 0; k < MLP_OUT; ++k)
        T[(size_t)(base + j) * MLP_OUT + k] = nd > 0.0f ? f.dsig[k * N + v] / nd - u[k] : 0.0f;
}

__global__ void k_mlp_rec_adv(int *__restrict__ count, int cap, int nsamp, const IterCtl *__restrict__ ctl) {
    if (ctl && ctl->done) return;
    if (threadIdx.x == 0 && *count + nsamp <= cap) *count += nsamp;
}

// ---------------------------------------------------------------------------
// Setup on the device: per voxel, TranMat_xt2sa from the Bunge angles (EulerToTransMatrix,
// kinematics.c:187-240, transposed as init.c:273-282 stores it), the initial stress
// sig = C_gr : (Udot dt) (+ the creep stress; InitFields, init.c:444-516) with C_gr = Q diag Q^T in
// the B basis, and per-block partial sums of C_gr (36) and sig (6) for the grid averages.
// ---------------------------------------------------------------------------
constexpr int SETUP_BLOCK = 256;
constexpr int SETUP_NPART = 42;

__global__ void __launch_bounds__(SETUP_BLOCK) k_setup(const DevConst *__restrict__ cst, Fields f, int nvox,
                                                       const double *__restrict__ eul, double deg, Vec9 e9,
                                                       Vec9 sc9, int creep, double *__restrict__ partials) {
    const int v = blockIdx.x * SETUP_BLOCK + threadIdx.x;
    const size_t N = (size_t)nvox;
    double acc[SETUP_NPART];
#pragma unroll
    for (int k = 0; k < SETUP_NPART; ++k) acc[k] = 0.0;
    if (v < nvox) {
        const int ph = f.phase[v];
        const bool xt = ph >= 1 && ph <= NPH && cst->ph[ph - 1].xtal;
        double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        double s6[6] = {0, 0, 0, 0, 0, 0};
        if (xt) {
            const double t1 = eul[3 * (size_t)v] * deg, th = eul[3 * (size_t)v + 1] * deg;
            const double om = eul[3 * (size_t)v + 2] * deg;
            const double ct1 = cos(t1), st1 = sin(t1), cph = cos(th), sph = sin(th), ct2 = cos(om), st2 = sin(om);
            // sample -> crystal rows, stored transposed (crystal -> sample)
            const double s2x[9] = {ct1 * ct2 - st1 * st2 * cph, st1 * ct2 + ct1 * st2 * cph, st2 * sph,
                                   -1.0 * ct1 * st2 - st1 * ct2 * cph, -1.0 * st1 * st2 + ct1 * ct2 * cph, ct2 * sph,
                                   st1 * sph, -1.0 * ct1 * sph, cph};
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) a[3 * i + j] = s2x[3 * j + i];
            double Q[5][5];
            q5_from(a, Q);
            const double *sinv = cst->ph[ph - 1].sinv;
            double c[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) c[k] = 1.0 / sinv[k];
            // C_gr deviatoric block; the hydrostatic entry is 1/sinv[5], the off-blocks zero
#pragma unroll
            for (int i = 0; i < 5; ++i)
#pragma unroll
                for (int j = 0; j < 5; ++j) {
                    double t = 0.0;
#pragma unroll
                    for (int k = 0; k < 5; ++k) t += Q[i][k] * c[k] * Q[j][k];
                    acc[6 * i + j] = t;
                }
            acc[35] = 1.0 / sinv[5];
            // B-basis strain e6 of sym(Udot dt), then sig = Q (c o (Q^T e6))
            double e6[6];
            e6[0] = SRX_R2 * (e9.v[4] - e9.v[0]);
            e6[1] = SRX_R6 * (2.0 * e9.v[8] - e9.v[0] - e9.v[4]);
            e6[2] = 2.0 * SRX_R2 * e9.v[5];
            e6[3] = 2.0 * SRX_R2 * e9.v[2];
            e6[4] = 2.0 * SRX_R2 * e9.v[1];
            e6[5] = SRX_R3 * (e9.v[0] + e9.v[4] + e9.v[8]);
            double t[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                double x = 0.0;
#pragma unroll
                for (int i = 0; i < 5; ++i) x += Q[i][k] * e6[i];
                t[k] = x * c[k];
            }
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                double x = 0.0;
#pragma unroll
                for (int k = 0; k < 5; ++k) x += Q[i][k] * t[k];
                s6[i] = x;
            }
            s6[5] = e6[5] / sinv[5];
            if (creep == 1) {
                const double sc[6] = {SRX_R2 * (sc9.v[4] - sc9.v[0]), SRX_R6 * (2.0 * sc9.v[8] - sc9.v[0] - sc9.v[4]),
                                      2.0 * SRX_R2 * sc9.v[5], 2.0 * SRX_R2 * sc9.v[2], 2.0 * SRX_R2 * sc9.v[1],
                                      SRX_R3 * (sc9.v[0] + sc9.v[4] + sc9.v[8])};
#pragma unroll
                for (int i = 0; i < 6; ++i) s6[i] += sc[i];
            }
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            f.tran[k * N + v] = a[k];
            f.tran0[k * N + v] = a[k];
        }
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            f.sig[i * N + v] = s6[i];
            acc[36 + i] = s6[i];
        }
    }
    // block sums in a fixed order: wavefronts, then the four wavefront sums in LDS
    __shared__ double red[SETUP_BLOCK / 64][SETUP_NPART];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < SETUP_NPART; ++k) {
        const double x = wave_sum(acc[k]);
        if (lane == 0) red[w][k] = x;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < SETUP_NPART; k += SETUP_BLOCK) {
        double x = 0.0;
        for (int j = 0; j < SETUP_BLOCK / 64; ++j) x += red[j][k];
        partials[(size_t)k * gridDim.x + blockIdx.x] = x;
    }
}

// ---------------------------------------------------------------------------
// host launch wrappers
// ---------------------------------------------------------------------------
static int min_radix(const FftPlan &p) {
    int m = 1 << 30;
    for (int i = 0; i < p.nst; ++i) m = p.rad[i] < m ? p.rad[i] : m;
    return p.nst ? m : 1;
}

bool make_plan(int n, FftPlan &p) {
    p.n = n;
    p.nst = 0;
    int m = n;
    if (n < 1 || n > 1024) return false;
    // powers of two: as many radix-8 stages as possible without leaving a radix-2 stage
    // (2^7 = 8*4*4, 2^8 = 8*8*4), then radix 3 and 5
    int k2 = 0;
    while (m % 2 == 0) { ++k2; m /= 2; }
    int n8 = k2 / 3, n4 = 0, n2 = 0;
    if (k2 % 3 == 1) { if (n8 > 0) { --n8; n4 = 2; } else { n2 = 1; } }
    else if (k2 % 3 == 2) n4 = 1;
    for (int i = 0; i < n8; ++i) p.rad[p.nst++] = 8;
    for (int i = 0; i < n4; ++i) p.rad[p.nst++] = 4;
    for (int i = 0; i < n2; ++i) p.rad[p.nst++] = 2;
    while (m % 3 == 0) { p.rad[p.nst++] = 3; m /= 3; }
    while (m % 5 == 0) { p.rad[p.nst++] = 5; m /= 5; }
    if (p.nst > MAX_RADIX_STAGES) return false;
    return m == 1;
}

// LDS budget of one FFT block (bytes)
constexpr int FFT_LDS_MAX = 48 * 1024;

// LDS of a y or z pass: t pencils at the padded pitch n + 1, then the twiddle table
static size_t padded_lds(int t, int n, size_t csz) { return ((size_t)t * (n + 1) + n) * csz; }

// largest tile count t with t*n/minrad <= maxb*FFT_BLOCK, t*comps*(n+1) complex values within
// FFT_LDS_MAX and t <= cap
static int tile_for(const FftPlan &p, int per_tile_comps, int maxb, int cap, int prec) {
    const int mr = min_radix(p);
    int t = (maxb * FFT_BLOCK * mr) / (p.n * per_tile_comps);
    const int csz = prec == 0 ? 16 : 8;
    const int tl = (FFT_LDS_MAX - p.n * csz) / (per_tile_comps * (p.n + 1) * csz);
    if (t > tl) t = tl;
    if (t > cap) t = cap;
    if (t < 1) t = 1;
    return t;
}

static void check_launch(const char *name) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) fprintf(stderr, "srx: launch of %s failed: %s\n", name, hipGetErrorString(e));
}

// Rows (pencils) per block of a z or y pass: `tile` if it is valid, else the largest that fits.
static int zy_tile(const FftPlan &p, int rows, int tile, int prec) {
    const int t = tile_for(p, 1, 2, rows < 16 ? rows : 16, prec);
    return tile >= 1 && tile <= t ? tile : t;
}

// Largest valid tile of pass `pass` (0 forward z, 1 forward y, 2 Green x, 3 inverse y, 4 inverse
// z); p is the plan of the pass's axis.
int fft_tile_max(int pass, const int n[3], const FftPlan &p, int prec) {
    switch (pass) {
    case 0: case 4: return zy_tile(p, n[1], 0, prec);
    case 1: case 3: return zy_tile(p, n[0], 0, prec);
    default: {   // Green x pass: 9 slots of t rows, up to 4 butterflies per thread
        int t = (4 * FFT_BLOCK * min_radix(p)) / (9 * p.n);
        const int csz = prec == 0 ? 16 : 8;
        const int tl = (FFT_LDS_MAX / csz - p.n) / (9 * p.n);
        t = t < tl ? t : tl;
        t = t < n[1] ? t : n[1];
        return t < 1 ? 1 : (t > 4 ? 4 : t);
    }
    }
}

void launch_fft_fwd_z(hipStream_t s, int prec, const double *sig, void *specB, const int n[3], const FftPlan &pz,
                      const void *twz, const IterCtl *ctl, int tile) {
    const int TY = zy_tile(pz, n[1], tile, prec);
    dim3 grid(n[0], (n[1] + TY - 1) / TY, 3);
    if (prec == 0) {
        size_t sm = padded_lds(TY, n[2], sizeof(cx<double>));
        SRX_LAUNCH(k_fwd_z<double>, grid, dim3(FFT_BLOCK), sm, s, sig, (cx<double> *)specB, 6, n[0], n[1],
                           n[2], pz, (const cx<double> *)twz, TY, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, ctl);
    } else {
        size_t sm = padded_lds(TY, n[2], sizeof(cx<float>));
        SRX_LAUNCH(k_fwd_z<float>, grid, dim3(FFT_BLOCK), sm, s, sig, (cx<float> *)specB, 6, n[0], n[1],
                           n[2], pz, (const cx<float> *)twz, TY, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, ctl);
    }
    check_launch("fwd_z");
}

void launch_fft_y(hipStream_t s, int prec, const void *in, void *out, int ncomp, int sign, int b_to_c, const int n[3],
                  const FftPlan &py, const void *twy, const IterCtl *ctl, int tile) {
    (void)sign;
    const int TX = zy_tile(py, n[0], tile, prec);
    const int H = n[2] / 2 + 1;
    dim3 grid((n[0] + TX - 1) / TX, H, ncomp);
    if (prec == 0) {
        size_t sm = padded_lds(TX, n[1], sizeof(cx<double>));
        SRX_LAUNCH(k_fft_y<double>, grid, dim3(FFT_BLOCK), sm, s, (const cx<double> *)in, (cx<double> *)out,
                           n[0], n[1], n[2], b_to_c, py, (const cx<double> *)twy, TX, ctl);
    } else {
        size_t sm = padded_lds(TX, n[1], sizeof(cx<float>));
        SRX_LAUNCH(k_fft_y<float>, grid, dim3(FFT_BLOCK), sm, s, (const cx<float> *)in, (cx<float> *)out,
                           n[0], n[1], n[2], b_to_c, py, (const cx<float> *)twy, TX, ctl);
    }
    check_launch("fft_y");
}

static void launch_green_generic(hipStream_t s, int prec, void *specC, const int n[3], const FftPlan &px,
                                 const void *twx, const DevConst *cst, const double *dc6, int mode, int ncomp,
                                 const IterCtl *ctl, int tile = 0) {
    const int slots = mode == 1 ? 9 : ncomp;
    int TY;
    bool wide;
    if (mode == 1 && tile >= 1 && tile <= fft_tile_max(2, n, px, prec)) {
        TY = tile;
        wide = (long)slots * TY * n[0] / min_radix(px) > 2 * FFT_BLOCK;
    } else {
        // two butterflies per thread when one row of `slots` pencils allows it, four otherwise
        wide = (long)slots * n[0] / min_radix(px) > 2 * FFT_BLOCK;
        TY = tile_for(px, slots, wide ? 4 : 2, n[1] < 2 ? n[1] : 2, prec);
    }
    const int H = n[2] / 2 + 1;
    dim3 grid((n[1] + TY - 1) / TY, H);
    const double z6[6] = {0, 0, 0, 0, 0, 0};
    const double *d = dc6 ? dc6 : z6;
    if (prec == 0) {
        size_t sm = (size_t)(slots * TY + 1) * n[0] * sizeof(cx<double>);
        if (wide)
            SRX_LAUNCH((k_green_x<double, 4>), grid, dim3(FFT_BLOCK), sm, s, (cx<double> *)specC, n[0], n[1], n[2],
                       px, (const cx<double> *)twx, TY, mode, ncomp, cst, d[0], d[1], d[2], d[3], d[4], d[5], ctl);
        else
            SRX_LAUNCH((k_green_x<double, 2>), grid, dim3(FFT_BLOCK), sm, s, (cx<double> *)specC, n[0], n[1], n[2],
                       px, (const cx<double> *)twx, TY, mode, ncomp, cst, d[0], d[1], d[2], d[3], d[4], d[5], ctl);
    } else {
        size_t sm = (size_t)(slots * TY + 1) * n[0] * sizeof(cx<float>);
        if (wide)
            SRX_LAUNCH((k_green_x<float, 4>), grid, dim3(FFT_BLOCK), sm, s, (cx<float> *)specC, n[0], n[1], n[2],
                       px, (const cx<float> *)twx, TY, mode, ncomp, cst, d[0], d[1], d[2], d[3], d[4], d[5], ctl);
        else
            SRX_LAUNCH((k_green_x<float, 2>), grid, dim3(FFT_BLOCK), sm, s, (cx<float> *)specC, n[0], n[1], n[2],
                       px, (const cx<float> *)twx, TY, mode, ncomp, cst, d[0], d[1], d[2], d[3], d[4], d[5], ctl);
    }
    check_launch("green_x");
}

// The wavefront-per-row x pass serves n0 = 64, 128, 256 (tile -1, and the default tile 0 there);
// other extents, and tiles >= 1, use the LDS kernel.
bool green_wave_ok(const int n[3]) { return n[0] == 64 || n[0] == 128 || n[0] == 256; }

template <typename T>
static void launch_green_wave(hipStream_t s, void *specC, const int n[3], const void *twx, const DevConst *cst,
                              const IterCtl *ctl) {
    const dim3 grid((n[1] + 3) / 4, n[2] / 2 + 1), block(256);
    cx<T> *sp = (cx<T> *)specC;
    const cx<T> *tw = (const cx<T> *)twx;
    if (n[0] == 64) SRX_LAUNCH((k_green_wave<T, 1>), grid, block, 0, s, sp, n[1], n[2], tw, cst, 0., 0., 0., 0., 0., 0., ctl);
    else if (n[0] == 128) SRX_LAUNCH((k_green_wave<T, 2>), grid, block, 0, s, sp, n[1], n[2], tw, cst, 0., 0., 0., 0., 0., 0., ctl);
    else SRX_LAUNCH((k_green_wave<T, 4>), grid, block, 0, s, sp, n[1], n[2], tw, cst, 0., 0., 0., 0., 0., 0., ctl);
}

void launch_green_x(hipStream_t s, int prec, void *specC, const int n[3], const FftPlan &px, const void *twx,
                    const DevConst *cst, const IterCtl *ctl, int tile) {
    if (green_wave_ok(n) && tile <= 0) {
        if (prec == 0) launch_green_wave<double>(s, specC, n, twx, cst, ctl);
        else launch_green_wave<float>(s, specC, n, twx, cst, ctl);
        check_launch("green_x");
        return;
    }
    launch_green_generic(s, prec, specC, n, px, twx, cst, nullptr, 1, 9, ctl, tile);
}

void launch_fft_inv_z_update(hipStream_t s, int prec, const void *specB, double *dg, const int n[3], const FftPlan &pz,
                             const void *twz, const IterCtl *ctl, double divisor, int tile) {
    const int TY = zy_tile(pz, n[1], tile, prec);
    dim3 grid(n[0], (n[1] + TY - 1) / TY, 5);
    const double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};   // dDisGradAvg comes from ctl
    if (prec == 0) {
        size_t sm = padded_lds(TY, n[2], sizeof(cx<double>));
        SRX_LAUNCH(k_inv_z<double>, grid, dim3(FFT_BLOCK), sm, s, (const cx<double> *)specB, dg, 9, n[0], n[1],
                           n[2], pz, (const cx<double> *)twz, TY, 1, divisor, a[0], a[1], a[2], a[3], a[4], a[5],
                           a[6], a[7], a[8], ctl);
    } else {
        size_t sm = padded_lds(TY, n[2], sizeof(cx<float>));
        SRX_LAUNCH(k_inv_z<float>, grid, dim3(FFT_BLOCK), sm, s, (const cx<float> *)specB, dg, 9, n[0], n[1],
                           n[2], pz, (const cx<float> *)twz, TY, 1, divisor, a[0], a[1], a[2], a[3], a[4], a[5], a[6],
                           a[7], a[8], ctl);
    }
    check_launch("inv_z");
}

void launch_mlp_record(hipStream_t s, const DevConst *cst, const Fields &f, int nvox, int stride, int after,
                       float *X, float *T, float *ND, int *count, int cap, const IterCtl *ctl) {
    const int nsamp = (nvox + stride - 1) / stride;
    const dim3 g((nsamp + 255) / 256), b(256);
    if (!after) {
        hipLaunchKernelGGL(k_mlp_rec_x, g, b, 0, s, cst, f, nvox, stride, nsamp, X, ND, count, cap, ctl);
    } else {
        hipLaunchKernelGGL(k_mlp_rec_t, g, b, 0, s, f, nvox, stride, nsamp, X, T, ND, count, cap, ctl);
        hipLaunchKernelGGL(k_mlp_rec_adv, dim3(1), dim3(64), 0, s, count, cap, nsamp, ctl);
    }
    check_launch("mlp_record");
}

int setup_block_count(int nvox) { return (nvox + SETUP_BLOCK - 1) / SETUP_BLOCK; }

void launch_setup(hipStream_t s, const DevConst *cst, const Fields &f, int nvox, const double *eul, double deg,
                  const double *e9, const double *sc9, int creep, double *partials, double *out) {
    Vec9 e, sc;
    for (int k = 0; k < 9; ++k) {
        e.v[k] = e9[k];
        sc.v[k] = sc9[k];
    }
    const int nb = setup_block_count(nvox);
    hipLaunchKernelGGL(k_setup, dim3(nb), dim3(SETUP_BLOCK), 0, s, cst, f, nvox, eul, deg, e, sc, creep, partials);
    check_launch("setup");
    launch_reduce(s, partials, nb, SETUP_NPART, out, nullptr);
}

int nr_block_count(int nvox) { return (nvox + NR_BLOCK - 1) / NR_BLOCK; }
int end_block_count(int nvox) { return (nvox + END_BLOCK - 1) / END_BLOCK; }

// Kernel variant of a Newton call: first call of an increment (F), rotated texture (R), Newton
// start (W: 0 reference, 1 extrapolated, 2 learned; the first call of an increment starts as 1),
// GND option (G, reference start only), crystalline second phase (P).
#define SRX_NR_VARIANTS(K, ARGS)                                                                           \
    do {                                                                                                 \
        const int ws_ = first_call && ws == 2 ? 1 : ws;                                                  \
        const int code_ = (first_call ? 16 : 0) | (rotated ? 8 : 0) | (gnd ? 4 : 0) | (ph2 ? 2 : 0);    \
        switch (code_ * 4 + ws_) {                                                                       \
        SRX_NR_CASE(K, ARGS, 1, 1, 0) SRX_NR_CASE(K, ARGS, 1, 0, 0) SRX_NR_CASE(K, ARGS, 0, 1, 0)        \
        SRX_NR_CASE(K, ARGS, 0, 0, 0) SRX_NR_CASE(K, ARGS, 1, 1, 1) SRX_NR_CASE(K, ARGS, 1, 0, 1)        \
        SRX_NR_CASE(K, ARGS, 0, 1, 1) SRX_NR_CASE(K, ARGS, 0, 0, 1) SRX_NR_CASE(K, ARGS, 0, 1, 2)        \
        SRX_NR_CASE(K, ARGS, 0, 0, 2)                                                                    \
        SRX_NR_GCASE(K, ARGS, 1, 1) SRX_NR_GCASE(K, ARGS, 1, 0) SRX_NR_GCASE(K, ARGS, 0, 1)              \
        SRX_NR_GCASE(K, ARGS, 0, 0)                                                                      \
        default: break;                                                                                  \
        }                                                                                                \
    } while (0)
#define SRX_NR_ONE(K, ARGS, F, R, W, G, P)                                                              \
    case ((F) * 16 + (R) * 8 + (G) * 4 + (P) * 2) * 4 + (W):                                             \
        SRX_APPLY(SRX_LAUNCH, (K<(F) != 0, (R) != 0, (W), (G) != 0, (P) != 0>), SRX_UNPACK ARGS);        \
        break;
#define SRX_NR_CASE(K, ARGS, F, R, W) SRX_NR_ONE(K, ARGS, F, R, W, 0, 0) SRX_NR_ONE(K, ARGS, F, R, W, 0, 1)
#define SRX_NR_GCASE(K, ARGS, F, R) SRX_NR_ONE(K, ARGS, F, R, 0, 1, 0) SRX_NR_ONE(K, ARGS, F, R, 0, 1, 1)
#define SRX_UNPACK(...) __VA_ARGS__
#define SRX_APPLY(M, ...) M(__VA_ARGS__)

void launch_nr(hipStream_t s, const DevConst *cst, const Fields &f, int nvox, int first_call, int rotated,
               double *partials, const IterCtl *ctl, int gnd, int ph2) {
    const dim3 g(nr_block_count(nvox)), b(NR_BLOCK);
    // the learned start needs the network; with only the previous change it is the extrapolated one
    const int ws = gnd ? 0 : (f.dsig && f.mlp) ? 2 : f.dsig ? 1 : 0;
    SRX_NR_VARIANTS(k_nr, (g, b, 0, s, cst, f, nvox, partials, ctl));
    check_launch("nr");
}

void launch_reduce(hipStream_t s, const double *partials, int nblocks, int nvals, double *out, const IterCtl *ctl) {
    const unsigned long long mask = nvals == NR_NPART ? (1ull << 15) : 0ull;
    SRX_LAUNCH(k_reduce, dim3(nvals), dim3(RED_BLOCK), 0, s, partials, nblocks, nvals, mask, out, ctl);
    check_launch("reduce");
}

void launch_reduce_nr(hipStream_t s, const double *partials, int nblocks, double *seg, const IterCtl *ctl) {
    SRX_LAUNCH(k_reduce_seg, dim3(NR_NPART, RED_SEGS), dim3(RED_SEG_BLOCK), 0, s, partials, nblocks, 1ull << 15, seg,
               ctl);
    check_launch("reduce");
}

void launch_nr_ens(hipStream_t s, const EnsMember *tab, int nmem, int nblocks, int first_call, int rotated, int ws,
                   int gnd, int ph2) {
    const dim3 g(nblocks), b(NR_BLOCK);
    if (gnd) ws = 0;
    SRX_NR_VARIANTS(k_nr_ens, (g, b, 0, s, tab, nmem));
    check_launch("nr_ens");
}

void launch_reduce_nr_ens(hipStream_t s, const EnsMember *tab, int nmem) {
    SRX_LAUNCH(k_reduce_seg_ens, dim3(NR_NPART, RED_SEGS, nmem), dim3(RED_SEG_BLOCK), 0, s, tab, 1ull << 15);
    check_launch("reduce_ens");
}

void launch_control_ens(hipStream_t s, const EnsMember *tab, int nmem, int *done) {
    SRX_LAUNCH(k_control_ens, dim3(nmem), dim3(64), 0, s, tab, done);
    check_launch("control_ens");
}

void launch_control(hipStream_t s, const DevConst *cst, const double *seg, IterCtl *ctl) {
    SRX_LAUNCH(k_control, dim3(1), dim3(64), 0, s, cst, seg, ctl);
    check_launch("control");
}

void launch_prologue(hipStream_t s, Fields f, int nvox, const double *a, int creep, const double *st,
                     const double *uo) {
    const int B = 256;
    Vec9 add, set, uold;
    for (int k = 0; k < 9; ++k) {
        add.v[k] = a[k];
        set.v[k] = st[k];
        uold.v[k] = uo[k];
    }
    SRX_LAUNCH(k_prologue, dim3((nvox + B - 1) / B), dim3(B), 0, s, f, nvox, add, creep, set, uold);
    check_launch("prologue");
}

void launch_end(hipStream_t s, const DevConst *cst, const Fields &f, int nvox, int lagged_possible, int,
                int do_orient, int, int istep, double *partials, int gnd) {
    if (gnd)
        SRX_LAUNCH(k_end<true>, dim3(end_block_count(nvox)), dim3(END_BLOCK), 0, s, cst, f, nvox, lagged_possible,
                   do_orient, istep, partials);
    else
        SRX_LAUNCH(k_end<false>, dim3(end_block_count(nvox)), dim3(END_BLOCK), 0, s, cst, f, nvox, lagged_possible,
                   do_orient, istep, partials);
    check_launch("end");
}

int gnd_block_count(int nvox) { return (nvox + GND_BLOCK - 1) / GND_BLOCK; }

void launch_gnd(hipStream_t s, const DevConst *cst, const DevConst &hc, const Fields &f, int nvox, int grad,
                double *partials) {
    SRX_LAUNCH(k_gnd, dim3(gnd_block_count(nvox)), dim3(GND_BLOCK), 0, s, cst, f.phase, f.gamdot, hc.gbc, hc.rho_g1,
               hc.rho_g2, hc.gnd_fp, nvox, grad, partials);
    check_launch("gnd");
}

// test path: forward FFT of ncomp real comps to layout C (x pass done), no Green
void launch_fft_real_forward_test(hipStream_t s, int prec, const double *in, int ncomp, void *specB, void *specC,
                                  const int n[3], const FftPlan *pl, void *const *tw) {
    const int TY = tile_for(pl[2], 1, 2, n[1] < 16 ? n[1] : 16, prec);
    dim3 grid(n[0], (n[1] + TY - 1) / TY, (ncomp + 1) / 2);
    if (prec == 0) {
        size_t sm = padded_lds(TY, n[2], sizeof(cx<double>));
        SRX_LAUNCH(k_fwd_z<double>, grid, dim3(FFT_BLOCK), sm, s, in, (cx<double> *)specB, ncomp, n[0], n[1],
                           n[2], pl[2], (const cx<double> *)tw[2], TY, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, nullptr);
    } else {
        size_t sm = padded_lds(TY, n[2], sizeof(cx<float>));
        SRX_LAUNCH(k_fwd_z<float>, grid, dim3(FFT_BLOCK), sm, s, in, (cx<float> *)specB, ncomp, n[0], n[1],
                           n[2], pl[2], (const cx<float> *)tw[2], TY, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, nullptr);
    }
    check_launch("fwd_z(test)");
    launch_fft_y(s, prec, specB, specC, ncomp, -1, 1, n, pl[1], tw[1], nullptr);
    launch_green_generic(s, prec, specC, n, pl[0], tw[0], nullptr, nullptr, 0, ncomp, nullptr);
}

// test path: inverse of the above for ncomp comps (layout C in, real out = IFFT * 1/divisor)
void launch_fft_real_inverse_test(hipStream_t s, int prec, void *specC, void *specB, int ncomp, double *out,
                                  const int n[3], const FftPlan *pl, void *const *tw, double divisor) {
    launch_green_generic(s, prec, specC, n, pl[0], tw[0], nullptr, nullptr, 2, ncomp, nullptr);
    launch_fft_y(s, prec, specC, specB, ncomp, +1, 0, n, pl[1], tw[1], nullptr);
    const int TY = tile_for(pl[2], 1, 2, n[1] < 16 ? n[1] : 16, prec);
    dim3 grid(n[0], (n[1] + TY - 1) / TY, (ncomp + 1) / 2);
    if (prec == 0) {
        size_t sm = padded_lds(TY, n[2], sizeof(cx<double>));
        SRX_LAUNCH(k_inv_z<double>, grid, dim3(FFT_BLOCK), sm, s, (const cx<double> *)specB, out, ncomp, n[0],
                           n[1], n[2], pl[2], (const cx<double> *)tw[2], TY, 0, divisor, 0., 0., 0., 0., 0., 0., 0.,
                           0., 0., nullptr);
    } else {
        size_t sm = padded_lds(TY, n[2], sizeof(cx<float>));
        SRX_LAUNCH(k_inv_z<float>, grid, dim3(FFT_BLOCK), sm, s, (const cx<float> *)specB, out, ncomp, n[0],
                           n[1], n[2], pl[2], (const cx<float> *)tw[2], TY, 0, divisor, 0., 0., 0., 0., 0., 0., 0., 0.,
                           0., nullptr);
    }
    check_launch("inv_z(test)");
}

// ---------------------------------------------------------------------------
// HBM ceiling: a plain streaming copy, 16 bytes per lane per access, U independent accesses in
// flight per lane, grid-stride over the buffer.  The measured rate is the ceiling the HBM-bound
// passes above are compared with.
// ---------------------------------------------------------------------------
typedef double dvec2 __attribute__((ext_vector_type(2)));
template <int U>
__global__ void __launch_bounds__(256) k_stream_copy(const dvec2 *__restrict__ in, dvec2 *__restrict__ out,
                                                      size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    for (; i + (U - 1) * stride < n; i += U * stride) {
        dvec2 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = __builtin_nontemporal_load(in + i + u * stride);
#pragma unroll
        for (int u = 0; u < U; ++u) __builtin_nontemporal_store(v[u], out + i + u * stride);
    }
    for (; i < n; i += stride) out[i] = in[i];
}

void launch_stream_copy(hipStream_t s, const void *in, void *out, size_t n16, int unroll, int blocks) {
    const dim3 grid(blocks), block(256);
    if (unroll == 4)
        hipLaunchKernelGGL(k_stream_copy<4>, grid, block, 0, s, (const dvec2 *)in, (dvec2 *)out, n16);
    else if (unroll == 2)
        hipLaunchKernelGGL(k_stream_copy<2>, grid, block, 0, s, (const dvec2 *)in, (dvec2 *)out, n16);
    else
        hipLaunchKernelGGL(k_stream_copy<1>, grid, block, 0, s, (const dvec2 *)in, (dvec2 *)out, n16);
    check_launch("stream_copy");
}

}  // namespace srx
