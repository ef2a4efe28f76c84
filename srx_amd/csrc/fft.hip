// fft.hip -- CDNA4 (gfx950) FFT and Green passes of the EVP-FFT increment loop
// (evolution.c:1157-1201 replaced): LDS-resident Stockham radix-2/3/4/5/8 transforms over
// batches of pencils, 256 threads per block, their global loads issued in chunks before the
// first wait.
//   fwd_z   real [c][x][y][z]      -> half spectrum B [c][kz][x][y]   (two real comps per complex FFT)
//   fft_y   B <-> C [c][kz][y][x]  (the y transform doubles as the pencil transpose)
//   green   C(6 stress comps) -> forward x -> Green operator -> inverse x -> C(9 comps), in place
//           (k_green_wave: one wavefront per row, cross-lane butterflies, for x extents 64/128/256;
//           k_green_x: LDS transforms otherwise)
//   inv_z   B(9 comps) -> c2r along z, fused with the FP64 DisGrad update (Eq. 15)
// Also here: the FFT-only test entry points and the streaming copy that measures the HBM ceiling.
#include "kern_common.hpp"

namespace srx {

// Events of the next launch (see SRX_LAUNCH).
thread_local hipEvent_t g_ev_start = nullptr, g_ev_stop = nullptr;
void set_launch_events(hipEvent_t a, hipEvent_t b) {
    g_ev_start = a;
    g_ev_stop = b;
}

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

// Position in LDS of element e of a pencil whose length is a multiple of 32: bits 0-4 of e XORed
// with b and 8b, b = bits 5-6 of e (a permutation within each aligned run of 32 elements).  A
// radix-8 first stage writes element 8j + r from butterfly j and a radix-4 second stage element
// 32 jl + k + 8 r (j = 8 jl + k); unswizzled, four lanes of a pencil share each LDS bank in both,
// swizzled the lanes of a pencil write distinct banks.  The passes' own LDS accesses use the same
// positions.
__device__ __forceinline__ int lds_swz(int e) {
    const int b = (e >> 5) & 3;
    return e ^ b ^ (b << 3);
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
template <typename T, int R, int MAXB, int N, int L, bool SW>
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
            const cx<T> *src = buf + p * pitch + (SW ? 0 : j);
#pragma unroll
            for (int r = 0; r < R; ++r) v[m][r] = src[SW ? lds_swz(j + r * nb) : r * nb];
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
            const int e0 = (j / L) * L * R + k;
            cx<T> *dst = buf + p * pitch + (SW ? 0 : e0);
#pragma unroll
            for (int r = 0; r < R; ++r) dst[SW ? lds_swz(e0 + r * L) : r * L] = v[m][r];
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

template <typename T, int MAXB, int N, int ST, int L, bool SW>
__device__ __forceinline__ void lds_fft_pow2(cx<T> *buf, int npen, const cx<T> *__restrict__ tw, int sign,
                                             int pitch) {
    constexpr int R = pow2_radix(N, ST);
    if constexpr (R != 0) {
        stockham_stage_ct<T, R, MAXB, N, L, SW>(buf, npen, tw, sign, pitch);
        lds_fft_pow2<T, MAXB, N, ST + 1, L * R, SW>(buf, npen, tw, sign, pitch);
    }
}

// whether plan pl is make_plan's schedule for the power-of-two length N (the one the compile-time
// stages run); a tuned or caller-set plan with other radices takes the runtime stages
template <int N>
__device__ __forceinline__ bool pow2_default(const FftPlan &pl) {
    constexpr int ns = pow2_radix(N, 0) == 0 ? 0 : pow2_radix(N, 1) == 0 ? 1 : pow2_radix(N, 2) == 0 ? 2
                     : pow2_radix(N, 3) == 0 ? 3 : 4;
    bool same = pl.nst == ns;
#pragma unroll
    for (int st = 0; st < ns; ++st) same = same && pl.rad[st] == pow2_radix(N, st);
    return same;
}

// Full 1-D transform of npen LDS-resident pencils.  Caller guarantees
// npen*N/min_radix <= MAXB*FFT_BLOCK and a __syncthreads() before the call.  SWZ: the compile-time
// stages store the pencils swizzled (lds_swz), and the caller's own accesses must do the same.
// POW2: lengths 32..256 with make_plan's radices use the compile-time stages (measured faster in
// the x, y and z passes; in the z passes they cost an occupancy step, which the swizzled, conflict
// free stages more than repay).
template <typename T, int MAXB, bool POW2 = true, bool SWZ = true>
__device__ void lds_fft(cx<T> *buf, int npen, const FftPlan &pl, const cx<T> *__restrict__ tw, int sign,
                        int pitch) {
    if constexpr (POW2) {
        switch (pl.n) {
        case 32: if (pow2_default<32>(pl)) { lds_fft_pow2<T, MAXB, 32, 0, 1, SWZ>(buf, npen, tw, sign, pitch); return; } break;
        case 64: if (pow2_default<64>(pl)) { lds_fft_pow2<T, MAXB, 64, 0, 1, SWZ>(buf, npen, tw, sign, pitch); return; } break;
        case 128: if (pow2_default<128>(pl)) { lds_fft_pow2<T, MAXB, 128, 0, 1, SWZ>(buf, npen, tw, sign, pitch); return; } break;
        case 256: if (pow2_default<256>(pl)) { lds_fft_pow2<T, MAXB, 256, 0, 1, SWZ>(buf, npen, tw, sign, pitch); return; } break;
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
template <typename T, bool SW>
__global__ void __launch_bounds__(FFT_BLOCK) SRX_ZPASS_ATTR k_fwd_z(const double *__restrict__ in, cx<T> *__restrict__ specB,
                                                     int ncomp, int n0, int n1, int n2, FftPlan pz,
                                                     const cx<T> *__restrict__ tw, int TY, double m0, double m1,
                                                     double m2, double m3, double m4, double m5,
                                                     const IterCtl *__restrict__ ctl) {
    if (ctl && ctl->done) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
    // SW: the plan runs the compile-time stages, which store the pencils swizzled (lds_swz) at pitch
    // n2 + 2, and the accesses here do the same; else the runtime stages, plain, at pitch n2 + 1
    constexpr bool sw = SW;
    const int P = SW ? n2 + 2 : n2 + 1;   // padded pencil pitch: rows fall on different LDS banks
    cx<T> *stw = buf + TY * P;            // twiddle table staged in LDS
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
                buf[yy * P + (sw ? lds_swz(z) : z)] = cx<T>{T(v.x - ma), T(two ? v.y - mb : 0.0)};
            });
        __syncthreads();
        if constexpr (SW) lds_fft<T, 2, true, true>(buf, ty, pz, stw, -1, P);
        else lds_fft<T, 2, false, false>(buf, ty, pz, stw, -1, P);
        for (int i = threadIdx.x; i < H * ty; i += FFT_BLOCK) {
            const int kz = dty.div(i), yy = i - kz * ty;
            const int e2 = kz == 0 ? 0 : n2 - kz;
            const cx<T> z1 = buf[yy * P + (sw ? lds_swz(kz) : kz)];
            const cx<T> z2 = buf[yy * P + (sw ? lds_swz(e2) : e2)];
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
template <typename T, bool SW>
__global__ void __launch_bounds__(FFT_BLOCK) k_fft_y(const cx<T> *__restrict__ in, cx<T> *__restrict__ out,
                                                     int n0, int n1, int n2, int b_to_c, FftPlan py,
                                                     const cx<T> *__restrict__ tw, int TX,
                                                     const IterCtl *__restrict__ ctl) {
    if (ctl && ctl->done) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
    // SW: compile-time stages, swizzled pencils at pitch n1 + 2; else runtime stages, plain, n1 + 1
    constexpr bool sw = SW;
    const int P = SW ? n1 + 2 : n1 + 1;   // padded pencil pitch
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
                                buf[xx * P + (sw ? lds_swz(y) : y)] = v;
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
                buf[xx * P + (sw ? lds_swz(y) : y)] = v;
            });
    }
    __syncthreads();
    if constexpr (SW) lds_fft<T, 2, true, true>(buf, tx, py, stw, b_to_c ? -1 : +1, P);
    else lds_fft<T, 2, false, false>(buf, tx, py, stw, b_to_c ? -1 : +1, P);
    if (b_to_c) {
        for (int i = threadIdx.x; i < tx * n1; i += FFT_BLOCK) {
            const int y = dtx.div(i), xx = i - y * tx;
            out[base + (size_t)y * n0 + x0 + xx] = buf[xx * P + (sw ? lds_swz(y) : y)];
        }
    } else {
        for (int i = threadIdx.x; i < tx * n1; i += FFT_BLOCK) {
            const int xx = dn1.div(i), y = i - xx * n1;
            out[base + (size_t)(x0 + xx) * n1 + y] = buf[xx * P + (sw ? lds_swz(y) : y)];
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

// The same operator in factored form along an x row, for rows with ky and kz both off their
// Nyquist indices and not both 0.  With t the unnormalised wave vector (unit_k), n = t / |t| and
// du_ij = -w_i n_j, the y and z columns are t_y Y_i and t_z Y_i with Y_i = -w_i / |t|, and t_y, t_z
// are constant along the row, so the row's inverse x transforms of those six components are
// t_y IFFT(Y_i) and t_z IFFT(Y_i).  X_i is the x column.  At the x Nyquist index the Hermitian
// average of green_point keeps this form: Y_i = -(w_i + w'_i) / (2 |t|), X_i = -(w_i - w'_i) n_x / 2.
template <typename T>
__device__ __forceinline__ void green_point_xy(const DevConst *__restrict__ cst, int kx, int ky, int kz, int n0,
                                               int n1, int n2, const cx<T> q[6], cx<T> X[3], cx<T> Y[3]) {
    using A = T;
    const A *G6;
    if constexpr (sizeof(T) == 4) G6 = cst->G6f;
    else G6 = cst->G6;
    A tzv, tyv, txv;
    unit_k(kz, n2, tzv);
    unit_k(ky, n1, tyv);
    unit_k(kx, n0, txv);
    const bool nyx = (n0 % 2 == 0) && (kx == n0 / 2);
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
    if (!nyx) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            X[a] = cx<T>{-wr[a] * nv[0], -wi[a] * nv[0]};
            Y[a] = cx<T>{-wr[a] * rg, -wi[a] * rg};
        }
    } else {
        const A np[3] = {-nv[0], nv[1], nv[2]};
        A wr2[3], wi2[3];
        green_w6(G6, np, sr, si, wr2, wi2);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            X[a] = cx<T>{A(0.5) * (-wr[a] + wr2[a]) * nv[0], A(0.5) * (-wi[a] + wi2[a]) * nv[0]};
            Y[a] = cx<T>{A(-0.5) * (wr[a] + wr2[a]) * rg, A(-0.5) * (wi[a] + wi2[a]) * rg};
        }
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
    if (mode != 2) lds_fft<T, MAXB, true, false>(buf, cin * ty, px, stw, -1, n0);
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
    if (mode != 0) lds_fft<T, MAXB, true, false>(buf, cout * ty, px, stw, +1, n0);
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
    const int kb = NPL * brev6(lane);
    // the row's Green operator in factored form (6 inverse transforms) unless ky or kz is at its
    // Nyquist index or the row holds k = 0 (wavefront-uniform)
    const bool factored = !((n1 % 2 == 0 && ky == n1 / 2) || (n2 % 2 == 0 && kz == n2 / 2) || (ky == 0 && kz == 0));
    if (factored) {
        cx<T> o[6][NPL];   // X_0..2, Y_0..2
#pragma unroll
        for (int r = 0; r < NPL; ++r) {
            cx<T> qi[6], xi[3], yi[3];
#pragma unroll
            for (int c = 0; c < 6; ++c) qi[c] = q[c][r];
            green_point_xy(cst, r + kb, ky, kz, N, n1, n2, qi, xi, yi);
#pragma unroll
            for (int a = 0; a < 3; ++a) { o[a][r] = xi[a]; o[3 + a][r] = yi[a]; }
        }
        wave_fft_inv<T, NPL, 6>(o, wl, ws, lane);
        T ty, tz;
        unit_k(ky, n1, ty);
        unit_k(kz, n2, tz);
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int r = 0; r < NPL; ++r) {
                const cx<T> y = o[3 + a][r];
                spec[(3 * a) * cstride + row + 64 * r] = o[a][r];
                spec[(3 * a + 1) * cstride + row + 64 * r] = cx<T>{ty * y.x, ty * y.y};
                spec[(3 * a + 2) * cstride + row + 64 * r] = cx<T>{tz * y.x, tz * y.y};
            }
        return;
    }
    cx<T> o[9][NPL];
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
template <typename T, bool SW>
__global__ void __launch_bounds__(FFT_BLOCK) SRX_ZPASS_ATTR k_inv_z(const cx<T> *__restrict__ specB, double *__restrict__ dg,
                                                     int ncomp, int n0, int n1, int n2, FftPlan pz,
                                                     const cx<T> *__restrict__ tw, int TY, int accumulate,
                                                     double divisor, double a0, double a1, double a2, double a3,
                                                     double a4, double a5, double a6, double a7, double a8,
                                                     const IterCtl *__restrict__ ctl) {
    if (ctl && ctl->done) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
    // SW: compile-time stages, swizzled pencils at pitch n2 + 2; else runtime stages, plain, n2 + 1
    constexpr bool sw = SW;
    const int P = SW ? n2 + 2 : n2 + 1;   // padded pencil pitch
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
                buf[yy * P + (sw ? lds_swz(kz) : kz)] = cx<T>{A.x - B.y, A.y + B.x};
                if (kz >= 1 && kz <= n2 - H)
                    buf[yy * P + (sw ? lds_swz(n2 - kz) : n2 - kz)] = cx<T>{A.x + B.y, B.x - A.y};
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
        if constexpr (SW) lds_fft<T, 2, true, true>(buf, ty, pz, stw, +1, P);
        else lds_fft<T, 2, false, false>(buf, ty, pz, stw, +1, P);
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
                    const cx<T> v = buf[yy * P + (sw ? lds_swz(z) : z)];
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

// Radix plans of length n the tuner may time: make_plan's first, then every split of the power of
// two into radix-8 and radix-4 stages with at most one radix-2 stage (more radix-2 stages only add
// passes over LDS), each in make_plan's order (8s, 4s, the 2, then 3s and 5s) and in the reverse
// order.  Returns the number written to out (at most cap); 0 when make_plan rejects n.
int plan_candidates(int n, FftPlan *out, int cap) {
    FftPlan d;
    if (cap < 1 || !make_plan(n, d)) return 0;
    int cnt = 0;
    auto add = [&](const FftPlan &p) {
        for (int i = 0; i < cnt; ++i) {
            bool same = out[i].nst == p.nst;
            for (int s = 0; same && s < p.nst; ++s) same = out[i].rad[s] == p.rad[s];
            if (same) return;
        }
        if (cnt < cap) out[cnt++] = p;
    };
    add(d);
    int m = n, k2 = 0, k3 = 0, k5 = 0;
    while (m % 2 == 0) { ++k2; m /= 2; }
    while (m % 3 == 0) { ++k3; m /= 3; }
    while (m % 5 == 0) { ++k5; m /= 5; }
    for (int n8 = k2 / 3; n8 >= 0; --n8)
        for (int n4 = (k2 - 3 * n8) / 2; n4 >= 0; --n4) {
            const int n2 = k2 - 3 * n8 - 2 * n4;
            if (n2 > 1 || n8 + n4 + n2 + k3 + k5 > MAX_RADIX_STAGES) continue;
            FftPlan p;
            p.n = n;
            p.nst = 0;
            for (int i = 0; i < n8; ++i) p.rad[p.nst++] = 8;
            for (int i = 0; i < n4; ++i) p.rad[p.nst++] = 4;
            for (int i = 0; i < n2; ++i) p.rad[p.nst++] = 2;
            for (int i = 0; i < k3; ++i) p.rad[p.nst++] = 3;
            for (int i = 0; i < k5; ++i) p.rad[p.nst++] = 5;
            add(p);
            FftPlan r = p;
            for (int s = 0; s < p.nst; ++s) r.rad[s] = p.rad[p.nst - 1 - s];
            add(r);
        }
    return cnt;
}

// LDS budget of one FFT block (bytes)
constexpr int FFT_LDS_MAX = 48 * 1024;

// LDS of a y or z pass: t pencils at the padded pitch (n + 2 swizzled, else n + 1), then the
// twiddle table
static size_t padded_lds(int t, int n, size_t csz, bool sw) { return ((size_t)t * (sw ? n + 2 : n + 1) + n) * csz; }

// Whether a z or y pass runs plan p with the compile-time stages and swizzled pencils (the kernels'
// SW = true instantiation): lengths 32..256 with make_plan's radices.  Other plans run the runtime
// stages on plain pencils (SW = false), a separate instantiation so that neither pays for the
// other's registers.
static bool plan_swizzled(const FftPlan &p) {
    if (p.n != 32 && p.n != 64 && p.n != 128 && p.n != 256) return false;
    FftPlan d;
    make_plan(p.n, d);
    if (d.nst != p.nst) return false;
    for (int s = 0; s < p.nst; ++s)
        if (d.rad[s] != p.rad[s]) return false;
    return true;
}

// largest tile count t with t*n/minrad <= maxb*FFT_BLOCK, t*comps*lds_pitch(n) complex values within
// FFT_LDS_MAX and t <= cap
static int tile_for(const FftPlan &p, int per_tile_comps, int maxb, int cap, int prec) {
    const int mr = min_radix(p);
    int t = (maxb * FFT_BLOCK * mr) / (p.n * per_tile_comps);
    const int csz = prec == 0 ? 16 : 8;
    const int tl = (FFT_LDS_MAX - p.n * csz) / (per_tile_comps * lds_pitch(p.n) * csz);
    if (t > tl) t = tl;
    if (t > cap) t = cap;
    if (t < 1) t = 1;
    return t;
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
    const bool sw = plan_swizzled(pz);
    dim3 grid(n[0], (n[1] + TY - 1) / TY, 3);
    if (prec == 0) {
        size_t sm = padded_lds(TY, n[2], sizeof(cx<double>), sw);
        if (sw)
            SRX_LAUNCH((k_fwd_z<double, true>), grid, dim3(FFT_BLOCK), sm, s, sig, (cx<double> *)specB, 6, n[0], n[1],
                           n[2], pz, (const cx<double> *)twz, TY, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, ctl);
        else
            SRX_LAUNCH((k_fwd_z<double, false>), grid, dim3(FFT_BLOCK), sm, s, sig, (cx<double> *)specB, 6, n[0], n[1],
                           n[2], pz, (const cx<double> *)twz, TY, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, ctl);
    } else {
        size_t sm = padded_lds(TY, n[2], sizeof(cx<float>), sw);
        if (sw)
            SRX_LAUNCH((k_fwd_z<float, true>), grid, dim3(FFT_BLOCK), sm, s, sig, (cx<float> *)specB, 6, n[0], n[1],
                           n[2], pz, (const cx<float> *)twz, TY, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, ctl);
        else
            SRX_LAUNCH((k_fwd_z<float, false>), grid, dim3(FFT_BLOCK), sm, s, sig, (cx<float> *)specB, 6, n[0], n[1],
                           n[2], pz, (const cx<float> *)twz, TY, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, ctl);
    }
    check_launch("fwd_z");
}

void launch_fft_y(hipStream_t s, int prec, const void *in, void *out, int ncomp, int sign, int b_to_c, const int n[3],
                  const FftPlan &py, const void *twy, const IterCtl *ctl, int tile) {
    (void)sign;
    const int TX = zy_tile(py, n[0], tile, prec);
    const bool sw = plan_swizzled(py);
    const int H = n[2] / 2 + 1;
    dim3 grid((n[0] + TX - 1) / TX, H, ncomp);
    if (prec == 0) {
        size_t sm = padded_lds(TX, n[1], sizeof(cx<double>), sw);
        if (sw)
            SRX_LAUNCH((k_fft_y<double, true>), grid, dim3(FFT_BLOCK), sm, s, (const cx<double> *)in, (cx<double> *)out,
                           n[0], n[1], n[2], b_to_c, py, (const cx<double> *)twy, TX, ctl);
        else
            SRX_LAUNCH((k_fft_y<double, false>), grid, dim3(FFT_BLOCK), sm, s, (const cx<double> *)in, (cx<double> *)out,
                           n[0], n[1], n[2], b_to_c, py, (const cx<double> *)twy, TX, ctl);
    } else {
        size_t sm = padded_lds(TX, n[1], sizeof(cx<float>), sw);
        if (sw)
            SRX_LAUNCH((k_fft_y<float, true>), grid, dim3(FFT_BLOCK), sm, s, (const cx<float> *)in, (cx<float> *)out,
                           n[0], n[1], n[2], b_to_c, py, (const cx<float> *)twy, TX, ctl);
        else
            SRX_LAUNCH((k_fft_y<float, false>), grid, dim3(FFT_BLOCK), sm, s, (const cx<float> *)in, (cx<float> *)out,
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
    const bool sw = plan_swizzled(pz);
    dim3 grid(n[0], (n[1] + TY - 1) / TY, 5);
    const double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};   // dDisGradAvg comes from ctl
    if (prec == 0) {
        size_t sm = padded_lds(TY, n[2], sizeof(cx<double>), sw);
        if (sw)
            SRX_LAUNCH((k_inv_z<double, true>), grid, dim3(FFT_BLOCK), sm, s, (const cx<double> *)specB, dg, 9, n[0], n[1],
                           n[2], pz, (const cx<double> *)twz, TY, 1, divisor, a[0], a[1], a[2], a[3], a[4], a[5],
                           a[6], a[7], a[8], ctl);
        else
            SRX_LAUNCH((k_inv_z<double, false>), grid, dim3(FFT_BLOCK), sm, s, (const cx<double> *)specB, dg, 9, n[0], n[1],
                           n[2], pz, (const cx<double> *)twz, TY, 1, divisor, a[0], a[1], a[2], a[3], a[4], a[5],
                           a[6], a[7], a[8], ctl);
    } else {
        size_t sm = padded_lds(TY, n[2], sizeof(cx<float>), sw);
        if (sw)
            SRX_LAUNCH((k_inv_z<float, true>), grid, dim3(FFT_BLOCK), sm, s, (const cx<float> *)specB, dg, 9, n[0], n[1],
                           n[2], pz, (const cx<float> *)twz, TY, 1, divisor, a[0], a[1], a[2], a[3], a[4], a[5], a[6],
                           a[7], a[8], ctl);
        else
            SRX_LAUNCH((k_inv_z<float, false>), grid, dim3(FFT_BLOCK), sm, s, (const cx<float> *)specB, dg, 9, n[0], n[1],
                           n[2], pz, (const cx<float> *)twz, TY, 1, divisor, a[0], a[1], a[2], a[3], a[4], a[5], a[6],
                           a[7], a[8], ctl);
    }
    check_launch("inv_z");
}

// test path: forward FFT of ncomp real comps to layout C (x pass done), no Green
void launch_fft_real_forward_test(hipStream_t s, int prec, const double *in, int ncomp, void *specB, void *specC,
                                  const int n[3], const FftPlan *pl, void *const *tw) {
    const int TY = tile_for(pl[2], 1, 2, n[1] < 16 ? n[1] : 16, prec);
    const bool sw = plan_swizzled(pl[2]);
    dim3 grid(n[0], (n[1] + TY - 1) / TY, (ncomp + 1) / 2);
    if (prec == 0) {
        size_t sm = padded_lds(TY, n[2], sizeof(cx<double>), sw);
        if (sw)
            SRX_LAUNCH((k_fwd_z<double, true>), grid, dim3(FFT_BLOCK), sm, s, in, (cx<double> *)specB, ncomp, n[0], n[1],
                           n[2], pl[2], (const cx<double> *)tw[2], TY, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, nullptr);
        else
            SRX_LAUNCH((k_fwd_z<double, false>), grid, dim3(FFT_BLOCK), sm, s, in, (cx<double> *)specB, ncomp, n[0], n[1],
                           n[2], pl[2], (const cx<double> *)tw[2], TY, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, nullptr);
    } else {
        size_t sm = padded_lds(TY, n[2], sizeof(cx<float>), sw);
        if (sw)
            SRX_LAUNCH((k_fwd_z<float, true>), grid, dim3(FFT_BLOCK), sm, s, in, (cx<float> *)specB, ncomp, n[0], n[1],
                           n[2], pl[2], (const cx<float> *)tw[2], TY, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, nullptr);
        else
            SRX_LAUNCH((k_fwd_z<float, false>), grid, dim3(FFT_BLOCK), sm, s, in, (cx<float> *)specB, ncomp, n[0], n[1],
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
    const bool sw = plan_swizzled(pl[2]);
    dim3 grid(n[0], (n[1] + TY - 1) / TY, (ncomp + 1) / 2);
    if (prec == 0) {
        size_t sm = padded_lds(TY, n[2], sizeof(cx<double>), sw);
        if (sw)
            SRX_LAUNCH((k_inv_z<double, true>), grid, dim3(FFT_BLOCK), sm, s, (const cx<double> *)specB, out, ncomp, n[0],
                           n[1], n[2], pl[2], (const cx<double> *)tw[2], TY, 0, divisor, 0., 0., 0., 0., 0., 0., 0.,
                           0., 0., nullptr);
        else
            SRX_LAUNCH((k_inv_z<double, false>), grid, dim3(FFT_BLOCK), sm, s, (const cx<double> *)specB, out, ncomp, n[0],
                           n[1], n[2], pl[2], (const cx<double> *)tw[2], TY, 0, divisor, 0., 0., 0., 0., 0., 0., 0.,
                           0., 0., nullptr);
    } else {
        size_t sm = padded_lds(TY, n[2], sizeof(cx<float>), sw);
        if (sw)
            SRX_LAUNCH((k_inv_z<float, true>), grid, dim3(FFT_BLOCK), sm, s, (const cx<float> *)specB, out, ncomp, n[0],
                           n[1], n[2], pl[2], (const cx<float> *)tw[2], TY, 0, divisor, 0., 0., 0., 0., 0., 0., 0., 0.,
                           0., nullptr);
        else
            SRX_LAUNCH((k_inv_z<float, false>), grid, dim3(FFT_BLOCK), sm, s, (const cx<float> *)specB, out, ncomp, n[0],
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
