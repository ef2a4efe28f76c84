// crystal_dev.hpp -- per-voxel crystal helpers of the Newton, end-of-increment and setup
// kernels: rotations in the B basis and the dislocation-density slip law.
#pragma once
#include "kern_common.hpp"

namespace srx {

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
// (a literal addend is first copied into a vector register at every use).  Not `static`: with
// internal linkage the compiler folds the values back into literals.  Each kernel file that
// includes this header holds its own copy.
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

}  // namespace srx
