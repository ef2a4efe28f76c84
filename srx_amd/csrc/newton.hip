// newton.hip -- CDNA4 (gfx950) Newton kernels of the EVP-FFT increment loop: the constitutive
// solve of UpdateStress (evolution.c:157-427), one voxel per lane in FP64, its fields read and
// written as raw buffer accesses (one 32-bit lane offset from a scalar base); the fixed-order
// reductions of its block partial sums (k_reduce, k_reduce_seg), so every sum is bitwise
// repeatable from run to run; the device-side outer-iteration control (k_control); the
// lockstep ensemble variants; and the recorder of training samples for the learned start.
#include "crystal_dev.hpp"
#include "mlp_dev.hpp"

namespace srx {

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
// A field's base pointer as an opaque value at this point: the epilogue's buffer descriptors are
// built after the Newton loop instead of being the prologue's identical ones, which the compiler
// would keep in scalar registers (spilled to vector lanes) across the loop.
template <typename T>
__device__ __forceinline__ T *fresh(T *p) {
    asm volatile("" : "+s"(p));
    return p;
}
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
        if (f.acrf) {
#pragma unroll
            for (int k = 0; k < 15; ++k) in.A[k] = (double)ldcf(f.acrf + (size_t)k * N, N, o8 / 2);
        } else {
#pragma unroll
            for (int k = 0; k < 15; ++k) in.A[k] = ldc(f.acr + (size_t)k * N, N, o8);
        }
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
        if (f.acrf) {
#pragma unroll
            for (int k = 0; k < 15; ++k) {
                const float a32 = (float)pro.A[k];
                pro.A[k] = (double)a32;
                stcf(f.acrf + (size_t)(k) * N, N, 4u * (unsigned)v, a32);
            }
        } else {
#pragma unroll
            for (int k = 0; k < 15; ++k) stc(f.acr + (size_t)(k) * N, N, o8, pro.A[k]);
        }
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
    // mixed precision: the constants are stored in FP32 (f.nrcf), and this call iterates with the
    // rounded values too, so that every call of the increment solves with the same constants
    const size_t fpitch = nrcf_pitch(N);
    if (FIRST) {
        const int nsys = P.nsys;
        const bool f32 = f.nrcf != nullptr;
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
                // an inactive system: e = |rss| ci - ct stays negative (also as FP32's largest value)
                ct = 1e300; ci = 0.0; cg = 0.0;
                tp[s * NR_BLOCK] = 1e300;
            }
            if (f32) {
                ct = (double)(float)fmin(ct, 3.0e38);
                ci = (double)(float)ci;
                cg = (double)(float)cg;
            }
            inv[s * NR_BLOCK] = ci;
            g0[s * NR_BLOCK] = cg;
            if (f32) {
                stcf(f.nrcf + (size_t)(NRC_TP + s) * fpitch, fpitch, 4u * (unsigned)v, (float)ct);
                stcf(f.nrcf + (size_t)(NRC_INV + s) * fpitch, fpitch, 4u * (unsigned)v, (float)ci);
                stcf(f.nrcf + (size_t)(NRC_G0 + s) * fpitch, fpitch, 4u * (unsigned)v, (float)cg);
            } else {
                stc(f.nrc + (size_t)(NRC_TP + s) * nrc_pitch(N), nrc_pitch(N), o8, ct);
                stc(f.nrc + (size_t)(NRC_INV + s) * nrc_pitch(N), nrc_pitch(N), o8, ci);
                stc(f.nrc + (size_t)(NRC_G0 + s) * nrc_pitch(N), nrc_pitch(N), o8, cg);
            }
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
            if (f.nrcf) {
                const float *nc = fresh(f.nrcf);
#pragma unroll
                for (int s = 0; s < NSYS; ++s)
                    tp[s * NR_BLOCK] = (double)ldcf(nc + (size_t)(NRC_TP + s) * fpitch, fpitch, 4u * (unsigned)v);
            } else {
                const double *nc = fresh(f.nrc);
#pragma unroll
                for (int s = 0; s < NSYS; ++s)
                    tp[s * NR_BLOCK] = ldc(nc + (size_t)(NRC_TP + s) * nrc_pitch(N), nrc_pitch(N), o8);
            }
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
        const double *tr = fresh(f.tran);
#pragma unroll
        for (int k = 0; k < 9; ++k) a[k] = ldc(tr + (size_t)(k) * N, N, o8);
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
    for (int k = 0; k < 5; ++k) stc(fresh(f.sig_pre) + (size_t)(k) * N, N, o8, pre[k]);
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
// The same for constants stored in FP32 at pitch P (a multiple of 4): one 16-byte load per lane moves
// four components (lanes 16 q .. 16 q + 15 component c + q, four voxels each) into the upper half of
// the NC FP64 slots: FP32 slot k at float offset (NC + k) 64 from their start (NC a multiple of 4);
// widen_slots then converts them in place.
template <int NC>
__device__ __forceinline__ void copy_field_f32(const float *base, size_t P, double *lds, int blk) {
    static_assert(NC % 4 == 0, "four components per load");
#if __HIP_DEVICE_COMPILE__
    const int lane = threadIdx.x;
    const int v0 = blk * NR_BLOCK;
    const unsigned pb = 4u * (unsigned)min((size_t)(v0 + 4 * (lane & 15)), P - 4);
    const unsigned q = 4u * (unsigned)P * (unsigned)(lane >> 4);
    float *half = reinterpret_cast<float *>(lds + NC / 2 * NR_BLOCK);
#pragma unroll
    for (int c = 0; c < NC; c += 4)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(
            __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(base + (size_t)c * P), (short)0,
                                              rsrc_bytes(4 * P * 4u), 0x00020000),
            (lds_void *)(half + c * NR_BLOCK), 16, pb + q, 0, 0, 0);
#endif
}
// After copy_field_f32 and copy_wait: FP32 slot k -> FP64 slot k of the lane's voxel.  Every FP32
// value of the lane is read before the first FP64 slot is written (the FP64 slots overlay the FP32
// ones; a block is one wavefront, whose LDS accesses complete in program order).
typedef float __attribute__((may_alias)) float_alias;
template <int NC>
__device__ __forceinline__ void widen_slots(double *lds) {
    static_assert(NR_BLOCK == 64, "one wavefront per block");
    const float_alias *half = reinterpret_cast<const float_alias *>(lds + NC / 2 * NR_BLOCK) + threadIdx.x;
    float x[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) x[c] = half[c * NR_BLOCK];
    asm volatile("" ::: "memory");
#pragma unroll
    for (int c = 0; c < NC; ++c) lds[c * NR_BLOCK + threadIdx.x] = (double)x[c];
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
        const bool f32 = f.nrcf != nullptr;
        if (f32)
            copy_field_f32<NRC_COMPS>(f.nrcf, nrcf_pitch(N), lds, blk);
        else
            copy_field<NRC_COMPS>(f.nrc, nrc_pitch(N), lds, blk);
        nr_prologue<false, ROT, 0, (WS != 0)>(cst, f, vv, N, in, pro, valid);
        pin_prologue(pro);
        copy_wait();
        if (f32) widen_slots<NRC_COMPS>(lds);
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
        double *sg = fresh(f.sig);
#pragma unroll
        for (int k = 0; k < 6; ++k) stc(sg + (size_t)(k) * N, N, o8, r.sig[k]);
        fresh(f.nr_it)[v] = r.it;
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
// value k of the block partials part[k][0 .. nblocks-1] into out[k]; sh: the kernel's RED_BLOCK doubles of LDS
__device__ __forceinline__ void reduce_value(const double *__restrict__ part, int nblocks, unsigned long long maxmask,
                                             double *__restrict__ out, unsigned k, double *sh) {
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
}

__global__ void __launch_bounds__(RED_BLOCK) k_reduce(const double *__restrict__ part, int nblocks, int nvals,
                                                      unsigned long long maxmask, double *__restrict__ out,
                                                      const IterCtl *__restrict__ ctl) {
    if (ctl && ctl->done) return;
    __shared__ double sh[RED_BLOCK];
    reduce_value(part, nblocks, maxmask, out, blockIdx.x, sh);
    (void)nvals;
}

// the end-of-increment partials of every member of a lockstep ensemble: block (k, m) reduces value
// k of member m into its `red`
__global__ void __launch_bounds__(RED_BLOCK) k_reduce_end_ens(const EnsIncr *__restrict__ tab) {
    const EnsIncr &e = tab[blockIdx.y];
    __shared__ double sh[RED_BLOCK];
    reduce_value(e.partials, e.nblk, 0ull, e.red, blockIdx.x, sh);
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
// host launch wrappers
// ---------------------------------------------------------------------------
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

int nr_block_count(int nvox) { return (nvox + NR_BLOCK - 1) / NR_BLOCK; }

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

void launch_reduce_end_ens(hipStream_t s, const EnsIncr *tab, int nmem) {
    SRX_LAUNCH(k_reduce_end_ens, dim3(END_NPART, nmem), dim3(RED_BLOCK), 0, s, tab);
    check_launch("reduce_end_ens");
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

}  // namespace srx
