// increment.hip -- CDNA4 (gfx950) per-increment kernels of the EVP-FFT loop, one voxel per
// lane in FP64: the increment prologue (evolution.c:1115-1128); the end of increment, Edot/Eps
// update, update_orient and DislocationEvolution (evolution.c:1268-1362); the GND option;
// and the device setup of orientations and initial stress.
#include "crystal_dev.hpp"

namespace srx {

// increment prologue (evolution.c:1115-1128): dg0 = dg; dg += Udot dt (or = set9 under creep)
struct Vec9 {
    double v[9];
};

// evolution.c:1115-1128.  In a creep run DisGrad is reset to the macroscopic value; the
// fluctuation it held (DisGrad minus the uniform part uold) is carried in fl0 so that the
// accumulated fluctuation of evolution.c:1198 survives the reset.
__device__ __forceinline__ void prologue_voxel(const Fields &f, int nvox, unsigned blk, const double *add, int creep,
                                               const double *set, const double *uold) {
    const int v = blk * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    const size_t N = (size_t)nvox;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const double d = f.dg[k * N + v];
        f.dg0[k * N + v] = d;
        f.dg[k * N + v] = creep ? set[k] : d + add[k];
        if (creep && f.fl0) f.fl0[k * N + v] += d - uold[k];
    }
}

constexpr int PRO_BLOCK = 256;

__global__ void k_prologue(Fields f, int nvox, Vec9 add, int creep, Vec9 set, Vec9 uold) {
    prologue_voxel(f, nvox, blockIdx.x, add.v, creep, set.v, uold.v);
}

// every member of a lockstep ensemble in one launch: member m owns blocks [pblk0, pblk0 + pnblk)
__global__ void __launch_bounds__(PRO_BLOCK) k_prologue_ens(const EnsIncr *__restrict__ tab, int nmem) {
    int m = 0;
    while (m + 1 < nmem && (int)blockIdx.x >= tab[m + 1].pblk0) ++m;
    const EnsIncr &e = tab[m];
    prologue_voxel(e.f, e.nvox, blockIdx.x - (unsigned)e.pblk0, e.add, e.creep, e.set, e.uold);
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
// Block blk of nblk (0: the launch's grid size) of one grid's end-of-increment launch.  The kernel declares the block's LDS
// (lv: END_NPART partial values per lane, lrs: NSYS densities per lane, both in columns of stride
// END_BLOCK; red: one row of partials per wavefront), so that every address is a constant.
template <bool GND>
__device__ __forceinline__ void end_block(const DevConst *__restrict__ cst, const Fields &f, int nvox,
                                          int lagged_possible, int do_orient, int istep,
                                          double *__restrict__ partials, unsigned blk, unsigned nblk, double *lv,
                                          double *lrs, double (*red)[END_NPART]) {
    const int v = blk * END_BLOCK + threadIdx.x;
    const size_t N = (size_t)nvox;
    const bool valid = v < nvox;
    const int ph = valid ? f.phase[v] : 0;
    // this lane's partial values, in an LDS column (stride END_BLOCK) rather than registers
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
        partials[(size_t)k * (nblk ? nblk : gridDim.x) + blk] = acc;
    }
}

template <bool GND>
__global__ void __launch_bounds__(END_BLOCK) __attribute__((amdgpu_waves_per_eu(SRX_END_WAVES_PER_EU, 8)))
k_end(const DevConst *__restrict__ cst, Fields f, int nvox, int lagged_possible, int do_orient, int istep,
      double *__restrict__ partials) {
    __shared__ double lv[END_NPART * END_BLOCK];
    __shared__ double lrs[NSYS * END_BLOCK];
    __shared__ double red[END_BLOCK / 64][END_NPART];
    end_block<GND>(cst, f, nvox, lagged_possible, do_orient, istep, partials, blockIdx.x, 0u, lv, lrs, red);
}

// every member of a lockstep ensemble in one launch: member m owns blocks [blk0, blk0 + nblk)
template <bool GND>
__global__ void __launch_bounds__(END_BLOCK) __attribute__((amdgpu_waves_per_eu(SRX_END_WAVES_PER_EU, 8)))
k_end_ens(const EnsIncr *__restrict__ tab, int nmem, int istep) {
    int m = 0;
    while (m + 1 < nmem && (int)blockIdx.x >= tab[m + 1].blk0) ++m;
    const EnsIncr &e = tab[m];
    __shared__ double lv[END_NPART * END_BLOCK];
    __shared__ double lrs[NSYS * END_BLOCK];
    __shared__ double red[END_BLOCK / 64][END_NPART];
    end_block<GND>(e.cst, e.f, e.nvox, e.lagged, e.orient, istep, e.partials, blockIdx.x - (unsigned)e.blk0,
                   (unsigned)e.nblk, lv, lrs, red);
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

int end_block_count(int nvox) { return (nvox + END_BLOCK - 1) / END_BLOCK; }

void launch_prologue(hipStream_t s, Fields f, int nvox, const double *a, int creep, const double *st,
                     const double *uo) {
    Vec9 add, set, uold;
    for (int k = 0; k < 9; ++k) {
        add.v[k] = a[k];
        set.v[k] = st[k];
        uold.v[k] = uo[k];
    }
    SRX_LAUNCH(k_prologue, dim3(prologue_block_count(nvox)), dim3(PRO_BLOCK), 0, s, f, nvox, add, creep, set, uold);
    check_launch("prologue");
}

int prologue_block_count(int nvox) { return (nvox + PRO_BLOCK - 1) / PRO_BLOCK; }

void launch_prologue_ens(hipStream_t s, const EnsIncr *tab, int nmem, int pblocks) {
    SRX_LAUNCH(k_prologue_ens, dim3(pblocks), dim3(PRO_BLOCK), 0, s, tab, nmem);
    check_launch("prologue_ens");
}

void launch_end_ens(hipStream_t s, const EnsIncr *tab, int nmem, int eblocks, int istep, int gnd) {
    if (gnd)
        SRX_LAUNCH(k_end_ens<true>, dim3(eblocks), dim3(END_BLOCK), 0, s, tab, nmem, istep);
    else
        SRX_LAUNCH(k_end_ens<false>, dim3(eblocks), dim3(END_BLOCK), 0, s, tab, nmem, istep);
    check_launch("end_ens");
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

}  // namespace srx
