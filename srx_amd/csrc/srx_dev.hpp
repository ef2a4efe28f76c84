// srx_dev.hpp -- data shared between the host driver and the HIP kernels.
//
// Device state is structure-of-arrays in FP64: every per-voxel component is
// a contiguous array of nvox values in the reference's voxel order
// (x slowest, index (x*N1+y)*N2+z, evp.h:60).  Symmetric second-order
// tensors are kept in the orthonormal B basis of kinematics.c:35-53
// (components 0..4 deviatoric, 5 hydrostatic).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace srx {

constexpr int NSYS = 12;
constexpr int NPH = 2;
constexpr int MAX_RADIX_STAGES = 12;
// Component pitch of Fields::nrc: the voxel count rounded up to even, so that every component
// starts 16-byte aligned.
__host__ __device__ inline size_t nrc_pitch(size_t n) { return (n + 1) & ~(size_t)1; }
// The same for the FP32 copy of the constants (mixed precision): a multiple of four voxels, so that
// one 16-byte load moves four of a component's values.  36 nrcf_pitch(n) floats fit in the
// 36 nrc_pitch(n) doubles of the FP64 array.
__host__ __device__ inline size_t nrcf_pitch(size_t n) { return (n + 3) & ~(size_t)3; }
// Largest grid (voxels): the per-voxel kernels address element v of a component, or of the
// component after it, by a 32-bit byte offset (8 v, 8 (N + v)) from the component's base.
constexpr long long SRX_MAX_VOXELS = (1ll << 28) - 1;

// Element v of one component of a field, read or written as a raw buffer access: the component's
// base address and byte length in a buffer resource (scalar registers, wavefront-uniform) and the
// lane's 32-bit byte offset o8 = 8 v.  A field access written as base[k * N + v] makes a 64-bit
// address per lane and component, which the compiler computes early and keeps live; here the only
// per-lane value is o8.  (An offset past the component reads 0 and writes nothing.)
#if defined(__HIP__)
// Byte length of a buffer resource: saturated at 2^32 - 1 (every offset used stays below 2^32).
__device__ __forceinline__ int rsrc_bytes(size_t bytes) {
    return (int)(unsigned)(bytes > 0xFFFFFFFFull ? 0xFFFFFFFFull : bytes);
}
__device__ __forceinline__ __amdgpu_buffer_rsrc_t comp_rsrc(const double *base, size_t n) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(base), (short)0, rsrc_bytes(n * 8u), 0x00020000);
}
__device__ __forceinline__ double ldc(const double *base, size_t n, unsigned o8) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(comp_rsrc(base, n), o8, 0, 0));
}
// The same for an FP32 field (lane byte offset o4 = 4 v).
__device__ __forceinline__ float ldcf(const float *base, size_t n, unsigned o4) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(base), (short)0, rsrc_bytes(n * 4u), 0x00020000), o4, 0,
        0));
}
__device__ __forceinline__ void stcf(float *base, size_t n, unsigned o4, float x) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x),
                                          __builtin_amdgcn_make_buffer_rsrc(base, (short)0, rsrc_bytes(n * 4u), 0x00020000),
                                          o4, 0, 0);
}
typedef unsigned int srx_u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void stc(double *base, size_t n, unsigned o8, double x) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(srx_u32x2, x), comp_rsrc(base, n), o8, 0, 0);
}
#endif

// Material constants of one phase, precomputed on the host in the
// reference's evaluation order where it matters for rounding.
struct DevPhase {
    int xtal;             // Type_phases == 0
    int nsys;
    double p[NSYS][5];    // Schmid tensors in the crystal frame, B basis (init.c:208-214)
    double bn[NSYS][3], bb[NSYS][3];
    double chi[NSYS][NSYS], chi1[NSYS][NSYS];
    double sinv[6];       // crystal-frame compliance in the B basis (diagonal)
    double c1Gb;          // c_1*G*b           (tau_pass prefix, constitutive.c:587)
    double kTf;           // kT/MY_J_IN_EV     (tau_cut prefix, constitutive.c:630)
    double c2, c3, b, tau0;
    double rhom_k;        // 2 kT/c1/c2/c3/G/b^3 (rho_m prefix, constitutive.c:632)
    double expmq;         // exp(-q_slip)
    double G, c4, c5, c6;
    double expmqb;        // exp(-q_bulk)
    double c1Gb3;         // c_1*G*b*1e-3: tau_pass = c1Gb3*sqrt(|rho_P + rho_m|)
    double k_tcut;        // kT/MY_J_IN_EV/c2/c3/b^2*0.1: tau_cut = k_tcut*sqrt(rho_f)
    double k_g0;          // 1e-3*TRIAL_FREQ*c2*b*exp(-q_slip): gamma_0 = k_g0*rho_m/sqrt(rho_f)
    // GND option: bp = bb x bn (init.c:200-202) and the projections of init.c:1118-1141
    double bp[NSYS][3];
    double cos_g1[NSYS][NSYS], cos_g2[NSYS][NSYS], sin_g1[NSYS][NSYS], sin_g2[NSYS][NSYS];
};

struct DevConst {
    DevPhase ph[NPH];
    double S066[6][6];    // inverse of the average stiffness, sample frame, B basis
    double C0[81];        // average stiffness, 4th order (init.c:370)
    double S0[81];        // its inverse, 4th order (init.c:371)
    double G6[36];        // C0 contracted over symmetric index pairs: A_ik = sum_b G6[(ik)][b] (n n)_b
    float G6f[36];        // the same in FP32 (the mixed-precision Green operator)
    double dt;            // TimeStep
    double nu;            // ElastConst_PhaseI[3], used for every phase (constitutive.c:1383)
    double T, D0;
    double Udot_a[9];     // applied spin (antisymmetric part of Udot)
    int n[3];
    int elast_bc;
    int gnd;              // 1: GND densities evolve and enter rho_F, rho_P, rho_tot (a DD_GND build)
    double L0;            // deck L0: length that scales the GND rates (constitutive.c:1437, 1470)
    // GND option (null otherwise).  Kept here rather than among the kernel arguments so that kernels
    // load them only inside their GND branch.
    double *rho_g1, *rho_g2;   // [12][N] GND densities, screw and edge-I (edge-II stays zero in the
                               //   first-order model)
    double *gnd_fp;       // [24][N] GND parts of rho_F (12) and rho_P (12) of the current densities
    unsigned char *gbc;   // [N] same-grain flags of the +x,-x,+y,-y,+z and 'z-' neighbours (bits 0..5)
};

// Outer fixed-point control kept on the device.  The mixed boundary-condition algebra
// (get_smacro, evolution.c:590-664) and the convergence test (evolution.c:1150, 1246-1252) run in
// a one-thread kernel after each iteration's reduction, so the host enqueues several iterations
// without reading anything back in between; every kernel of an iteration returns at once when
// `done` is set.
struct IterCtl {
    // set by the host at the start of each increment
    double wgt, wgt_ph1, err, e_vm_prev, svm_stray, nvox;
    double Scauchy[9];
    int vbc[6], sbc[6];
    int svm_mode, itermax, fixed_iters, mixed;
    // advanced by the control kernel
    int iter, done, nan, pad0;
    long long nrtot;
    int nrmax, pad1;
    double nrwave;
    double Err_s, Err_e;
    double sig_sum6[6];   // grid sums of the stress (B basis) after the last iteration
    double mean6[6];      // their means: subtracted before the FP32 cast (mixed precision)
    double dcs6[6];       // the sums reinjected at k = 0 (mixed precision), else 0
    double ddg[9];        // dDisGradAvg, applied by the next inverse-z pass
    double acum[9], uniform[9];
    double SigAvg[9], SigAvg1[9], s_vm, s_vm1;
};

// B basis of kinematics.c:35-53: component a,b of basis tensor i
__host__ __device__ constexpr double bbasis(int i, int a, int b) {
    constexpr double R2 = 0.70710678118654744, R3 = 0.57735026918962584, R6 = 0.40824829046386304;
    switch (i) {
    case 0: return a == b ? (a == 0 ? -R2 : (a == 1 ? R2 : 0.0)) : 0.0;
    case 1: return a == b ? (a == 2 ? 2. * R6 : -R6) : 0.0;
    case 2: return (a == 1 && b == 2) || (a == 2 && b == 1) ? R2 : 0.0;
    case 3: return (a == 0 && b == 2) || (a == 2 && b == 0) ? R2 : 0.0;
    case 4: return (a == 0 && b == 1) || (a == 1 && b == 0) ? R2 : 0.0;
    default: return a == b ? R3 : 0.0;
    }
}
__host__ __device__ inline void ctl_b6_to_t9(const double v[6], double t[9], int dim) {
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            double s = 0.0;
            for (int i = 0; i < dim; ++i) s += v[i] * bbasis(i, a, b);
            t[3 * a + b] = s;
        }
}
__host__ __device__ inline void ctl_t9_to_b6(const double t[9], double v[6]) {
    for (int i = 0; i < 6; ++i) {
        double s = 0.0;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) s += t[3 * a + b] * bbasis(i, a, b);
        v[i] = s;
    }
}

// get_smacro (evolution.c:590-664): volume averages, the strain correction dDisGradAvg of the
// stress-controlled components (upper-triangle entries only), and the two von Mises measures
// (the first with the out-of-bounds SigDevAvg[3][3] term of evolution.c:648 taken as svm_stray).
__host__ __device__ inline void smacro_update(IterCtl &c, const double *S0, const double sum6[6],
                                              const double sum6_ph1[6]) {
    const int IJV[2][6] = {{0, 1, 2, 1, 0, 0}, {0, 1, 2, 2, 2, 1}};
    double t[9], t1[9];
    ctl_b6_to_t9(sum6, t, 6);
    ctl_b6_to_t9(sum6_ph1, t1, 6);
    for (int k = 0; k < 9; ++k) {
        c.SigAvg[k] = t[k] * c.wgt;
        c.SigAvg1[k] = t1[k] * (c.wgt / c.wgt_ph1);
    }
    for (int i = 0; i < 6; ++i) {
        const int ii = IJV[0][i], jj = IJV[1][i];
        double acc = 0.0;
        if (c.vbc[i] == 0) {
            for (int k = 0; k < 6; ++k) {
                const int kk = IJV[0][k], ll = IJV[1][k];
                acc += S0[((ii * 3 + jj) * 3 + kk) * 3 + ll] * c.sbc[k] * (c.Scauchy[3 * kk + ll] - c.SigAvg[3 * kk + ll]);
            }
        }
        c.ddg[3 * ii + jj] = acc;
    }
    for (int k = 0; k < 9; ++k) c.acum[k] += c.ddg[k];
    double sav6[6], sda[9];
    ctl_t9_to_b6(c.SigAvg, sav6);
    ctl_b6_to_t9(sav6, sda, 5);
    const double tr = c.svm_mode == 1 ? (sda[0] + sda[4] + sda[8]) : (sda[4] + sda[8] + c.svm_stray);
    double s = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const double x = sda[3 * i + j];
            s += i != j ? x * x : (x - tr / 3.0) * (x - tr / 3.0);
        }
    c.s_vm = sqrt(3. / 2. * s);
    ctl_t9_to_b6(c.SigAvg1, sav6);
    ctl_b6_to_t9(sav6, sda, 5);
    s = 0.0;
    for (int k = 0; k < 9; ++k) s += sda[k] * sda[k];
    c.s_vm1 = sqrt(3. / 2. * s);
}

// whether the outer loop stops before iteration c.iter + 1 (evolution.c:1150; fixed_iters > 0:
// exactly that many iterations)
__host__ __device__ inline int ctl_stop(const IterCtl &c) {
    if (c.fixed_iters > 0) return c.iter >= c.fixed_iters;
    const double e = fabs(c.Err_s) > fabs(c.Err_e) ? fabs(c.Err_s) : fabs(c.Err_e);
    return !(c.iter < c.itermax && e > c.err);
}

// Largest pitch (complex values) of one LDS pencil of length n in the z and y passes, for sizing:
// a pencil the compile-time stages transform is stored swizzled at pitch n + 2 (a multiple of 32
// long), so that the pencils a wavefront covers start on different LDS banks; others at n + 1.
__host__ __device__ inline int lds_pitch(int n) { return n % 32 == 0 ? n + 2 : n + 1; }

// Pointers to the per-voxel device arrays (component-major).
struct MlpPacked;   // bf16 operands of the learned Newton start (mlp_dev.hpp)

struct Fields {
    double *sig;       // [6][N] stress, B basis, sample frame
    double *sig_pre;   // [5][N] last NR pre-update iterate, crystal frame of tran (deviatoric)
    double *eps;       // [5][N] plastic strain, B basis (deviatoric)
    double *edot;      // [5][N] plastic strain rate, B basis
    double *dg;        // [9][N] displacement gradient DisGrad, row-major tensor comps
    double *dg0;       // [9][N] DisGrad at the start of the increment
    double *tran;      // [9][N] TranMat_xt2sa (crystal -> sample), current
    double *tran0;     // [9][N] TranMat_xt2sa at setup (C_gr is never rotated)
    double *rho_s, *rho_m, *rho_m_old, *rho_F, *rho_P, *gamdot, *gam_acum;   // [12][N]
    double *rho_dipole, *rho_forest, *rho_inplane;                          // [12][N]
    double *rho_tot;   // [N] rho_tot of the last hardening update (evolution.c:1030-1048)
    double *fl0;       // [9][N] fluctuation carried over DisGrad resets (creep runs only, else null)
    double *acr;       // [15][N] crystal-frame Newton matrix A = Q^T S066 Q + S' (packed symmetric 5x5) of the
                       //   increment, written by its first Newton call, read by the later ones
    double *nrc;       // [36][nrc_pitch(N)] per-system Newton constants of the increment, written by its first call:
                       //   tau_pass/(tau_0+tau_cut) [12], 1/(tau_0+tau_cut) [12], gamma_0 dt [12]
    float *acrf;       // mixed precision: A rounded to FP32, [15][N] in acr's memory; null in FP64
    float *nrcf;       // mixed precision: the same constants rounded to FP32, [36][nrcf_pitch(N)] in nrc's memory
                       //   (every Newton call of the increment uses the rounded values); null in FP64
    float *dsig;       // [6][N] last Newton call's stress change, B basis, crystal frame (warm start only)
    float *ecp;        // [6][N] eps_c = sym(DisGrad) (B basis, sample frame) of the last Newton call (warm start only)
    const MlpPacked *mlp;   // learned Newton start (non-null: the start comes from the network)
    int *phase;        // [N]
    int *nr_it;        // [N]
};

// One member of an ensemble whose per-voxel kernels run in one launch: the member owns the grid
// blocks [blk0, blk0 + nblk) of the Newton kernel.
struct EnsMember {
    const DevConst *cst;
    Fields f;
    int nvox, blk0, nblk, pad;
    double *partials;   // NR_NPART x nblk block partials
    double *redseg;     // NR_NPART x 16 segment sums
    IterCtl *ctl;
};

// One member of a lockstep ensemble's once-per-increment launches (k_prologue_ens, k_end_ens and
// the reduction of the end-of-increment partials): the member owns the prologue blocks
// [pblk0, pblk0 + pnblk) and the end blocks [blk0, blk0 + nblk) of those launches.
struct EnsIncr {
    const DevConst *cst;
    Fields f;
    int nvox, pblk0, pnblk, blk0, nblk;
    int creep, lagged, orient;   // the prologue's creep flag; k_end's lagged_possible and do_orient
    double add[9], set[9], uold[9];   // the prologue's arguments
    double *partials;   // END_NPART x nblk block partials
    double *red;        // END_NPART reduced values
};

// Radix plan of one 1-D FFT length.
struct FftPlan {
    int n;
    int nst;
    int rad[MAX_RADIX_STAGES];
};

// Per-block partial sums of the NR kernel epilogue.
constexpr int NR_NPART = 18;   // conv_s, conv_e, sig6[6], sig6_ph1[6], nr_sum, nr_max, nan_count, wave nr_max sum
// Per-block partial sums of the end-of-increment kernel.
constexpr int END_NPART = 24;  // eps6[6], edot6[6], rhoF, rhoP, rhos, rhom, rhodip, rhofor, rhoinp, rsl, rlc, nan, pad..

bool make_plan(int n, FftPlan &p);
// radix plans of length n the autotuner may time, make_plan's first (fft.hip)
int plan_candidates(int n, FftPlan *out, int cap);
// whether the FFT passes take an n0 x n1 x n2 grid in precision prec (host.cpp)
bool srx_grid_supported(const FftPlan pl[3], const int n[3], int prec);
void set_launch_events(hipEvent_t start, hipEvent_t stop);

// ---- launch wrappers (fft.hip, newton.hip, increment.hip) ---------------
// Spectrum buffers: specB [c][kz][x][y], specC [c][kz][y][x]; T = double or float.
// With a non-null `ctl` the iteration's scalars (mean stress, k = 0 sums, dDisGradAvg) come from
// the device control block and the kernels return at once when its `done` is set.
// `tile`: rows (pencils) per block, 0 for the default; fft_tile_max gives the largest valid one of
// a pass (0 forward z, 1 forward y, 2 Green x, 3 inverse y, 4 inverse z; p: the pass axis's plan).
int fft_tile_max(int pass, const int n[3], const FftPlan &p, int prec);
// The solver's Green x pass can run one wavefront per row (tile -1; the default tile 0 selects it)
// for n0 = 64, 128, 256.
bool green_wave_ok(const int n[3]);
void launch_fft_fwd_z(hipStream_t s, int prec, const double *sig, void *specB, const int n[3],
                      const FftPlan &pz, const void *twz, const IterCtl *ctl, int tile = 0);
void launch_fft_y(hipStream_t s, int prec, const void *in, void *out, int ncomp, int sign, int b_to_c,
                  const int n[3], const FftPlan &py, const void *twy, const IterCtl *ctl = nullptr, int tile = 0);
void launch_green_x(hipStream_t s, int prec, void *specC, const int n[3], const FftPlan &px, const void *twx,
                    const DevConst *cst, const IterCtl *ctl, int tile = 0);
void launch_fft_inv_z_update(hipStream_t s, int prec, const void *specB, double *dg, const int n[3],
                             const FftPlan &pz, const void *twz, const IterCtl *ctl, double scale, int tile = 0);
// ph2: the second phase is crystalline (its Newton path is compiled in only then)
void launch_nr(hipStream_t s, const DevConst *cst, const Fields &f, int nvox, int first_call, int rotated,
               double *partials, const IterCtl *ctl = nullptr, int gnd = 0, int ph2 = 1);
void launch_reduce(hipStream_t s, const double *partials, int nblocks, int nvals, double *out,
                   const IterCtl *ctl = nullptr);
// The per-iteration reduction of the Newton kernel's block partials in two fixed-order stages:
// launch_reduce_nr writes NR_NPART x 16 segment results to seg, launch_control sums them and does
// the iteration's bookkeeping (get_smacro and the convergence test).
constexpr int NR_RED_SEG_DOUBLES = NR_NPART * 16;
void launch_reduce_nr(hipStream_t s, const double *partials, int nblocks, double *seg, const IterCtl *ctl);
void launch_control(hipStream_t s, const DevConst *cst, const double *seg, IterCtl *ctl);
// The same three kernels over every member of an ensemble in one launch each (tab: device array of
// nmem members; nblocks: their Newton blocks in total; ws: the Newton start option; done[m]: each
// member's stop flag after the iteration).
void launch_nr_ens(hipStream_t s, const EnsMember *tab, int nmem, int nblocks, int first_call, int rotated, int ws,
                   int gnd, int ph2);
void launch_reduce_nr_ens(hipStream_t s, const EnsMember *tab, int nmem);
void launch_control_ens(hipStream_t s, const EnsMember *tab, int nmem, int *done);
void launch_prologue(hipStream_t s, Fields f, int nvox, const double *add9, int creep, const double *set9,
                     const double *uold9);
void launch_end(hipStream_t s, const DevConst *cst, const Fields &f, int nvox, int lagged_possible, int rotated,
                int do_orient, int do_hardening, int istep, double *partials, int gnd = 0);
// The increment prologue, the end-of-increment kernel and the reduction of its partials (into each
// member's `red`) over every member of a lockstep ensemble in one launch each (tab: device array of
// nmem members; pblocks, eblocks: their prologue and end blocks in total).
int prologue_block_count(int nvox);
void launch_prologue_ens(hipStream_t s, const EnsIncr *tab, int nmem, int pblocks);
void launch_end_ens(hipStream_t s, const EnsIncr *tab, int nmem, int eblocks, int istep, int gnd);
void launch_reduce_end_ens(hipStream_t s, const EnsIncr *tab, int nmem);
// GND option: gradient of gamdot (grad = 1), GND density update, GND parts of rho_F / rho_P and one
// partial per block (gnd_block_count) of sum sqrt(rho_g1^2 + rho_g2^2)
int gnd_block_count(int nvox);
void launch_gnd(hipStream_t s, const DevConst *cst, const DevConst &hc, const Fields &f, int nvox, int grad,
                double *partials);
// Learned Newton start: operand packing (mlp.hip) and the recording of training samples around a
// Newton call (newton.hip): `after` = 0 writes the features of the sampled voxels before the
// call, 1 their targets after it and advances the sample count.
void mlp_pack(hipStream_t s, const float *dblob, MlpPacked *M);
size_t mlp_packed_bytes();
void launch_mlp_record(hipStream_t s, const DevConst *cst, const Fields &f, int nvox, int stride, int after,
                       float *X, float *T, float *ND, int *count, int cap, const IterCtl *ctl);
// Device setup: TranMat (and its initial copy) from Bunge angles in degrees (times deg = pi/180),
// the initial stress from e9 = Udot dt (+ sc9 for creep runs), and out[0..36) = sum of C_gr
// (B basis, row-major 6x6), out[36..42) = sum of the initial stress.  partials: 42 x
// setup_block_count(nvox) doubles.
int setup_block_count(int nvox);
void launch_setup(hipStream_t s, const DevConst *cst, const Fields &f, int nvox, const double *eul, double deg,
                  const double *e9, const double *sc9, int creep, double *partials, double *out);
int nr_block_count(int nvox);
// Streaming copy of n16 16-byte elements (the HBM ceiling measurement): `unroll` (1, 2, 4)
// accesses in flight per lane, `blocks` blocks of 256 threads, grid-stride.
void launch_stream_copy(hipStream_t s, const void *in, void *out, size_t n16, int unroll, int blocks);
int end_block_count(int nvox);
// Field conversion helpers used by srx_get_field (device -> AoS tensor on device scratch)
void launch_b6_to_tensor(hipStream_t s, const double *v, int ncomp, int nvox, double *out9);
void launch_fft_real_forward_test(hipStream_t s, int prec, const double *in, int ncomp, void *specB, void *specC,
                                  const int n[3], const FftPlan *pl, void *const *tw);

}  // namespace srx
