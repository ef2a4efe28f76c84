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
};

struct DevConst {
    DevPhase ph[NPH];
    double S066[6][6];    // inverse of the average stiffness, sample frame, B basis
    double C0[81];        // average stiffness, 4th order (init.c:370)
    double S0[81];        // its inverse, 4th order (init.c:371)
    double G6[36];        // C0 contracted over symmetric index pairs: A_ik = sum_b G6[(ik)][b] (n n)_b
    double dt;            // TimeStep
    double nu;            // ElastConst_PhaseI[3], used for every phase (constitutive.c:1383)
    double T, D0;
    double Udot_a[9];     // applied spin (antisymmetric part of Udot)
    int n[3];
    int elast_bc;
    int pad;
};

// Pointers to the per-voxel device arrays (component-major).
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
    float *dsig;       // [6][N] last Newton call's stress change, B basis, crystal frame (warm start only)
    int *phase;        // [N]
    int *nr_it;        // [N]
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
void set_launch_events(hipEvent_t start, hipEvent_t stop);

// ---- launch wrappers (kernels.hip) -------------------------------------
// Spectrum buffers: specB [c][kz][x][y], specC [c][kz][y][x]; T = double or float.
void launch_fft_fwd_z(hipStream_t s, int prec, const double *sig, void *specB, const int n[3],
                      const FftPlan &pz, const void *twz, const double *mean6);
void launch_fft_y(hipStream_t s, int prec, const void *in, void *out, int ncomp, int sign, int b_to_c,
                  const int n[3], const FftPlan &py, const void *twy);
void launch_green_x(hipStream_t s, int prec, void *specC, const int n[3], const FftPlan &px, const void *twx,
                    const DevConst *cst, const double *dc_sig6_sum);
void launch_fft_inv_z_update(hipStream_t s, int prec, const void *specB, double *dg, const int n[3],
                             const FftPlan &pz, const void *twz, const double *ddg9, double scale);
void launch_nr(hipStream_t s, const DevConst *cst, const Fields &f, int nvox, int first_call, int rotated,
               double *partials);
void launch_reduce(hipStream_t s, const double *partials, int nblocks, int nvals, double *out);
void launch_prologue(hipStream_t s, Fields f, int nvox, const double *add9, int creep, const double *set9,
                     const double *uold9);
void launch_end(hipStream_t s, const DevConst *cst, const Fields &f, int nvox, int lagged_possible, int rotated,
                int do_orient, int do_hardening, int istep, double *partials);
int nr_block_count(int nvox);
int end_block_count(int nvox);
// Field conversion helpers used by srx_get_field (device -> AoS tensor on device scratch)
void launch_b6_to_tensor(hipStream_t s, const double *v, int ncomp, int nvox, double *out9);
void launch_fft_real_forward_test(hipStream_t s, int prec, const double *in, int ncomp, void *specB, void *specC,
                                  const int n[3], const FftPlan *pl, void *const *tw);

}  // namespace srx
