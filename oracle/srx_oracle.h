/*
 * srx_oracle.h -- CPU FP64 restatement of the SRX EVP-FFT increment loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU product is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product path never links or calls it.
 *
 * It restates, in plain C with OpenMP, what the reference does in
 * evolution.c:157-427 (UpdateStress), 590-664 (get_smacro), 1030-1515
 * (Evolution), constitutive.c:142-273 (update_orient), 507-728
 * (StrainRate_Orowan, FORWARD_PLUS_BACKWARD branch), 1353-1529
 * (DislocationEvolution), 1681-1808 (AvgDislocation), kinematics.c:5-483,
 * init.c:10-703,728-1216 (non-DRX parts) and io.c:25-164 (GetNameList).
 * The FFT is a plain complex DFT (radix-2 for powers of two, direct DFT
 * otherwise) with FFTW's conventions: forward sign -1, unnormalised.
 */
#ifndef SRX_ORACLE_H
#define SRX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_state orc_state;

/* Same layout as srx_options in include/srx.h. */
typedef struct {
    int    precision;      /* ignored by the oracle (always FP64) */
    int    fixed_iters;    /* >0: run exactly this many outer iterations per increment */
    double err_override;   /* >0: replaces the deck's Err */
    int    itermax_override; /* >0: replaces the deck's IterMax */
    int    nsteps_override;  /* >0: replaces the deck's N_steps */
    int    tau0_mode;      /* 0: tau0xf of the phase; 1: use tau0_value */
    double tau0_value;
    int    svm_mode;       /* 0: evolution.c:648 quirk with stray term = svm_stray; 1: standard von Mises */
    double svm_stray;
    int    write_files;    /* 1: write vm.out, dd.out, err.out, drx.out into workdir */
    int    write_fields;   /* 1: also write the .iout field dumps at PrintControl steps */
    int    nthreads;       /* OpenMP threads (0 = default) */
    int    verbose;
    int    warm_start;     /* 1: extrapolated Newton start (no counterpart in the reference) */
    int    gnd;            /* 1: geometrically necessary dislocations as a DD_GND build (first order) */
    int    reserved[7];
} orc_options;

/* Same layout as srx_step_report in include/srx.h. */
typedef struct {
    int    istep;
    int    iters;               /* outer iterations run */
    int    converged;           /* 1 if max(|Err_s|,|Err_e|) <= Err at exit */
    int    nr_iters_max;        /* largest NR iteration count of any voxel in any outer iteration */
    long   nr_iters_total;      /* sum over voxels and outer iterations */
    double err_s, err_e;        /* normalised errors at exit */
    double time_tot, e_vm, evmp, d_vm, dvmp, s_vm, s_vm1;
    double rho_F_avg, rho_P_avg, rho_s_avg, rho_m_avg, rho_g_avg;
    double rho_dipole_avg, rho_forest_avg, rho_inplane_avg;
    double sig_avg[9], sig_avg1[9], eps_avg[9], edot_avg[9];
    double disgrad_avg_actual[9], velgrad_avg[9];
    double nr_wave_iters;       /* sum over groups of 64 consecutive voxels of the group's largest NR count */
    double reserved[15];
} orc_report;

/* Field ids, same as include/srx.h; all returned voxel-major, x slowest. */
enum {
    ORC_F_SIG = 0, ORC_F_EPS = 1, ORC_F_EDOT = 2, ORC_F_DISGRAD = 3,
    ORC_F_VELGRAD = 4, ORC_F_TRANMAT = 5,
    ORC_F_RHO_S = 6, ORC_F_RHO_M = 7, ORC_F_RHO_F = 8, ORC_F_RHO_P = 9,
    ORC_F_GAMDOT = 10, ORC_F_GAM_ACUM = 11, ORC_F_RHO_DIPOLE = 12,
    ORC_F_RHO_FOREST = 13, ORC_F_RHO_INPLANE = 14, ORC_F_NR_ITERS = 15,
    ORC_F_FLUCT = 16, ORC_F_TRIAL_GAM_ACUM = 17, ORC_F_RHO_TOT = 18,
    ORC_F_PHASE = 19, ORC_F_GRAIN = 20, ORC_F_RHO_G1 = 21, ORC_F_RHO_G2 = 22
};

void orc_default_options(orc_options *o);
int  orc_open(const char *deck, const char *workdir, const orc_options *opt, orc_state **out);
int  orc_step(orc_state *s, orc_report *rep);
int  orc_run(orc_state *s);
int  orc_get_field(orc_state *s, int field, double *out, long n);
int  orc_set_field(orc_state *s, int field, const double *in, long n);
int  orc_dims(orc_state *s, int *n3);
void orc_close(orc_state *s);
/* Scalars and tables of the setup (for the known-answer tests):
   which: 0 C066[36] 1 S066[36] 2 kT,q_slip,q_bulk,bb,shear_G 3 Schm_xt[12*5] of phase 1,
          4 Chi[144] 5 Chi_1[144] 6 elastic constants c11,c12,c44,nu of phase 1 7 Udot[9] 8 Scauchy[9]
          9 d_vm,e_vm (initial) 10 Wgt_ph1 */
int  orc_get_setup(orc_state *s, int which, double *out, long n);

/* Single-function entry points used by the unit tests. */
void orc_chg_basis(double *v6, double *t33, double *c66, double *t3333, int opt, int dim);
void orc_lu_inv66(double *c66);
void orc_euler_to_matrix(double *t1, double *ph, double *t2, double *a33, int opt);
void orc_fft3d(double *data, int n0, int n1, int n2, int sign, int nthreads);
/* Green contraction of one k-space point exactly as CalcGAMMA + evolution.c:1169-1183.
   c0: 81 doubles; nvec: 3; ks: 9 complex (re,im) input; du: 9 complex output */
void orc_green_point(const double *c0, const double *nvec, int dc, int elast_bc,
                     const double *s0, const double *ks, double *du);
/* StrainRate_Orowan for one voxel of the state (idx), at stress sig6.  Writes
   edot6[6] and dedot[36] and updates the voxel's rho_m/rho_F/rho_P/gamdot. */
int  orc_strain_rate(orc_state *s, int idx, const double *sig6, double *edot6, double *dedot);
/* One NR solve (UpdateStress body) for voxel idx; returns NR iterations. */
int  orc_nr_voxel(orc_state *s, int idx, double *err_s, double *err_e);
/* update_schmid (kinematics.c:430-463) and the sample-frame Schmid tensors of a voxel */
void orc_update_schmid(orc_state *s);
/* Orientation (constitutive.c:142-194) on one 3x3 TranMat and spin increment (rot is overwritten
   with the rotation applied); update_orient (constitutive.c:196-273) over the current fields. */
void orc_orientation(double *tran9, double *rot9);
void orc_update_orient(orc_state *s);
int  orc_get_schmid(orc_state *s, int idx, double *out60);

#ifdef __cplusplus
}
#endif
#endif
