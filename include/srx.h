/*
 * srx.h -- C ABI of libsrx, the MI355X EVP-FFT crystal-plasticity solver.
 *
 * The reference (mesoOSU/SRX) has no plugin or FFI interface: its hot path is
 * reached through the process entry `trial <deck>` (evp.c:7-31), which calls
 * three void(void) functions over process globals.  libsrx replaces that
 * entry with a context-based C ABI.  Each declaration below names the
 * reference interface it replaces.
 *
 * Conventions: plain pointers and sizes; every call returns 0 or an error
 * code (never aborts the process); the reference's own PError/exit codes are
 * kept (4, 7, 28, 35, 114, 116, 230, 445, 604, 720, 1110, 1139) and new codes
 * start at 2000.  One srx_ctx per host thread; contexts are independent.
 * Field arrays use the reference's voxel order (x slowest, index
 * (x*N1 + y)*N2 + z, evp.h:60) and its per-voxel layout (3x3 row-major
 * tensors, 12 values per slip-system array).
 */
#ifndef SRX_H
#define SRX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct srx_ctx srx_ctx;

/* Precision of the FFT + Green-operator passes. */
enum { SRX_PREC_FP64 = 0, SRX_PREC_MIXED = 1 };

/* Error codes added by libsrx (the reference aborts or fails silently). */
enum {
    SRX_OK = 0,
    SRX_E_FILE = 2001,          /* deck, slip-mode or mesh file cannot be opened */
    SRX_E_GRID = 2002,          /* CellDim not supported: each extent must be 2^a 3^b 5^c, <= 1024; the x
                                   extent CellDim[0] also within the x pass's LDS row (10 n0 complex values
                                   in 48 KB, 9 n0 / smallest radix <= 1024): up to 300 in FP64 (not 250,
                                   270) and up to 512 in mixed precision (not 250, 270, 360, 375, 384, 405,
                                   432, 450, 480, 486, 500); at most 2^28 - 1 voxels in all */
    SRX_E_HIP = 2003,           /* HIP runtime error */
    SRX_E_SYSTEMS = 2004,       /* more than 12 slip systems in a mode */
    SRX_E_NOMEM = 2005,         /* allocation failed */
    SRX_E_DONE = 2006,          /* srx_step after the last increment */
    SRX_E_CREEP_STOP = 2007,    /* creep run reached e_vm > 0.3 (evolution.c:1020-1022) */
    SRX_E_SIZE = 2008,          /* caller buffer too small */
    SRX_E_FIELD = 2009,         /* unknown field id */
    SRX_E_ARG = 2010,           /* invalid argument */
    SRX_E_SCRATCH = 2011,       /* problem larger than the static scratch of srx_fft_forward_noalloc */
    SRX_E_NAN = 3001,           /* NaN in the constitutive update (the reference asserts, constitutive.c:660,683,707) */
    SRX_E_UPDATE_FLAG = 3002    /* Update_Flag 0 leaves Schm_gr uninitialised at istep 0 (init.c:815): rejected */
};

typedef struct {
    int    precision;        /* SRX_PREC_FP64 (default) or SRX_PREC_MIXED: FP32 FFT/Green passes and
                                FP32 storage of each increment's Newton matrix and per-system
                                constants, FP64 Newton arithmetic, hardening and residual */
    int    fixed_iters;      /* >0: exactly this many outer iterations per increment (parity runs) */
    double err_override;     /* >0: replaces the deck's Err */
    int    itermax_override; /* >0: replaces the deck's IterMax */
    int    nsteps_override;  /* >0: replaces the deck's N_steps */
    int    tau0_mode;        /* 0: lattice resistance = tau0xf of the phase; 1: tau0_value */
    double tau0_value;
    int    svm_mode;         /* 0: reproduce evolution.c:648 (out-of-bounds term := svm_stray); 1: standard */
    double svm_stray;
    int    write_files;      /* 1: write vm.out, dd.out, err.out, drx.out into the work directory */
    int    write_fields;     /* 1: also the .iout field dumps at the deck's PrintControl steps (evolution.c:1456-1483) */
    int    nthreads;         /* host threads for setup (0: default) */
    int    verbose;          /* 1: per-iteration errors on stdout */
    int    warm_start;       /* 1: each Newton solve after the first of an increment starts from
                                the stress extrapolated from the last two outer iterations instead of
                                the last one (converges to the same tolerance in fewer iterations;
                                results differ from the reference within that tolerance).
                                2: the start comes from the learned network installed by srx_mlp_set
                                (same convergence test; no counterpart in the reference) */
    int    gnd;              /* 1: geometrically necessary dislocations, as a build of the reference with
                                DD_GND (evpFlags.h, which then also selects DD_GND_FIRST_ORDER): after each
                                increment the grain-boundary-aware gradient of the slip rates
                                (Gradient_ShearRate, constitutive.c:882-1115) updates the screw and edge-I
                                GND densities (constitutive.c:1428-1519, deck key L0), which enter rho_F,
                                rho_P (constitutive.c:565-576) and rho_tot (evolution.c:1040-1046);
                                read them with SRX_F_RHO_G1 / SRX_F_RHO_G2. Needs warm_start = 0 (else
                                srx_open returns SRX_E_ARG). 0 (default): the shipped build */
    int    reserved[7];      /* [0]: kernel timing (srx_kernel_times); [1] = 1: enqueue the outer
                                iterations launch by launch instead of replaying each batch of 8
                                from a HIP graph captured on first use (the default); [2] = 1: keep
                                the default FFT tiles instead of timing the candidates at srx_open;
                                [3] = 1: also time the FFT radix plans (srx_fft_plan) */
} srx_options;

/* Everything the reference prints or writes per increment, plus counters it
   does not keep (iterations, NR statistics). */
typedef struct {
    int    istep;
    int    iters;               /* outer fixed-point iterations run */
    int    converged;           /* max(|Err_s|,|Err_e|) <= Err at exit */
    int    nr_iters_max;        /* largest per-voxel NR iteration count in the increment */
    long   nr_iters_total;      /* sum of per-voxel NR iterations over the increment */
    double err_s, err_e;        /* normalised errors at exit (evolution.c:1246-1252) */
    double time_tot, e_vm, evmp, d_vm, dvmp, s_vm, s_vm1;   /* vm.out columns */
    double rho_F_avg, rho_P_avg, rho_s_avg, rho_m_avg, rho_g_avg;
    double rho_dipole_avg, rho_forest_avg, rho_inplane_avg; /* AvgDislocation */
    double sig_avg[9], sig_avg1[9], eps_avg[9], edot_avg[9];
    double disgrad_avg_actual[9], velgrad_avg[9];
    double nr_wave_iters;       /* sum over groups of 64 consecutive voxels and outer iterations of the
                                   group's largest NR iteration count (64x this = lane-iterations a
                                   wavefront-per-group solver executes) */
    double reserved[15];
} srx_step_report;

/* Field ids for srx_get_field / srx_set_field (values per voxel in brackets). */
enum {
    SRX_F_SIG = 0,        /* [9] Cauchy stress Sig (evp.h:429) */
    SRX_F_EPS = 1,        /* [9] plastic strain Eps */
    SRX_F_EDOT = 2,       /* [9] plastic strain rate Edot */
    SRX_F_DISGRAD = 3,    /* [9] displacement gradient DisGrad */
    SRX_F_VELGRAD = 4,    /* [9] velocity gradient VelGrad of the last increment */
    SRX_F_TRANMAT = 5,    /* [9] crystal->sample matrix TranMat_xt2sa */
    SRX_F_RHO_S = 6,      /* [12] SSD density */
    SRX_F_RHO_M = 7,      /* [12] mobile density */
    SRX_F_RHO_F = 8,      /* [12] forest density of the last constitutive call */
    SRX_F_RHO_P = 9,      /* [12] parallel density of the last constitutive call */
    SRX_F_GAMDOT = 10,    /* [12] slip rates */
    SRX_F_GAM_ACUM = 11,  /* [12] accumulated slip */
    SRX_F_RHO_DIPOLE = 12, SRX_F_RHO_FOREST = 13, SRX_F_RHO_INPLANE = 14,   /* [12] */
    SRX_F_NR_ITERS = 15,  /* [1] NR iterations of the last outer iteration */
    SRX_F_FLUCT = 16,     /* [9] accumulated strain fluctuation (evolution.c:1198) */
    SRX_F_TRIAL_GAM_ACUM = 17, /* [12] */
    SRX_F_RHO_TOT = 18,   /* [1] rho_tot used by the last hardening update */
    SRX_F_PHASE = 19, SRX_F_GRAIN = 20,  /* [1] as doubles */
    SRX_F_RHO_G1 = 21, SRX_F_RHO_G2 = 22 /* [12] GND densities rho_g1 (screw), rho_g2 (edge-I) (evp.h:349);
                                            gnd option only, read-only */
};

/* Replaces GetNameList + SetupJob (io.c:96-164, init.c:728-1216) and
   OpenFiles (io.c:213-228).  `workdir` is the directory relative paths in the
   deck are resolved against and outputs are written to (the reference uses
   the CWD); NULL means the deck's directory.  `opt` may be NULL. */
int srx_open(const char *deck_path, const char *workdir, const srx_options *opt, srx_ctx **out);

/* One loading increment: evolution.c:726-1520 (one pass of the istep loop). */
int srx_step(srx_ctx *ctx, srx_step_report *rep);

/* All remaining increments (the istep loop of Evolution, evolution.c:667-1551). */
int srx_run(srx_ctx *ctx);

/* Copy a field to a caller buffer of n doubles (see the field ids). */
int srx_get_field(srx_ctx *ctx, int field_id, double *host, size_t n);

/* Overwrite a state field from a caller buffer (SIG, EPS, DISGRAD, TRANMAT,
   RHO_S, RHO_M, GAM_ACUM); used by tests to start from a given state. */
int srx_set_field(srx_ctx *ctx, int field_id, const double *host, size_t n);

/* Grid extents CellDim[3]. */
int srx_dims(srx_ctx *ctx, int *n3);

/* Replaces DestroyJob + CloseFiles (init.c:1218-1416, io.c:230-245). */
void srx_close(srx_ctx *ctx);

/* Result of one member of an ensemble run. */
typedef struct {
    int    index;            /* position in the decks array */
    int    rc;               /* 0, or the error code srx_open/srx_step returned */
    int    increments;       /* increments completed */
    long   voxels;           /* CellDim product */
    double setup_seconds;    /* srx_open */
    double run_seconds;      /* the member's increments */
    double run_phase_seconds;/* wall time of the whole run phase (every member's increments) */
    srx_step_report last;    /* report of the last completed increment */
} srx_ensemble_result;

/* Runs every increment of n independent decks (an ensemble of RVEs, each the equivalent of one
   `trial <deck>` job) on this GPU.  Every member is set up first (all stay resident on the GPU),
   then the increments run, then all are closed.  concurrency <= 0: lockstep -- the members are
   dealt into max(1, -concurrency) groups; the members of a group advance one increment at a time
   together, the Newton solve, its reduction and the outer-iteration control of all of them running
   as one launch each over the group, their FFT passes on up to 8 / groups streams; the groups run
   concurrently, each from its own host thread.  concurrency k > 0: k members at a time, each on
   its own stream from its own host thread, largest grids first.  Every mode gives each member the
   results of its deck run alone.
   Output files go to workdirs[i] (NULL: each deck's directory).  results[i] is filled for every
   deck; returns 0 or the first member's error code.  No counterpart in the reference, which runs
   one RVE per mpirun job. */
int srx_run_ensemble(const char *const *decks, const char *const *workdirs, int n, const srx_options *opt,
                     int concurrency, srx_ensemble_result *results);

/* ---- learned Newton start (no counterpart in the reference) ----------------------------
   A 24-64-64-6 ReLU perceptron, evaluated in bf16 on the matrix cores with FP32 accumulation
   inside the Newton kernel, sets the step length of each voxel's extrapolated start: the solve
   starts at lam + alpha d, d the previous Newton call's stress change, alpha = 1 + c . d / |d|
   with c the network's output, limited to 0 <= alpha <= 1 (c = 0 is the warm_start 1 start).
   The solve still iterates to its own
   tolerance (evolution.c:217-223) and is repeated from lam if it fails to converge, so the
   network only changes how many iterations it takes.
   Parameters travel as one FP32 blob of SRX_MLP_BLOB floats: W1[24][64], b1[64], W2[64][64],
   b2[64], W3[64][6], b3[6] (W[k][i]: input k -> unit i), feature mean[24] and inverse spread[24],
   target mean[6] and spread[6].  Features per voxel (raw): orientation quaternion (w >= 0),
   direction of d (crystal frame, B basis), the start stress lam in the crystal frame / 100 MPa,
   the forcing of the coming solve e / (|d| + |e|), e = 1e5 MPa x (eps_c - eps_c of the previous
   call) in the crystal frame (eps_c = sym(DisGrad), the strain the solve is driven by,
   evolution.c:192-201), log10(|d| / |lam|) and log10 of the mean rho_s.  Target: the coming
   call's change over |d|, minus d / |d|. */
#define SRX_MLP_BLOB 6210
#define SRX_MLP_NPARAM 6150

typedef struct {
    int epochs;              /* passes over the samples (default 20) */
    int batch;               /* samples per Adam step, a multiple of 64 (default 4096) */
    float lr;                /* initial learning rate (default 2e-3), cosine-decayed to lr_final */
    float lr_final;          /* (default 1e-4) */
    unsigned long long seed; /* shuffling and initial weights */
    int init;                /* 1: compute the normalisation from the samples and draw initial weights;
                                0: start from the blob passed in */
    int reserved[7];
} srx_mlp_train_opts;

void srx_mlp_default_train_options(srx_mlp_train_opts *opt);

/* Trains the network on the GPU (forward on the matrix cores, backward in FP32, Adam) on n samples
   x[n][24] (raw features) -> t[n][6] (targets); blob is read (init 0) and written;
   loss_hist[epochs] (may be NULL) receives the mean squared error of the normalised outputs per epoch. */
int srx_mlp_train(const float *x, const float *t, size_t n, const srx_mlp_train_opts *opt, float *blob,
                  float *loss_hist);

/* Gradient of mean((y' - t')^2) over n samples and 6 outputs (normalised outputs and targets)
   with respect to the SRX_MLP_NPARAM parameters, and that loss (one GPU forward/backward pass). */
int srx_mlp_grad(const float *blob, const float *x, const float *t, size_t n, float *grad, double *loss);

/* Network outputs y[n][6] (de-normalised) for raw features x[n][24]. */
int srx_mlp_forward(const float *blob, const float *x, size_t n, float *y);

/* Installs the network of a context opened with warm_start 2. */
int srx_mlp_set(srx_ctx *ctx, const float *blob);

/* Starts (stride > 0) or stops (stride 0) recording training samples: every Newton call after the
   first of an increment stores, for voxels 0, stride, 2 stride, ..., the features before the call
   and the target (the correction defined above) after it, until
   `capacity` samples are held.  Needs warm_start 1 or 2. */
int srx_mlp_record(srx_ctx *ctx, int stride, size_t capacity);

/* Copies up to max recorded samples with a non-zero previous change (x[.][24], t[.][6]; NULL
   pointers: count only) and returns their number in *n. */
int srx_mlp_samples(srx_ctx *ctx, float *x, float *t, size_t max, size_t *n);

/* FFT tiles chosen at srx_open (rows per block of the forward z, forward y, Green x, inverse y
   and inverse z passes; 0: default; -1: the Green x pass runs one wavefront per row, possible for
   n0 = 64, 128 and 256, where it is also the default) and their timed launch durations in
   microseconds. */
int srx_fft_tiles(srx_ctx *ctx, int *tiles, double *us);

/* Most radix stages of one FFT plan. */
#define SRX_FFT_MAX_STAGES 12

/* FFT radix plans (no counterpart in the reference, whose FFTW plans are its own).  A plan is the
   sequence of radices (2, 3, 4, 5 or 8) of the Stockham stages of one axis (0: x, 1: y, 2: z); the
   default is one per extent (radix 8 as often as possible, then 4, one 2, 3s and 5s).  With
   options.reserved[3] = 1, srx_open also times the candidates of srx_fft_plan_candidates on each
   axis whose passes take a plan (not x where the wavefront Green pass runs) and keeps the fastest,
   the same plan for every context of one grid shape and precision in the process.  Another radix
   order rounds differently: results then differ from the default plans within the FFT's rounding
   error, not bitwise.  The x pass of n0 = 64, 128 and 256 takes no plan (one wavefront per row). */
int srx_fft_plan(srx_ctx *ctx, int axis, int *radices, int *nstages);

/* Sets the plan of one axis: radices[0..nstages-1] from {2, 3, 4, 5, 8} whose product is the
   extent (else SRX_E_ARG), within the passes' LDS and butterfly limits (else SRX_E_GRID).  That
   axis's passes return to their default tiles, and the next increment captures its batches anew. */
int srx_set_fft_plan(srx_ctx *ctx, int axis, const int *radices, int nstages);

/* Candidate plans for extent n (host only, no GPU needed), the default first: every split of the
   power of two into radix-8 and radix-4 stages with at most one radix-2 stage, times the 3s and 5s,
   in the default order and reversed.  Writes up to max_plans rows of SRX_FFT_MAX_STAGES ints
   (radices, then 0s) and their total number in *nplans; SRX_E_SIZE when there are more than
   max_plans, SRX_E_GRID when n is not 2^a 3^b 5^c up to 1024. */
int srx_fft_plan_candidates(int n, int *radices, int max_plans, int *nplans);

/* Measurement helper (no counterpart in the reference): the HBM ceiling of a plain streaming
   copy on the current GPU, the reference point for the HBM-bound passes.  Copies `bytes` (at
   least 1 MiB; two buffers of that size are allocated and freed) with 16-byte accesses for each of
   12 launch shapes (1, 2 or 4 accesses in flight per lane; 4, 8, 16 or 32 blocks of 256 threads
   per compute unit), `reps` timed launches each after a warm-up, and returns the best launch's
   read + write GB/s, the mean over that shape's launches, and the shape. */
int srx_hbm_stream(size_t bytes, int reps, double *best_gbs, double *mean_gbs, int *best_unroll,
                   int *best_blocks_per_cu);

/* Text for an error code. */
const char *srx_strerror(int code);

void srx_default_options(srx_options *opt);

/* Per-kernel timing accumulated since srx_open, from HIP events recorded on
   the solver's stream around each launch (options.reserved[0] = 1: every kernel;
   2: the Newton kernel only, which keeps the event overhead off the other launches):
   names[i], total milliseconds ms[i] and launch count calls[i], i < *count. */
int srx_kernel_times(srx_ctx *ctx, const char **names, double *ms, int *calls, int *count);

/* ---- single-kernel entry points (for the parity tests) ---------------- */

/* Forward 3D FFT of `ncomp` real fields (each n0*n1*n2, x slowest) to the
   half spectrum along z: out[c][kx][ky][kz] complex (re,im), kz < n2/2+1.
   Unnormalised, sign -1, the DFT the reference gets from fftwnd_mpi
   (evolution.c:1163).  precision: SRX_PREC_FP64 or SRX_PREC_MIXED. */
int srx_fft_forward(const double *in, int ncomp, int n0, int n1, int n2, int precision, double *out);

/* Inverse of srx_fft_forward for real output: out = IFFT(in) * scale. */
int srx_fft_inverse(const double *in, int ncomp, int n0, int n1, int n2, int precision, double scale, double *out);

/* srx_fft_forward without any allocation.  No counterpart in the reference: it is for callers that
   cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  d_in (ncomp*n0*n1*n2 doubles) and d_out (ncomp*n0*n1*(n2/2+1) complex values as
   (re, im) doubles, srx_fft_forward's order) are device buffers of the caller; the two half spectra
   and the twiddle tables live in one static device array of SRX_NOALLOC_SCRATCH_BYTES per GPU, held
   until the process ends.  Needs 3 tables of n complex values (each rounded up to 256 bytes) plus
   2 x ncomp*n0*n1*(n2/2+1) complex values (16 B FP64, 8 B mixed); larger problems return
   SRX_E_SCRATCH.  Calls that overlap on one GPU share that array and race.  Synchronous: the
   result is in d_out on return.  Twiddles are computed on the device, so results agree with
   srx_fft_forward to rounding (not bitwise). */
#define SRX_NOALLOC_SCRATCH_BYTES (32u << 20)
int srx_fft_forward_noalloc(const double *d_in, int ncomp, int n0, int n1, int n2, int precision,
                            double *d_out);

#ifdef __cplusplus
}
#endif
#endif
