/*
 * srx_oracle.c -- CPU FP64 restatement of the SRX EVP-FFT increment loop.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the GPU product (see the header).
 * Every routine cites the reference file:line it restates.  Parity pinning of
 * each routine is done in tests/test_oracle_*.py (closed forms, numpy/scipy
 * library routines, finite differences, invariants).
 *
 * Storage follows the reference: AoS per voxel, x slowest, index
 * (x*N1 + y)*N2 + z (evp.h:60).  Tensors are 3x3 row-major.
 */
#include "srx_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* evp.h:84-94 */
#define O_PI 3.14159265359
#define O_TWO_PI 6.28318530718
#define O_TRIAL_FREQ_DRX 1.0E10
#define O_EV_IN_J 1.60217733e-19
#define O_J_IN_EV (1.0 / O_EV_IN_J)
#define O_MY_J_IN_EV 0.62415063631
#define O_BOLZ 1.380658e-23
#define O_EPS_DIPOLE 1.0E-4     /* evp.h:21 EPSILON */
#define O_LU_ROWMIN 1E-5        /* evp.h:94 EPSILON_FFT, kinematics.c:306 */
#define O_LU_TINY 1.0E-20       /* kinematics.c:3 */
#define NSYS 12
#define NPH 2

typedef double t33[3][3];

/* ------------------------------------------------------------------ */
/* B basis (kinematics.c:14-93, chg_basis opt 0 literals at 22-24)     */
/* ------------------------------------------------------------------ */
static double Bb[6][3][3];
static int Bb_ready = 0;

static void make_basis(void)
{
    const double r2 = 0.70710678118654744, r3 = 0.57735026918962584, r6 = 0.40824829046386304;
    memset(Bb, 0, sizeof(Bb));
    Bb[0][0][0] = -r2;      Bb[0][1][1] = r2;                    /* (-e11+e22)/sqrt2 */
    Bb[1][0][0] = -r6;      Bb[1][1][1] = -r6; Bb[1][2][2] = 2. * r6;
    Bb[2][1][2] = r2;       Bb[2][2][1] = r2;
    Bb[3][0][2] = r2;       Bb[3][2][0] = r2;
    Bb[4][0][1] = r2;       Bb[4][1][0] = r2;
    Bb[5][0][0] = r3;       Bb[5][1][1] = r3; Bb[5][2][2] = r3;
    Bb_ready = 1;
}

/* chg_basis / chg_basis5 (kinematics.c:14-174); dim = 6 or 5 */
void orc_chg_basis(double *v6, double *t2, double *c2, double *t4, int opt, int dim)
{
    int i, j, a, b, c, d;
    if (!Bb_ready) make_basis();
    if (opt == 1) {
        for (a = 0; a < 3; a++) for (b = 0; b < 3; b++) {
            double s = 0.0;
            for (i = 0; i < dim; i++) s += v6[i] * Bb[i][a][b];
            t2[a * 3 + b] = s;
        }
    } else if (opt == 2) {
        for (i = 0; i < dim; i++) {
            double s = 0.0;
            for (a = 0; a < 3; a++) for (b = 0; b < 3; b++) s += t2[a * 3 + b] * Bb[i][a][b];
            v6[i] = s;
        }
    } else if (opt == 3) {
        for (a = 0; a < 3; a++) for (b = 0; b < 3; b++) for (c = 0; c < 3; c++) for (d = 0; d < 3; d++) {
            double s = 0.0;
            for (i = 0; i < dim; i++) for (j = 0; j < dim; j++) s += c2[i * 6 + j] * Bb[i][a][b] * Bb[j][c][d];
            t4[((a * 3 + b) * 3 + c) * 3 + d] = s;
        }
    } else if (opt == 4) {
        for (i = 0; i < dim; i++) for (j = 0; j < dim; j++) {
            double s = 0.0;
            for (a = 0; a < 3; a++) for (b = 0; b < 3; b++) for (c = 0; c < 3; c++) for (d = 0; d < 3; d++)
                s += t4[((a * 3 + b) * 3 + c) * 3 + d] * Bb[i][a][b] * Bb[j][c][d];
            c2[i * 6 + j] = s;
        }
    }
}

static void v6_to_t(const double *v, int dim, t33 t) { orc_chg_basis((double *)v, &t[0][0], 0, 0, 1, dim); }
static void t_to_v6(const t33 t, double *v, int dim) { orc_chg_basis(v, (double *)&t[0][0], 0, 0, 2, dim); }

/* ------------------------------------------------------------------ */
/* Crout LU with implicit partial pivoting; 6x6 inverse                 */
/* (kinematics.c:282-428).  Returns 35 on a row whose largest entry is  */
/* below 1e-5, the reference's exit(35).                                */
/* ------------------------------------------------------------------ */
static int lu_dcmp6(double a[6][6], int *indx)
{
    double vv[6];
    int i, j, k, imax = 0;
    for (i = 0; i < 6; i++) {
        double big = 0.0;
        for (j = 0; j < 6; j++) if (fabs(a[i][j]) > big) big = fabs(a[i][j]);
        if (fabs(big) < O_LU_ROWMIN) return 35;
        vv[i] = 1.0 / big;
    }
    for (j = 0; j < 6; j++) {
        double big;
        for (i = 0; i < j; i++) {
            double sum = a[i][j];
            for (k = 0; k < i; k++) sum -= a[i][k] * a[k][j];
            a[i][j] = sum;
        }
        big = 0.0;
        for (i = j; i < 6; i++) {
            double sum = a[i][j], dum;
            for (k = 0; k < j; k++) sum -= a[i][k] * a[k][j];
            a[i][j] = sum;
            if ((dum = vv[i] * fabs(sum)) >= big) { big = dum; imax = i; }
        }
        if (j != imax) {
            for (k = 0; k < 6; k++) { double d = a[imax][k]; a[imax][k] = a[j][k]; a[j][k] = d; }
            vv[imax] = vv[j];
        }
        indx[j] = imax;
        if (fabs(a[j][j]) < 1E-5) a[j][j] = O_LU_TINY;
        if (j != 5) {
            double dum = 1.0 / a[j][j];
            for (i = j + 1; i < 6; i++) a[i][j] *= dum;
        }
    }
    return 0;
}

static void lu_bksb6(double a[6][6], const int *indx, double *b)
{
    int i, j, ii = -1;
    for (i = 0; i < 6; i++) {
        int ip = indx[i];
        double sum = b[ip];
        b[ip] = b[i];
        if (ii >= 0) for (j = ii; j < i; j++) sum -= a[i][j] * b[j];
        else if (fabs(sum) > 1E-5) ii = i;
        b[i] = sum;
    }
    for (i = 5; i >= 0; i--) {
        double sum = b[i];
        for (j = i + 1; j < 6; j++) sum -= a[i][j] * b[j];
        b[i] = sum / a[i][i];
    }
}

static int lu_inv66(double c[6][6])
{
    int indx[6], i, j, rc;
    double y[6][6], col[6];
    rc = lu_dcmp6(c, indx);
    if (rc) return rc;
    for (j = 0; j < 6; j++) {
        for (i = 0; i < 6; i++) col[i] = 0.0;
        col[j] = 1.0;
        lu_bksb6(c, indx, col);
        for (i = 0; i < 6; i++) y[i][j] = col[i];
    }
    memcpy(c, y, sizeof(y));
    return 0;
}

void orc_lu_inv66(double *c66) { lu_inv66((double (*)[6])c66); }

/* EulerToTransMatrix (kinematics.c:187-240), Bunge angles, sample->crystal */
void orc_euler_to_matrix(double *p1, double *pp, double *p2, double *a, int opt)
{
    if (opt == 1) {
        double ph = acos(a[8]), t1, t2;
        if (fabs(a[8]) > 0.9999) {
            t2 = 0.;
            t1 = atan2(a[1], a[0]);
        } else {
            double sph = sin(ph);
            t1 = atan2(a[6] / sph, -1.0 * a[7] / sph);
            t2 = atan2(a[2] / sph, a[5] / sph);
        }
        *pp = ph * (180. / O_PI);
        *p1 = t1 * (180. / O_PI);
        *p2 = t2 * (180. / O_PI);
    } else {
        double ct1 = cos(*p1), st1 = sin(*p1), cph = cos(*pp), sph = sin(*pp), ct2 = cos(*p2), st2 = sin(*p2);
        a[0] = ct1 * ct2 - st1 * st2 * cph;
        a[1] = st1 * ct2 + ct1 * st2 * cph;
        a[2] = st2 * sph;
        a[3] = -1.0 * ct1 * st2 - st1 * ct2 * cph;
        a[4] = -1.0 * st1 * st2 + ct1 * ct2 * cph;
        a[5] = ct2 * sph;
        a[6] = st1 * sph;
        a[7] = -1.0 * ct1 * sph;
        a[8] = cph;
    }
}

/* Von Mises of a non-symmetric tensor (kinematics.c:465-483) */
static double von_mises(const t33 t)
{
    double tr = t[0][0] + t[1][1] + t[2][2], vm = 0.0;
    int i, j;
    for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) {
        double d = (t[i][j] + t[j][i]) / 2. - (double)(i == j) * tr / 3.0;
        vm += d * d;
    }
    return sqrt(2. / 3. * vm);
}

/* ------------------------------------------------------------------ */
/* FFT: 3D complex, in place, unnormalised, FFTW sign convention        */
/* ------------------------------------------------------------------ */
static void fft1d(double *re, double *im, int n, int sign, double *wr, double *wi)
{
    /* wr/wi: table of cos/sin(2*pi*m/n) for m < n; work in re/im (length n) */
    int i, j, k, len;
    if (n <= 1) return;
    if ((n & (n - 1)) == 0) {
        for (i = 1, j = 0; i < n; i++) {
            int bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) {
                double t = re[i]; re[i] = re[j]; re[j] = t;
                t = im[i]; im[i] = im[j]; im[j] = t;
            }
        }
        for (len = 2; len <= n; len <<= 1) {
            int step = n / len;
            for (i = 0; i < n; i += len) {
                for (k = 0; k < len / 2; k++) {
                    double c = wr[k * step], s = sign * wi[k * step];
                    int p = i + k, q = i + k + len / 2;
                    double xr = re[q] * c - im[q] * s, xi = re[q] * s + im[q] * c;
                    re[q] = re[p] - xr; im[q] = im[p] - xi;
                    re[p] += xr; im[p] += xi;
                }
            }
        }
    } else {
        double *tr = (double *)malloc(sizeof(double) * 2 * n), *ti = tr + n;
        for (k = 0; k < n; k++) {
            double sr = 0.0, si = 0.0;
            for (j = 0; j < n; j++) {
                int m = (int)(((long)k * j) % n);
                double c = wr[m], s = sign * wi[m];
                sr += re[j] * c - im[j] * s;
                si += re[j] * s + im[j] * c;
            }
            tr[k] = sr; ti[k] = si;
        }
        memcpy(re, tr, sizeof(double) * n);
        memcpy(im, ti, sizeof(double) * n);
        free(tr);
    }
}

void orc_fft3d(double *data, int n0, int n1, int n2, int sign, int nthreads)
{
    int ax;
    int nn[3] = {n0, n1, n2};
    (void)nthreads;
    for (ax = 0; ax < 3; ax++) {
        int n = nn[ax];
        long stride = ax == 0 ? (long)n1 * n2 : (ax == 1 ? n2 : 1);
        long npen = (long)n0 * n1 * n2 / n;
        double *wr, *wi;
        long p;
        if (n <= 1) continue;
        wr = (double *)malloc(sizeof(double) * 2 * n); wi = wr + n;
        for (p = 0; p < n; p++) { wr[p] = cos(2.0 * M_PI * (double)p / n); wi[p] = sin(2.0 * M_PI * (double)p / n); }
        #pragma omp parallel for schedule(static)
        for (p = 0; p < npen; p++) {
            double *re = (double *)malloc(sizeof(double) * 2 * n), *im = re + n;
            long base, i;
            if (ax == 0) base = p;                                  /* p = y*n2+z */
            else if (ax == 1) base = (p / n2) * (long)n1 * n2 + p % n2;  /* p = x*n2+z */
            else base = p * n2;                                     /* p = x*n1+y */
            for (i = 0; i < n; i++) { re[i] = data[2 * (base + i * stride)]; im[i] = data[2 * (base + i * stride) + 1]; }
            fft1d(re, im, n, sign, wr, wi);
            for (i = 0; i < n; i++) { data[2 * (base + i * stride)] = re[i]; data[2 * (base + i * stride) + 1] = im[i]; }
            free(re);
        }
        free(wr);
    }
}

/* ------------------------------------------------------------------ */
/* Deck (io.c:25-164), slip-mode file (init.c:52-231), mesh (253-301)   */
/* ------------------------------------------------------------------ */
typedef struct {
    int CellDim[3], N_phases, Type_phases[2], N_steps, Update_Flag, Hard_Flag, Tex_Flag, IterMax;
    int PrintControl[2], VelGrad_BC_Flag[6], Stress_BC_Flag[6], ElastBC, nucleus_radius, CellDim_pf[3];
    int RandSeed, Nucl_Static_Flag;
    double TimeStep, Err, VelGrad_BC[6], Stress_BC[6], ElastConst_PhaseI[4], ElastConst_PhaseII[4];
    char Slip_PhaseI[100], Slip_PhaseII[100], initial_ms[100];
    double a0[2], L0, GND_ScaleRatio, T__K, Q_slip[2], Q_bulk[2];
    double C_1[2], C_2[2], C_3[2], C_4[2], C_5[2], C_6[2], C_7[2], C_8[2];
    double Selfinter0[2], Coplanar0[2], CrossSlip0[2], GlissileJunction0[2], HirthLock0[2], LomerCottrellLock0[2];
    double rho_SSD_initial, D_0;
    double E_gb, M_gb, Scale_Fdeform, k_c, alpha, zeta_kc, age_drx, ssdScaler_DRX, gndScaler_DRX, sigScaler_DRX;
    double tau_DRX, DeltaQ_DRX, DeltaTau_DRX, pf_length_scale, pf_time_step;
} orc_deck;

typedef struct { const char *name; void *ptr; int type; int len; } key_t_;

static int parse_deck(const char *path, orc_deck *d)
{
#define KI(x, n) {#x, (void *)&d->x, 0, n}
#define KR(x, n) {#x, (void *)&d->x, 1, n}
#define KS(x) {#x, (void *)d->x, 2, 1}
    key_t_ keys[] = {
        KI(CellDim, 3), KI(N_phases, 1), KI(Type_phases, 2), KR(TimeStep, 1), KI(N_steps, 1),
        KI(Update_Flag, 1), KI(Hard_Flag, 1), KI(Tex_Flag, 1), KI(IterMax, 1), KI(PrintControl, 2),
        KR(Err, 1), KI(VelGrad_BC_Flag, 6), KR(VelGrad_BC, 6), KI(Stress_BC_Flag, 6), KR(Stress_BC, 6),
        KI(ElastBC, 1), KR(ElastConst_PhaseI, 4), KR(ElastConst_PhaseII, 4), KS(Slip_PhaseI), KS(Slip_PhaseII),
        KS(initial_ms), KR(a0, 2), KR(L0, 1), KR(GND_ScaleRatio, 1), KR(T__K, 1), KR(Q_slip, 2), KR(Q_bulk, 2),
        KR(C_1, 2), KR(C_2, 2), KR(C_3, 2), KR(C_4, 2), KR(C_5, 2), KR(C_6, 2), KR(C_7, 2), KR(C_8, 2),
        KR(Selfinter0, 2), KR(Coplanar0, 2), KR(CrossSlip0, 2), KR(GlissileJunction0, 2), KR(HirthLock0, 2),
        KR(LomerCottrellLock0, 2), KR(rho_SSD_initial, 1), KR(D_0, 1), KI(nucleus_radius, 1),
        KI(CellDim_pf, 3), KI(RandSeed, 1), KR(E_gb, 1), KR(M_gb, 1), KR(Scale_Fdeform, 1), KR(k_c, 1),
        KI(Nucl_Static_Flag, 1), KR(alpha, 1), KR(zeta_kc, 1), KR(age_drx, 1), KR(ssdScaler_DRX, 1),
        KR(gndScaler_DRX, 1), KR(sigScaler_DRX, 1), KR(tau_DRX, 1), KR(DeltaQ_DRX, 1), KR(DeltaTau_DRX, 1),
        KR(pf_length_scale, 1), KR(pf_time_step, 1),
    };
#undef KI
#undef KR
#undef KS
    int nk = (int)(sizeof(keys) / sizeof(keys[0])), k, j;
    int seen[80] = {0};
    char buff[200], *tok;
    FILE *fp = fopen(path, "r");
    if (!fp) return 2001;
    for (;;) {
        if (!fgets(buff, 200, fp)) break;
        if (feof(fp)) break;                       /* io.c:112-113: a last line without newline is dropped */
        if (buff[0] == '#' || buff[0] == '\n') continue;
        tok = strtok(buff, " \t\n");
        if (!tok) break;                           /* io.c:116-117 */
        for (k = 0; k < nk; k++) {
            if (strcmp(tok, keys[k].name) != 0) continue;
            if (!seen[k]) {
                seen[k] = 1;
                for (j = 0; j < keys[k].len; j++) {
                    tok = strtok(NULL, ", \t\n");
                    if (!tok) break;
                    if (keys[k].type == 0) ((int *)keys[k].ptr)[j] = (int)atol(tok);
                    else if (keys[k].type == 1) ((double *)keys[k].ptr)[j] = atof(tok);
                    else { strncpy((char *)keys[k].ptr, tok, 99); ((char *)keys[k].ptr)[99] = 0; }
                }
            }
            break;
        }
    }
    fclose(fp);
    return 0;
}

typedef struct {
    int crystalline, nsys;
    double c11, c12, c44, nu, shear_g;
    double C4[81];          /* crystal frame Cijkl */
    double bn[NSYS][3], bb[NSYS][3], schm_xt[NSYS][5];
    double bburg, tau0xf;
    double c1, c2, c3, c4, c5, c6, c7, c8, q_slip, q_bulk;
    double chi[NSYS][NSYS], chi1[NSYS][NSYS];
    /* projections of the GND densities (init.c:1118-1141) and bp = bb x bn (init.c:200-202) */
    double cos_g1[NSYS][NSYS], sin_g1[NSYS][NSYS], cos_g2[NSYS][NSYS], sin_g2[NSYS][NSYS];
    double cos_g3[NSYS][NSYS], sin_g3[NSYS][NSYS], bp[NSYS][3];
} orc_phase;

static void resolve(char *out, const char *dir, const char *name)
{
    if (name[0] == '/' || !dir || !dir[0]) snprintf(out, 1024, "%s", name);
    else snprintf(out, 1024, "%s/%s", dir, name);
}

/* PlasticInit (init.c:52-231) */
static int read_sx(const char *path, orc_phase *ph)
{
    char buf[80];
    int i, j, nmodes_max = 0, nmodes = 0, imode[16], mx, nsmx;
    double nrsx, gamd0x, twshx, tau0xf, tau0xb, tau1x, thet0, thet1, hs, hl;
    int isectwx;
    FILE *fp = fopen(path, "r");
    if (!fp) return 2001;
#define GL() do { if (!fgets(buf, 80, fp)) buf[0] = 0; } while (0)
    GL(); GL(); GL(); GL(); GL(); GL(); GL();
    GL(); sscanf(buf, "%d", &nmodes_max);
    GL(); GL();
    GL(); sscanf(buf, "%d", &nmodes);
    if (nmodes > 16) nmodes = 16;
    GL(); GL();
    for (i = 0; i < nmodes; i++) { GL(); sscanf(buf, "%d", &imode[i]); }
    GL();
    ph->nsys = 0;
    for (i = 0; i < nmodes_max; i++) {
        int match = 0;
        GL(); mx = -1; sscanf(buf, "%*s %*s %d %*s", &mx);
        for (j = 0; j < nmodes; j++) if (mx == imode[j]) { match = 1; break; }
        GL(); GL(); nsmx = 0; sscanf(buf, "%d %lf %lf %lf %d", &nsmx, &nrsx, &gamd0x, &twshx, &isectwx);
        GL(); GL(); sscanf(buf, "%lf %lf %lf %lf %lf", &tau0xf, &tau0xb, &tau1x, &thet0, &thet1);
        GL(); GL(); sscanf(buf, "%lf %lf", &hs, &hl);
        GL();
        if (match) {
            if (nsmx > NSYS) { fclose(fp); return 2004; }
            ph->nsys = nsmx;
            if (thet0 < thet1) { fclose(fp); return 230; }
            ph->tau0xf = tau0xf;
            for (j = 0; j < nsmx; j++) {
                GL();
                sscanf(buf, "%lf %lf %lf %lf %lf %lf", &ph->bn[j][0], &ph->bn[j][1], &ph->bn[j][2],
                       &ph->bb[j][0], &ph->bb[j][1], &ph->bb[j][2]);
            }
        } else {
            for (j = 0; j < nsmx; j++) GL();
        }
    }
#undef GL
    fclose(fp);
    for (j = 0; j < ph->nsys; j++) {
        double nrm = sqrt(pow(ph->bn[j][0], 2.0) + pow(ph->bn[j][1], 2.0) + pow(ph->bn[j][2], 2.0)), dot;
        t33 s;
        int a, b;
        for (a = 0; a < 3; a++) ph->bn[j][a] /= nrm;
        nrm = sqrt(pow(ph->bb[j][0], 2.0) + pow(ph->bb[j][1], 2.0) + pow(ph->bb[j][2], 2.0));
        for (a = 0; a < 3; a++) ph->bb[j][a] /= nrm;
        dot = ph->bn[j][0] * ph->bb[j][0] + ph->bn[j][1] * ph->bb[j][1] + ph->bn[j][2] * ph->bb[j][2];
        if (dot > 1E-3) return 7;
        for (a = 0; a < 3; a++) for (b = 0; b < 3; b++) s[a][b] = (ph->bn[j][a] * ph->bb[j][b] + ph->bn[j][b] * ph->bb[j][a]) / 2.0;
        t_to_v6(s, ph->schm_xt[j], 5);
    }
    return 0;
}

/* Arsenlis-Parks interaction table (evolution.c:3-112 / init.c:594-703):
   S self, C coplanar, X cross slip, G glissile, H Hirth, L Lomer; upper triangle by rows. */
static const char *ISM_ROWS[12] = {
    "SCCHGLXGGHLG", "SCGXGGHLLHG", "SLGHGLHGGX", "SCCHLGXGG", "SCLHGGHL",
    "SGGXGLH", "SCCHGL", "SCGXG", "SLGH", "SCC", "SC", "S"};

static void interaction_matrix(double self, double cop, double cross, double gli, double hirth, double lomer,
                               double chi[NSYS][NSYS])
{
    int i, j;
    for (i = 0; i < 12; i++) {
        for (j = i; j < 12; j++) {
            char c = ISM_ROWS[i][j - i];
            double v = c == 'S' ? self : c == 'C' ? cop : c == 'X' ? cross : c == 'G' ? gli : c == 'H' ? hirth : lomer;
            chi[i][j] = v;
            chi[j][i] = v;
        }
    }
}

/* ------------------------------------------------------------------ */
/* State                                                               */
/* ------------------------------------------------------------------ */
struct orc_state {
    orc_options opt;
    orc_deck dk;
    int n[3];
    long nxyz;
    double wgt, wgt_ph1;
    orc_phase ph[NPH];
    double kT;
    double C066[6][6], S066[6][6], C0[81], S0[81];
    double Udot[3][3], Udot_s[3][3], Udot_a[3][3], Scauchy[3][3];
    double DisGradAvg[3][3], dDisGradAvg[3][3], dDisGradAvg_acum[3][3], DisGradAvg_t[3][3];
    double SigAvg[3][3], SigAvg1[3][3];
    double d_vm, e_vm, s_vm, s_vm1, Err_e, Err_s, TimeTot, err, d_vm0, e_vm0;
    int creep, itermax, nsteps, istep;
    int *phase, *grain;
    double *tran;         /* TranMat_xt2sa 9 */
    double *cgr;          /* C_gr 36 */
    double *schm;         /* Schm_gr 60 */
    double *sig, *eps, *edot, *disgrad, *velgrad, *fluct;   /* 9 each */
    double *rho_s, *rho_m, *rho_F, *rho_P, *trial_rho_s, *gamdot, *gam_acum, *trial_gam_acum;
    double *rho_dipole, *rho_forest, *rho_inplane;          /* 12 each */
    double *rho_tot;
    double *rho_g1, *rho_g2, *rho_g3;   /* GND densities, 12 each (gnd option; rho_g3 stays 0 in first order) */
    double *gadot_grad;                 /* 36: gradient of gamdot per system (constitutive.c:882-1115) */
    unsigned char *gbc;                 /* GB_checkX/Y/Z[0..1] as bits 0..5 (constitutive.c:731-880) */
    float *dsig;      /* warm_start 1: last NR call's stress change, B basis (float, as the GPU keeps it) */
    int ws_use;       /* warm_start 1: this sweep starts from the extrapolated stress */
    double *nrit;
    double *work;         /* 2*nxyz complex scratch */
    double *ksig;         /* 9 comps x nxyz complex */
    double *nvec;         /* 3 per voxel: unit k vectors (init.c:776-797) */
    double *conv_s, *conv_e;
    char workdir[1024];
    FILE *fvm, *fdd, *ferr, *fdrx;
    orc_report last;
};

void orc_default_options(orc_options *o)
{
    memset(o, 0, sizeof(*o));
    o->svm_mode = 0;
    o->svm_stray = 0.0;
    o->write_files = 0;
}

int orc_dims(orc_state *s, int *n3) { n3[0] = s->n[0]; n3[1] = s->n[1]; n3[2] = s->n[2]; return 0; }

static int is_xtal(const orc_state *s, int jph) { return jph >= 1 && jph <= NPH && s->ph[jph - 1].crystalline; }

/* ElastStiffnessMatrix (init.c:10-50) */
static void cubic_c4(double c11, double c12, double c44, double *C)
{
    int i, j, k, l;
    for (i = 0; i < 81; i++) C[i] = 0.0;
    for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) for (k = 0; k < 3; k++) for (l = 0; l < 3; l++) {
        double v = 0.0;
        if (i == j && k == l) v = (i == k) ? c11 : c12;
        else if ((i == k && j == l) || (i == l && j == k)) v = c44;
        C[((i * 3 + j) * 3 + k) * 3 + l] = v;
    }
}

static void vn_to_full(const double *v, double m[3][3])   /* VoigtToFull kinematics.c:5-12 */
{
    m[0][0] = v[0]; m[0][1] = v[5]; m[0][2] = v[4];
    m[1][0] = v[5]; m[1][1] = v[1]; m[1][2] = v[3];
    m[2][0] = v[4]; m[2][1] = v[3]; m[2][2] = v[2];
}

static void update_schmid(orc_state *s)   /* kinematics.c:430-463 */
{
    long p;
    #pragma omp parallel for schedule(static)
    for (p = 0; p < s->nxyz; p++) {
        int jph = s->phase[p], i, a, b, c, d;
        const double *T = s->tran + 9 * p;
        if (!is_xtal(s, jph)) continue;
        for (i = 0; i < s->ph[jph - 1].nsys; i++) {
            t33 m, r;
            v6_to_t(s->ph[jph - 1].schm_xt[i], 5, m);
            for (a = 0; a < 3; a++) for (b = 0; b < 3; b++) {
                double acc = 0.0;
                for (c = 0; c < 3; c++) for (d = 0; d < 3; d++) acc += T[a * 3 + c] * T[b * 3 + d] * m[c][d];
                r[a][b] = acc;
            }
            t_to_v6(r, s->schm + 60 * p + 5 * i, 5);
        }
    }
}

/* Gradient_ExchangeGrainID (constitutive.c:731-880), one rank: periodic neighbours.  Bits 0..5 are
   GB_checkX[0], GB_checkX[1], GB_checkY[0], GB_checkY[1], GB_checkZ[0], GB_checkZ[1] (1: the +/-
   neighbour is in the same grain).  GB_checkZ[1] compares with the y- neighbour, as the reference
   does (823, 842, 860). */
static void gradient_exchange_grain_id(orc_state *s)
{
    const int n0 = s->n[0], n1 = s->n[1], n2 = s->n[2];
    long p;
    for (p = 0; p < s->nxyz; p++) {
        int x = (int)(p / ((long)n1 * n2)), y = (int)((p / n2) % n1), z = (int)(p % n2);
        int xp = x + 1 >= n0 ? 0 : x + 1, xm = x - 1 < 0 ? n0 - 1 : x - 1;
        int yp = y + 1 >= n1 ? 0 : y + 1, ym = y - 1 < 0 ? n1 - 1 : y - 1;
        int zp = z + 1 >= n2 ? 0 : z + 1;
        int g = s->grain[p];
        unsigned char b = 0;
        b |= (g == s->grain[((long)xp * n1 + y) * n2 + z]) << 0;
        b |= (g == s->grain[((long)xm * n1 + y) * n2 + z]) << 1;
        b |= (g == s->grain[((long)x * n1 + yp) * n2 + z]) << 2;
        b |= (g == s->grain[((long)x * n1 + ym) * n2 + z]) << 3;
        b |= (g == s->grain[((long)x * n1 + y) * n2 + zp]) << 4;
        b |= (g == s->grain[((long)x * n1 + ym) * n2 + z]) << 5;
        s->gbc[p] = b;
    }
}

/* one-sided or central difference inside the grain (constitutive.c:970-1010): c0, c1 = same grain on
   the + / - side */
static double gb_diff(int c0, int c1, double v, double vp, double vm)
{
    if (!c0 && c1) return v - vm;
    if (c0 && !c1) return vp - v;
    if (c0 && c1) return (vp - vm) / 2.0;
    return 0.0;
}

/* Gradient_ShearRate (constitutive.c:882-1115), one rank */
static void gradient_shear_rate(orc_state *s)
{
    const int n0 = s->n[0], n1 = s->n[1], n2 = s->n[2];
    long p;
    #pragma omp parallel for schedule(static)
    for (p = 0; p < s->nxyz; p++) {
        int x = (int)(p / ((long)n1 * n2)), y = (int)((p / n2) % n1), z = (int)(p % n2), i;
        int xp = x + 1 >= n0 ? 0 : x + 1, xm = x - 1 < 0 ? n0 - 1 : x - 1;
        int yp = y + 1 >= n1 ? 0 : y + 1, ym = y - 1 < 0 ? n1 - 1 : y - 1;
        int zp = z + 1 >= n2 ? 0 : z + 1, zm = z - 1 < 0 ? n2 - 1 : z - 1;
        long pxp = ((long)xp * n1 + y) * n2 + z, pxm = ((long)xm * n1 + y) * n2 + z;
        long pyp = ((long)x * n1 + yp) * n2 + z, pym = ((long)x * n1 + ym) * n2 + z;
        long pzp = ((long)x * n1 + y) * n2 + zp, pzm = ((long)x * n1 + y) * n2 + zm;
        const unsigned char b = s->gbc[p];
        for (i = 0; i < NSYS; i++) {
            const double *gd = s->gamdot;
            double *g = s->gadot_grad + 36 * p + 3 * i;
            g[0] = gb_diff(b & 1, (b >> 1) & 1, gd[12 * p + i], gd[12 * pxp + i], gd[12 * pxm + i]);
            g[1] = gb_diff((b >> 2) & 1, (b >> 3) & 1, gd[12 * p + i], gd[12 * pyp + i], gd[12 * pym + i]);
            g[2] = gb_diff((b >> 4) & 1, (b >> 5) & 1, gd[12 * p + i], gd[12 * pzp + i], gd[12 * pzm + i]);
        }
    }
}

static int setup(orc_state *s)
{
    orc_deck *d = &s->dk;
    long p;
    int i, j, ph, rc;
    double c066[6][6];
    char path[1024];
    s->n[0] = d->CellDim[0]; s->n[1] = d->CellDim[1]; s->n[2] = d->CellDim[2];
    if (s->n[0] < 1 || s->n[1] < 1 || s->n[2] < 1) return 2002;
    s->nxyz = (long)s->n[0] * s->n[1] * s->n[2];
    s->wgt = 1. / s->nxyz;
    s->kT = (d->T__K * O_BOLZ * O_J_IN_EV);
    if (!Bb_ready) make_basis();

    /* unit k vectors, init.c:776-797 */
    s->nvec = (double *)malloc(sizeof(double) * 3 * s->nxyz);
    for (p = 0; p < s->nxyz; p++) {
        int x = (int)(p / ((long)s->n[1] * s->n[2])), y = (int)((p / s->n[2]) % s->n[1]), z = (int)(p % s->n[2]);
        double t[3], g2;
        int m;
        t[0] = (double)x * (double)(1.0 / s->n[0]);
        t[1] = (double)y * (double)(1.0 / s->n[1]);
        t[2] = (double)z * (double)(1.0 / s->n[2]);
        for (m = 0; m < 3; m++) { t[m] -= round(t[m]); t[m] *= O_TWO_PI; }
        g2 = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
        if (x + y + z == 0) { g2 = 1.0; t[0] = t[1] = t[2] = 0.0; }
        for (m = 0; m < 3; m++) s->nvec[3 * p + m] = t[m] * (1. / g2);
    }

    /* phases, init.c:815-856 */
    memset(s->ph, 0, sizeof(s->ph));
    for (ph = 0; ph < NPH; ph++) s->ph[ph].crystalline = (ph < d->N_phases) && (d->Type_phases[ph] == 0);
    if (s->ph[0].crystalline) {
        /* DD_CU override (init.c:822-826) */
        d->ElastConst_PhaseI[0] = (1.818 - 0.0004045 * d->T__K) * 100 * 1000;
        d->ElastConst_PhaseI[1] = (1.285 - 0.0002 * d->T__K) * 100 * 1000;
        d->ElastConst_PhaseI[2] = (0.8396 - 0.0002711 * d->T__K) * 100 * 1000;
    }
    for (ph = 0; ph < NPH; ph++) {
        orc_phase *P = &s->ph[ph];
        double *E = ph == 0 ? d->ElastConst_PhaseI : d->ElastConst_PhaseII;
        if (!P->crystalline) continue;
        P->c11 = E[0]; P->c12 = E[1]; P->c44 = E[2]; P->nu = E[3];
        P->shear_g = 0.1 * (E[0] - E[1] + 3 * E[2]) + (0.5) * (5 * (E[0] - E[1]) * E[2]) / (4 * E[2] + 3 * (E[0] - E[1]));
        cubic_c4(E[0], E[1], E[2], P->C4);
        resolve(path, s->workdir, ph == 0 ? d->Slip_PhaseI : d->Slip_PhaseII);
        rc = read_sx(path, P);
        if (rc) return rc;
        P->bburg = d->a0[ph] / sqrt(2.0);
        P->c1 = d->C_1[ph]; P->c2 = d->C_2[ph]; P->c3 = d->C_3[ph]; P->c4 = d->C_4[ph];
        P->c5 = d->C_5[ph]; P->c6 = d->C_6[ph]; P->c7 = d->C_7[ph]; P->c8 = d->C_8[ph];
        P->q_slip = d->Q_slip[ph] * O_MY_J_IN_EV / s->kT;
        P->q_bulk = d->Q_bulk[ph] * O_MY_J_IN_EV / s->kT;
    }
    for (ph = 0; ph < NPH; ph++) {   /* init.c:1100-1116 */
        interaction_matrix(d->Selfinter0[ph], d->Coplanar0[ph], 0.0, 0.0, 0.0, 0.0, s->ph[ph].chi);
        interaction_matrix(0.0, 0.0, d->CrossSlip0[ph], d->GlissileJunction0[ph], d->HirthLock0[ph],
                           d->LomerCottrellLock0[ph], s->ph[ph].chi1);
    }
    for (ph = 0; ph < NPH; ph++) {   /* init.c:1118-1141 (sin and cos of the GND projections), 200-202 (bp) */
        orc_phase *P = &s->ph[ph];
        int i, j;
        for (j = 0; j < P->nsys; j++) {
            P->bp[j][0] = P->bb[j][1] * P->bn[j][2] - P->bb[j][2] * P->bn[j][1];
            P->bp[j][1] = P->bb[j][2] * P->bn[j][0] - P->bb[j][0] * P->bn[j][2];
            P->bp[j][2] = P->bb[j][0] * P->bn[j][1] - P->bb[j][1] * P->bn[j][0];
        }
        for (i = 0; i < P->nsys; i++) for (j = 0; j < P->nsys; j++) {
            double sv[3], cs;
            sv[0] = P->bn[j][1] * P->bb[j][2] - P->bn[j][2] * P->bb[j][1];
            sv[1] = P->bn[j][2] * P->bb[j][0] - P->bn[j][0] * P->bb[j][2];
            sv[2] = P->bn[j][0] * P->bb[j][1] - P->bn[j][1] * P->bb[j][0];
            cs = fabs(P->bn[i][0] * sv[0] + P->bn[i][1] * sv[1] + P->bn[i][2] * sv[2]);
            P->cos_g1[i][j] = fabs(P->bn[i][0] * P->bb[j][0] + P->bn[i][1] * P->bb[j][1] + P->bn[i][2] * P->bb[j][2]);
            P->sin_g1[i][j] = fabs(sqrt(fabs(1.0 - P->cos_g1[i][j] * P->cos_g1[i][j])));
            P->cos_g2[i][j] = cs;
            P->sin_g2[i][j] = fabs(sqrt(fabs(1.0 - cs * cs)));
            P->cos_g3[i][j] = fabs(P->bn[i][0] * P->bn[j][0] + P->bn[i][1] * P->bn[j][1] + P->bn[i][2] * P->bn[j][2]);
            P->sin_g3[i][j] = fabs(sqrt(fabs(1.0 - P->cos_g3[i][j] * P->cos_g3[i][j])));
        }
    }

    /* allocation */
#define AL(x, n) s->x = (double *)calloc((size_t)(n) * s->nxyz, sizeof(double))
    AL(tran, 9); AL(cgr, 36); AL(schm, 60); AL(sig, 9); AL(eps, 9); AL(edot, 9); AL(disgrad, 9);
    AL(velgrad, 9); AL(fluct, 9); AL(rho_s, 12); AL(rho_m, 12); AL(rho_F, 12); AL(rho_P, 12);
    AL(trial_rho_s, 12); AL(gamdot, 12); AL(gam_acum, 12); AL(trial_gam_acum, 12); AL(rho_dipole, 12);
    AL(rho_forest, 12); AL(rho_inplane, 12); AL(rho_tot, 1); AL(nrit, 1); AL(work, 2); AL(ksig, 18);
    if (s->opt.warm_start == 1) {
        s->dsig = (float *)calloc((size_t)6 * s->nxyz, sizeof(float));
        if (!s->dsig) return 2005;
    }
    AL(conv_s, 1); AL(conv_e, 1);
    if (s->opt.gnd) {
        AL(rho_g1, 12); AL(rho_g2, 12); AL(rho_g3, 12); AL(gadot_grad, 36);
        s->gbc = (unsigned char *)calloc(s->nxyz, 1);
        if (!s->rho_g1 || !s->rho_g2 || !s->rho_g3 || !s->gadot_grad || !s->gbc) return 2005;
    }
#undef AL
    s->phase = (int *)calloc(s->nxyz, sizeof(int));
    s->grain = (int *)calloc(s->nxyz, sizeof(int));
    if (!s->ksig || !s->phase) return 2005;

    /* InitMicroStruct (init.c:233-396): lines are read in order, the per-voxel
       rotations run in parallel, C066 is summed in voxel order. */
    {
        FILE *fp;
        char buf[80];
        long nph1 = 0;
        double *eul = (double *)malloc(sizeof(double) * 3 * s->nxyz);
        resolve(path, s->workdir, d->initial_ms);
        fp = fopen(path, "r");
        if (!fp) { free(eul); return 2001; }
        memset(c066, 0, sizeof(c066));
        for (p = 0; p < s->nxyz; p++) {
            double t1 = 0, th = 0, om = 0;
            int ii = 0, jj = 0, kk = 0, jgr = 0, jph = 0;
            long idx;
            if (!fgets(buf, 80, fp)) { fclose(fp); free(eul); return 1139; }
            sscanf(buf, "%lf %lf %lf %d %d %d %d %d", &t1, &th, &om, &ii, &jj, &kk, &jgr, &jph);
            idx = ((long)(ii - 1) * s->n[1] + jj - 1) * s->n[2] + kk - 1;
            if (idx != p || jph < 1 || jph > NPH) { fclose(fp); free(eul); return 1139; }
            if (jph == 1) nph1++;
            s->grain[p] = jgr;
            s->phase[p] = jph;
            eul[3 * p] = t1; eul[3 * p + 1] = th; eul[3 * p + 2] = om;
        }
        fclose(fp);
        #pragma omp parallel for schedule(static)
        for (p = 0; p < s->nxyz; p++) {
            double t1 = eul[3 * p], th = eul[3 * p + 1], om = eul[3 * p + 2], sa2xt[9], Cs[81];
            int jph = s->phase[p], a, b, c, e;
            if (!s->ph[jph - 1].crystalline) continue;
            t1 *= O_PI / 180.; th *= O_PI / 180.; om *= O_PI / 180.;
            orc_euler_to_matrix(&t1, &th, &om, sa2xt, 2);
            for (a = 0; a < 3; a++) for (b = 0; b < 3; b++) s->tran[9 * p + 3 * a + b] = sa2xt[3 * b + a];
            /* Ten4thTransform opt 2 (kinematics.c:262-275): C_s = sum a_mi a_nj a_pk a_ql C_xt */
            {
                const double *Cx = s->ph[jph - 1].C4;
                for (a = 0; a < 3; a++) for (b = 0; b < 3; b++) for (c = 0; c < 3; c++) for (e = 0; e < 3; e++) {
                    double acc = 0.0;
                    int m, n2, q, r;
                    for (m = 0; m < 3; m++) for (n2 = 0; n2 < 3; n2++) for (q = 0; q < 3; q++) for (r = 0; r < 3; r++)
                        acc += sa2xt[3 * m + a] * sa2xt[3 * n2 + b] * sa2xt[3 * q + c] * sa2xt[3 * r + e] *
                               Cx[((m * 3 + n2) * 3 + q) * 3 + r];
                    Cs[((a * 3 + b) * 3 + c) * 3 + e] = acc;
                }
            }
            orc_chg_basis(0, 0, s->cgr + 36 * p, Cs, 4, 6);
        }
        free(eul);
        for (p = 0; p < s->nxyz; p++) {
            int a;
            if (!s->ph[s->phase[p] - 1].crystalline) continue;
            for (a = 0; a < 36; a++) c066[a / 6][a % 6] += s->cgr[36 * p + a];
        }
        s->wgt_ph1 = (1.0 * (double)nph1) / s->nxyz;
        for (i = 0; i < 6; i++) for (j = 0; j < 6; j++) s->C066[i][j] = c066[i][j] * s->wgt;
        memcpy(s->S066, s->C066, sizeof(s->C066));
        rc = lu_inv66(s->S066);
        if (rc) return rc;
        orc_chg_basis(0, 0, &s->C066[0][0], s->C0, 3, 6);
        orc_chg_basis(0, 0, &s->S066[0][0], s->S0, 3, 6);
    }

    /* boundary conditions (init.c:865-901) */
    {
        int fs = d->VelGrad_BC_Flag[0] + d->VelGrad_BC_Flag[1] + d->VelGrad_BC_Flag[2];
        double vg[6], d6[6], dev[3][3];
        if (fs == 2) return 116;
        for (i = 0; i < 6; i++) vg[i] = d->VelGrad_BC[i] * d->VelGrad_BC_Flag[i];
        vn_to_full(vg, s->Udot);
        for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) {
            s->Udot_s[i][j] = (s->Udot[i][j] + s->Udot[j][i]) / 2.;
            s->Udot_a[i][j] = (s->Udot[i][j] - s->Udot[j][i]) / 2.;
        }
        t_to_v6(s->Udot_s, d6, 6);
        v6_to_t(d6, 5, dev);
        s->d_vm = 0.0;
        for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) s->d_vm += pow(dev[i][j], 2.0);
        s->d_vm = sqrt(2. / 3. * s->d_vm);
        for (i = 0; i < 6; i++)
            if ((d->Stress_BC_Flag[i] * d->VelGrad_BC_Flag[i] != 0) || ((d->Stress_BC_Flag[i] + d->VelGrad_BC_Flag[i]) != 1))
                return 114;
        s->creep = 0;
        for (i = 0; i < 6; i++) s->creep += d->Stress_BC_Flag[i];
        s->creep /= 6;
        vn_to_full(d->Stress_BC, s->Scauchy);
    }

    /* dislocation state (init.c:1068-1098) */
    for (p = 0; p < 12 * s->nxyz; p++) { s->rho_s[p] = d->rho_SSD_initial; s->trial_rho_s[p] = d->rho_SSD_initial; }

    /* InitFields (init.c:444-516) */
    for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) {
        s->DisGradAvg[i][j] = s->Udot[i][j] * d->TimeStep;
        s->dDisGradAvg[i][j] = 0.0;
        s->dDisGradAvg_acum[i][j] = 0.0;
    }
    #pragma omp parallel for schedule(static)
    for (p = 0; p < s->nxyz; p++) {
        int jph = s->phase[p], a, b, c, e;
        double cg[81];
        if (!is_xtal(s, jph)) continue;
        orc_chg_basis(0, 0, s->cgr + 36 * p, cg, 3, 6);
        for (a = 0; a < 3; a++) for (b = 0; b < 3; b++) {
            double acc = 0.0;
            for (c = 0; c < 3; c++) for (e = 0; e < 3; e++) acc += cg[((a * 3 + b) * 3 + c) * 3 + e] * s->DisGradAvg[c][e];
            if (s->creep == 1) acc += s->Scauchy[a][b];
            s->sig[9 * p + 3 * a + b] = acc;
        }
    }
    s->e_vm = s->d_vm * d->TimeStep;
    s->d_vm0 = s->d_vm;
    s->e_vm0 = s->e_vm;
    s->err = s->opt.err_override > 0 ? s->opt.err_override : d->Err;
    s->itermax = s->opt.itermax_override > 0 ? s->opt.itermax_override : d->IterMax;
    s->nsteps = s->opt.nsteps_override > 0 ? s->opt.nsteps_override : d->N_steps;
    if (s->opt.tau0_mode == 1) for (ph = 0; ph < NPH; ph++) s->ph[ph].tau0xf = s->opt.tau0_value;
    if (s->opt.gnd) gradient_exchange_grain_id(s);
    return 0;
}

/* ------------------------------------------------------------------ */
/* StrainRate_Orowan, FORWARD_PLUS_BACKWARD (constitutive.c:507-728)    */
/* Without the gnd option the GND densities are identically zero (DD_GND */
/* off), so the GND terms of 565-576 contribute exactly +0.             */
/* ------------------------------------------------------------------ */
static void strain_rate(orc_state *s, long idx, int jph, const double *stress, double *edot, double d_edot[6][6])
{
    const orc_phase *P = &s->ph[jph - 1];
    const int ns = P->nsys;
    const double G = P->shear_g, b = P->bburg, kT = s->kT;
    const double tau_0 = P->tau0xf;     /* reference reads tau[jph-1][i][0] with i uninitialised (537) */
    double gam[NSYS], gam2[NSYS], rho_f[NSYS];
    const double *rs = s->trial_rho_s + 12 * idx;
    double *rF = s->rho_F + 12 * idx, *rP = s->rho_P + 12 * idx, *rm = s->rho_m + 12 * idx;
    const double *sc = s->schm + 60 * idx;
    int i, j, k;
    for (i = 0; i < ns; i++) {
        double aF = 0.0, aP = 0.0;
        if (s->opt.gnd) {
            const double *g1 = s->rho_g1 + 12 * idx, *g2 = s->rho_g2 + 12 * idx, *g3 = s->rho_g3 + 12 * idx;
            for (j = 0; j < ns; j++) {
                aF += P->chi1[i][j] * (rs[j] + fabs(g1[j] * P->cos_g1[i][j]) + fabs(g2[j] * P->cos_g2[i][j]) +
                                       fabs(g3[j] * P->cos_g3[i][j]));
                aP += P->chi[i][j] * (rs[j] + fabs(g1[j] * P->sin_g1[i][j]) + fabs(g2[j] * P->sin_g2[i][j]) +
                                      fabs(g3[j] * P->sin_g3[i][j]));
            }
        } else {
            for (j = 0; j < ns; j++) {
                aF += P->chi1[i][j] * (rs[j] + 0.0 + 0.0 + 0.0);
                aP += P->chi[i][j] * (rs[j] + 0.0 + 0.0 + 0.0);
            }
        }
        rF[i] = aF; rho_f[i] = aF; rP[i] = aP;
    }
    for (i = 0; i < ns; i++) {
        double tau_pass, rss, tau_eff = 0.0, tau_cut, gam_o, tmp;
        const double *si = sc + 5 * i;
        tau_pass = P->c1 * G * b * sqrt(fabs(rP[i] + rm[i])) * 1.0E-3;
        /* the DRX passing-stress term (constitutive.c:588-590) is DeltaTau_DRX*exp(-1000-...) == 0 */
        rss = si[0] * stress[0] + si[1] * stress[1] + si[2] * stress[2] + si[3] * stress[3] + si[4] * stress[4];
        if (fabs(rss) > tau_pass) tau_eff = (fabs(rss) - tau_pass);
        tau_cut = (kT / O_MY_J_IN_EV) * sqrt(rho_f[i]) / P->c2 / P->c3 / b / b * 0.1;
        rm[i] = 2.0 * kT / P->c1 / P->c2 / P->c3 / G / (pow(b, 3.0)) * sqrt(rP[i] * rho_f[i]) * 160.2177;
        gam_o = 1E-3 * (rm[i]) * O_TRIAL_FREQ_DRX * P->c2 * b * exp(-1.0 * P->q_slip) / sqrt(rho_f[i]);
        tmp = fabs(rss) > tau_pass ? gam_o * sinh(tau_eff / (tau_0 + tau_cut)) : 0.0;
        gam[i] = tmp * ((double)(2 * (rss > 0) - 1));
        tmp = fabs(rss) > tau_pass ? gam_o * cosh(tau_eff / (tau_0 + tau_cut)) / (tau_0 + tau_cut) : 0.0;
        gam2[i] = tmp * (fabs(rss) > tau_pass);
        s->gamdot[12 * idx + i] = gam[i];
        s->trial_gam_acum[12 * idx + i] = s->gam_acum[12 * idx + i] + gam[i] * s->dk.TimeStep;
    }
    for (i = 0; i < 5; i++) {
        double acc = 0.;
        for (j = 0; j < ns; j++) acc += sc[5 * j + i] * gam[j];
        edot[i] = acc;
    }
    edot[5] = 0.0;
    for (i = 0; i < 5; i++) for (j = 0; j < 5; j++) {
        double acc = 0.0;
        for (k = 0; k < ns; k++) acc += sc[5 * k + i] * sc[5 * k + j] * gam2[k];
        d_edot[i][j] = acc;
    }
    for (i = 0; i < 6; i++) { d_edot[i][5] = 0.0; d_edot[5][i] = 0.0; }
}

int orc_strain_rate(orc_state *s, int idx, const double *sig6, double *edot6, double *dedot)
{
    int jph = s->phase[idx];
    if (!is_xtal(s, jph)) return -1;
    strain_rate(s, idx, jph, sig6, edot6, (double (*)[6])dedot);
    return 0;
}

/* UpdateStress body for one voxel (evolution.c:157-427, NR_MODIFIED off) */
static int nr_voxel(orc_state *s, long p, double *cs, double *ce)
{
    int jph = s->phase[p], i, j, k, iterl = 0;
    double sg66[6][6], lam6[6], sig6[6], eps6[6], str6[6], edot6[6], d_edot[6][6], tot[6], old[6], res[6];
    double jac[6][6], signorm = 0.0, erral, conv_s = 0.0, conv_e = 0.0;
    t33 lam, ep, st;
    const double dt = s->dk.TimeStep;
    if (!is_xtal(s, jph)) {
        for (i = 0; i < 9; i++) { s->sig[9 * p + i] = 0.0; s->edot[9 * p + i] = 0.0; }
        *cs = 0.0; *ce = 0.0;
        return 0;
    }
    for (i = 0; i < 36; i++) (&sg66[0][0])[i] = s->cgr[36 * p + i];
    lu_inv66(sg66);
    for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) {
        lam[i][j] = s->sig[9 * p + 3 * i + j];
        ep[i][j] = s->eps[9 * p + 3 * i + j];
        st[i][j] = (s->disgrad[9 * p + 3 * i + j] + s->disgrad[9 * p + 3 * j + i]) / 2.0;
    }
    t_to_v6(lam, lam6, 6);
    t_to_v6(lam, sig6, 6);
    /* warm_start 1 (no counterpart in the reference): start from lam + (lam - previous lam) */
    if (s->dsig && s->ws_use)
        for (i = 0; i < 6; i++) sig6[i] += (double)s->dsig[6 * p + i];
    t_to_v6(ep, eps6, 6);
    t_to_v6(st, str6, 6);
    for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) signorm += lam[i][j] * lam[i][j];
    signorm = sqrt(signorm);
    erral = 10 * 1E-10;
    while ((iterl < 500) && (fabs(erral) > fabs(1E-10))) {
        double cnr = 0.0, cee = 0.0, css = 0.0;
        iterl++;
        for (i = 0; i < 6; i++) old[i] = sig6[i];
        strain_rate(s, p, jph, sig6, edot6, d_edot);
        for (i = 0; i < 6; i++) {
            tot[i] = eps6[i] + edot6[i] * dt;
            for (j = 0; j < 6; j++) tot[i] += sg66[i][j] * sig6[j];
        }
        for (i = 0; i < 6; i++) {
            res[i] = sig6[i] - lam6[i];
            for (j = 0; j < 6; j++) res[i] += s->C066[i][j] * (tot[j] - str6[j]);
        }
        for (i = 0; i < 6; i++) for (j = 0; j < 6; j++) {
            jac[i][j] = (double)(i == j);
            for (k = 0; k < 6; k++) jac[i][j] += s->C066[i][k] * (sg66[k][j] + d_edot[k][j] * dt);
        }
        lu_inv66(jac);
        for (i = 0; i < 6; i++) for (j = 0; j < 6; j++) sig6[i] -= jac[i][j] * res[j];
        for (i = 0; i < 6; i++) {
            cnr += (sig6[i] - old[i]) * (sig6[i] - old[i]);
            cee += (tot[i] - str6[i]) * (tot[i] - str6[i]);
            css += (sig6[i] - lam6[i]) * (sig6[i] - lam6[i]);
        }
        erral = sqrt(cnr) / signorm;
        if (s->opt.verbose >= 2 && p == s->opt.reserved[0])
            printf("  nr p=%ld it=%d erral=%e sig=%e %e %e %e %e %e edotmax=%e\n", p, iterl, erral, sig6[0], sig6[1],
                   sig6[2], sig6[3], sig6[4], sig6[5], fmax(fabs(edot6[0]), fabs(edot6[1])));
        conv_s = sqrt(css);
        conv_e = sqrt(cee);
    }
    if (s->dsig)
        for (i = 0; i < 6; i++) s->dsig[6 * p + i] = (float)(sig6[i] - lam6[i]);
    {
        t33 so, eo;
        v6_to_t(sig6, 6, so);
        v6_to_t(edot6, 6, eo);
        for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) {
            s->sig[9 * p + 3 * i + j] = so[i][j];
            s->edot[9 * p + 3 * i + j] = eo[i][j];
        }
    }
    *cs = conv_s;
    *ce = conv_e;
    return iterl;
}

int orc_nr_voxel(orc_state *s, int idx, double *err_s, double *err_e) { return nr_voxel(s, idx, err_s, err_e); }

void orc_update_schmid(orc_state *s) { update_schmid(s); }

int orc_get_schmid(orc_state *s, int idx, double *out60)
{
    memcpy(out60, s->schm + 60 * (long)idx, 60 * sizeof(double));
    return 0;
}

/* One k point of CalcGAMMA (init.c:542-592) contracted as evolution.c:1169-1183 */
static void green_point(const double *C0, const double *nv, int dc, int ebc, const double *S0,
                        const double *ks, double *du)
{
    double G[81];
    int i, j, k, l;
    if (dc) {
        for (i = 0; i < 81; i++) G[i] = ebc == 0 ? -1.0 * S0[i] : 0.0;
    } else {
        double A[3][3], a[3][3], det;
        for (i = 0; i < 3; i++) for (k = 0; k < 3; k++) {
            double acc = 0.0;
            for (j = 0; j < 3; j++) for (l = 0; l < 3; l++) acc += C0[((i * 3 + j) * 3 + k) * 3 + l] * nv[j] * nv[l];
            A[i][k] = acc;
        }
        /* M3Inverse (init.c:518-540) */
        det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
              A[0][2] * (A[1][0] * A[2][1] - A[2][0] * A[1][1]);
        a[0][0] = A[1][1] * A[2][2] - A[1][2] * A[2][1];
        a[0][1] = A[0][2] * A[2][1] - A[0][1] * A[2][2];
        a[0][2] = A[0][1] * A[1][2] - A[0][2] * A[1][1];
        a[1][0] = A[1][2] * A[2][0] - A[1][0] * A[2][2];
        a[1][1] = A[0][0] * A[2][2] - A[0][2] * A[2][0];
        a[1][2] = A[0][2] * A[1][0] - A[0][0] * A[1][2];
        a[2][0] = A[1][0] * A[2][1] - A[1][1] * A[2][0];
        a[2][1] = A[0][1] * A[2][0] - A[0][0] * A[2][1];
        a[2][2] = A[0][0] * A[1][1] - A[0][1] * A[1][0];
        for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) a[i][j] /= det;
        for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) for (k = 0; k < 3; k++) for (l = 0; l < 3; l++)
            G[((i * 3 + j) * 3 + k) * 3 + l] = -1.0 * a[i][k] * nv[j] * nv[l];
    }
    for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) {
        double r = 0.0, im = 0.0;
        for (k = 0; k < 3; k++) for (l = 0; l < 3; l++) {
            r += G[((i * 3 + j) * 3 + k) * 3 + l] * ks[2 * (3 * k + l)];
            im += G[((i * 3 + j) * 3 + k) * 3 + l] * ks[2 * (3 * k + l) + 1];
        }
        du[2 * (3 * i + j)] = r;
        du[2 * (3 * i + j) + 1] = im;
    }
}

void orc_green_point(const double *c0, const double *nvec, int dc, int elast_bc, const double *s0,
                     const double *ks, double *du)
{
    green_point(c0, nvec, dc, elast_bc, s0, ks, du);
}

/* get_smacro (evolution.c:590-664) */
static void get_smacro(orc_state *s)
{
    static const int IJV[2][6] = {{0, 1, 2, 1, 0, 0}, {0, 1, 2, 2, 2, 1}};
    int i, j, k;
    long p;
    double sav6[6], sda[3][3];
    for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) {
        double a = 0.0, a1 = 0.0;
        for (p = 0; p < s->nxyz; p++) {
            a += s->sig[9 * p + 3 * i + j];
            if (s->phase[p] == 1) a1 += s->sig[9 * p + 3 * i + j];
        }
        s->SigAvg[i][j] = a * s->wgt;
        s->SigAvg1[i][j] = a1 * (s->wgt / s->wgt_ph1);
    }
    for (i = 0; i < 6; i++) {
        int ii = IJV[0][i], jj = IJV[1][i];
        s->dDisGradAvg[ii][jj] = 0.0;
        if (s->dk.VelGrad_BC_Flag[i] == 0) {
            for (k = 0; k < 6; k++) {
                int kk = IJV[0][k], ll = IJV[1][k];
                s->dDisGradAvg[ii][jj] += s->S0[((ii * 3 + jj) * 3 + kk) * 3 + ll] * s->dk.Stress_BC_Flag[k] *
                                          (s->Scauchy[kk][ll] - s->SigAvg[kk][ll]);
            }
        }
    }
    for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) s->dDisGradAvg_acum[i][j] += s->dDisGradAvg[i][j];
    t_to_v6(s->SigAvg, sav6, 6);
    v6_to_t(sav6, 5, sda);
    s->s_vm = 0.0;
    {
        /* evolution.c:642-651: the mean subtracted from the diagonal uses SigDevAvg[1][1],
           [2][2] and the out-of-bounds [3][3]; svm_stray stands in for the last. */
        double tr = s->opt.svm_mode == 1 ? (sda[0][0] + sda[1][1] + sda[2][2]) : (sda[1][1] + sda[2][2] + s->opt.svm_stray);
        for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) {
            if (i != j) s->s_vm += sda[i][j] * sda[i][j];
            else s->s_vm += (sda[i][j] - tr / 3.0) * (sda[i][j] - tr / 3.0);
        }
    }
    s->s_vm = sqrt(3. / 2. * s->s_vm);
    t_to_v6(s->SigAvg1, sav6, 6);
    v6_to_t(sav6, 5, sda);
    s->s_vm1 = 0.0;
    for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) s->s_vm1 += sda[i][j] * sda[i][j];
    s->s_vm1 = sqrt(3. / 2. * s->s_vm1);
}

/* Orientation (constitutive.c:142-194) */
static void orientation(double tran[3][3], double rot[3][3])
{
    double w[3], wn = 0.0, wt, om[3][3], om2[3][3], tn[3][3];
    int i, a, b;
    w[0] = rot[2][1]; w[1] = rot[0][2]; w[2] = rot[1][0];
    for (i = 0; i < 3; i++) wn += w[i] * w[i];
    wn = sqrt(wn);
    wt = tan(wn / 2.0);
    if (fabs(wn) < 1E-6) wn = 1.0;
    for (i = 0; i < 3; i++) w[i] *= wt / wn;
    om[0][0] = 0.; om[0][1] = -1. * w[2]; om[0][2] = w[1];
    om[1][0] = w[2]; om[1][1] = 0.; om[1][2] = -1. * w[0];
    om[2][0] = -1. * w[1]; om[2][1] = w[0]; om[2][2] = 0.;
    for (a = 0; a < 3; a++) for (b = 0; b < 3; b++) {
        om2[a][b] = 0.0;
        for (i = 0; i < 3; i++) om2[a][b] += om[a][i] * om[i][b];
    }
    wt = wt * wt;
    for (a = 0; a < 3; a++) for (b = 0; b < 3; b++) rot[a][b] = (double)(a == b) + 2. * (om[a][b] + om2[a][b]) / (1. + wt);
    for (a = 0; a < 3; a++) for (b = 0; b < 3; b++) {
        tn[a][b] = 0.0;
        for (i = 0; i < 3; i++) tn[a][b] += rot[a][i] * tran[i][b];
    }
    memcpy(tran, tn, sizeof(tn));
}

/* update_orient (constitutive.c:196-273) */
static void update_orient(orc_state *s)
{
    long p;
    const double dt = s->dk.TimeStep;
    #pragma omp parallel for schedule(static)
    for (p = 0; p < s->nxyz; p++) {
        int jph = s->phase[p], is, i, j, a, b;
        double tm[3][3], rl[3][3], lp[3][3], rs[3][3], rot[3][3];
        const double *vg = s->velgrad + 9 * p;
        if (!is_xtal(s, jph)) continue;
        for (a = 0; a < 3; a++) for (b = 0; b < 3; b++) {
            rl[a][b] = (vg[3 * a + b] - vg[3 * b + a]) / 2.;
            lp[a][b] = 0.0;
            tm[a][b] = s->tran[9 * p + 3 * a + b];
        }
        for (is = 0; is < s->ph[jph - 1].nsys; is++) {
            double bs[3], ns[3];
            for (i = 0; i < 3; i++) {
                bs[i] = 0.0; ns[i] = 0.0;
                for (j = 0; j < 3; j++) {
                    bs[i] += tm[i][j] * s->ph[jph - 1].bb[is][j];
                    ns[i] += tm[i][j] * s->ph[jph - 1].bn[is][j];
                }
            }
            for (a = 0; a < 3; a++) for (b = 0; b < 3; b++) lp[a][b] += bs[a] * ns[b] * s->gamdot[12 * p + is];
        }
        for (a = 0; a < 3; a++) for (b = 0; b < 3; b++) rs[a][b] = (lp[a][b] - lp[b][a]) / 2.;
        for (a = 0; a < 3; a++) for (b = 0; b < 3; b++) rot[a][b] = (s->Udot_a[a][b] + rl[a][b] - rs[a][b]) * dt;
        orientation(tm, rot);
        for (a = 0; a < 3; a++) for (b = 0; b < 3; b++) s->tran[9 * p + 3 * a + b] = tm[a][b];
    }
}

void orc_orientation(double *tran9, double *rot9)
{
    orientation((double (*)[3])tran9, (double (*)[3])rot9);
}

void orc_update_orient(orc_state *s) { update_orient(s); }

/* DislocationEvolution (constitutive.c:1353-1529), istep <= 2e6 branch */
static void dislocation_evolution(orc_state *s, int istep)
{
    long p;
    const double dt = s->dk.TimeStep, T = s->dk.T__K, D0 = s->dk.D_0;
    const double nu = s->dk.ElastConst_PhaseI[3];
    #pragma omp parallel for schedule(static)
    for (p = 0; p < s->nxyz; p++) {
        int jph = s->phase[p], i, a, b;
        const orc_phase *P;
        double sig6[6], rss[NSYS], G, dot_SSD = 0.0;
        t33 sa;
        if (!is_xtal(s, jph)) continue;
        P = &s->ph[jph - 1];
        G = P->shear_g;
        for (a = 0; a < 3; a++) for (b = 0; b < 3; b++) sa[a][b] = s->sig[9 * p + 3 * a + b];
        t_to_v6(sa, sig6, 6);
        for (i = 0; i < P->nsys; i++) {
            const double *sc = s->schm + 60 * p + 5 * i;
            rss[i] = sc[0] * sig6[0] + sc[1] * sig6[1] + sc[2] * sig6[2] + sc[3] * sig6[3] + sc[4] * sig6[4];
        }
        for (i = 0; i < P->nsys; i++) {
            const double bb = P->bburg;
            double *rs = s->rho_s + 12 * p + i;
            double gd = s->gamdot[12 * p + i], climb;
            double d_dip = sqrt(3.0) * G * bb / 16.0 / O_PI / (1 - nu) / (fabs(rss[i]) + O_EPS_DIPOLE);
            double hydro = G * bb / O_PI / (1 - nu) / d_dip;
            if (istep <= 2000000) {
                climb = 1E-17 * s->rho_tot[p] * 0.5 * 1E24 * (d_dip) * (D0 / bb) * pow(*rs, 2.0) *
                        (exp(fabs(hydro) * bb * bb * bb * 100 / (1.38 * T)) - 1.0) * exp(-1.0 * P->q_bulk);
                s->rho_dipole[12 * p + i] += climb * dt;
                s->rho_forest[12 * p + i] += fabs(gd) * (P->c4 * sqrt(s->rho_F[12 * p + i])) * dt;
                s->rho_inplane[12 * p + i] += fabs(gd) * (P->c5 * (*rs)) * dt;
                dot_SSD = fabs(gd) * (P->c4 * sqrt(s->rho_F[12 * p + i]) + P->c6 * (d_dip) * s->rho_m[12 * p + i] -
                                      P->c5 * (*rs)) - climb;
                *rs += dot_SSD * dt;
            }
            s->trial_rho_s[12 * p + i] = *rs;
            s->gam_acum[12 * p + i] = s->trial_gam_acum[12 * p + i];
            if (s->opt.gnd) {
                /* DD_GND with DD_GND_FIRST_ORDER (evpFlags.h): screw, edge-I, edge-II (1428-1519) */
                const double *gg = s->gadot_grad + 36 * p + 3 * i;
                const double L0 = s->dk.L0;
                double gnd1, dot_GND;
                gnd1 = gg[0] * P->bp[i][0] + gg[1] * P->bp[i][1] + gg[2] * P->bp[i][2];
                dot_GND = gnd1 / bb / L0;
                s->rho_g1[12 * p + i] += dot_GND * dt;
                gnd1 = gg[0] * P->bb[i][0] + gg[1] * P->bb[i][1] + gg[2] * P->bb[i][2];
                dot_GND = -1.0 * gnd1 / bb / L0;
                s->rho_g2[12 * p + i] += dot_GND * dt;
                gnd1 = 0.0;
                dot_GND = gnd1;
                s->rho_g3[12 * p + i] += dot_GND * dt;
            }
        }
        (void)dot_SSD;
    }
}

static double avg12(const orc_state *s, const double *f)   /* AvgDislocation (constitutive.c:1681-1808) */
{
    double acc = 0.0;
    long p;
    int i;
    for (p = 0; p < s->nxyz; p++) {
        int jph = s->phase[p];
        if (!is_xtal(s, jph)) continue;
        for (i = 0; i < s->ph[jph - 1].nsys; i++) acc += f[12 * p + i];
    }
    return acc * (1.0 / s->nxyz);
}

static void write_header_files(orc_state *s)
{
    char path[1100];
    if (!s->opt.write_files) return;
    snprintf(path, sizeof(path), "%s/vm.out", s->workdir[0] ? s->workdir : ".");
    s->fvm = fopen(path, "w+");
    snprintf(path, sizeof(path), "%s/dd.out", s->workdir[0] ? s->workdir : ".");
    s->fdd = fopen(path, "w+");
    snprintf(path, sizeof(path), "%s/drx.out", s->workdir[0] ? s->workdir : ".");
    s->fdrx = fopen(path, "w+");
    snprintf(path, sizeof(path), "%s/err.out", s->workdir[0] ? s->workdir : ".");
    s->ferr = fopen(path, "w+");
    /* init.c:497-505 */
    if (s->fvm) fprintf(s->fvm, "TIME\tEVM\tEVMP\tDVM\tDVMP\tSVM\tSVMSVM_1\n");
    if (s->fdd) fprintf(s->fdd, "TIME\tEVM\tDD_F\tDD_P\tDD_s\tDD_m\tDD_g\n");
    if (s->fdrx) fprintf(s->fdrx, "TIME\tEVM\tNEW_IDs\n");
    if (s->ferr) fprintf(s->ferr, "IT\tERRE\tERRS\tSVM\n");
}

int orc_open(const char *deck, const char *workdir, const orc_options *opt, orc_state **out)
{
    orc_state *s = (orc_state *)calloc(1, sizeof(orc_state));
    int rc;
    if (!s) return 2005;
    if (opt) s->opt = *opt; else orc_default_options(&s->opt);
    if (workdir) snprintf(s->workdir, sizeof(s->workdir), "%s", workdir);
#ifdef _OPENMP
    if (s->opt.nthreads > 0) omp_set_num_threads(s->opt.nthreads);
#endif
    rc = parse_deck(deck, &s->dk);
    if (!rc) rc = setup(s);
    if (rc) { orc_close(s); return rc; }
    write_header_files(s);
    *out = s;
    return 0;
}

/* One increment: evolution.c:726-1520 */
int orc_step(orc_state *s, orc_report *rep)
{
    const double dt = s->dk.TimeStep;
    const long N = s->nxyz;
    const int n0 = s->n[0], n1 = s->n[1], n2 = s->n[2];
    long p;
    int iter = 0, i, j, nrmax = 0;
    long nrtot = 0;
    double nrwave = 0.0;
    orc_report r;
    memset(&r, 0, sizeof(r));
    if (s->istep >= s->nsteps) return 2006;
    if (s->creep == 1 && s->e_vm > 0.3) return 2007;
    if (s->ferr && s->nsteps != 1) fprintf(s->ferr, "STEP = %d\n", s->istep);

    /* rho_tot with first_pf == 0 (evolution.c:1030-1048) */
    for (p = 0; p < N; p++) {
        int jph = s->phase[p];
        double acc = 0.0;
        if (is_xtal(s, jph))
            for (i = 0; i < s->ph[jph - 1].nsys; i++) {
                const double g1 = s->opt.gnd ? s->rho_g1[12 * p + i] : 0.0, g2 = s->opt.gnd ? s->rho_g2[12 * p + i] : 0.0;
                const double g3 = s->opt.gnd ? s->rho_g3[12 * p + i] : 0.0;
                acc += s->rho_s[12 * p + i] + s->rho_m[12 * p + i] + sqrt(pow(g1, 2.0) + pow(g2, 2.0) + pow(g3, 2.0));
            }
        s->rho_tot[p] = acc;
    }
    /* evolution.c:1115-1128 */
    for (p = 0; p < N; p++) for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) {
        s->velgrad[9 * p + 3 * i + j] = s->disgrad[9 * p + 3 * i + j];
        if (s->creep == 1) s->disgrad[9 * p + 3 * i + j] = s->DisGradAvg[i][j] + s->dDisGradAvg_acum[i][j];
        else s->disgrad[9 * p + 3 * i + j] += s->Udot[i][j] * dt;
    }
    if (s->istep == 1 || s->dk.Update_Flag == 1) update_schmid(s);
    memset(s->dDisGradAvg, 0, sizeof(s->dDisGradAvg));
    memset(s->dDisGradAvg_acum, 0, sizeof(s->dDisGradAvg_acum));
    s->Err_e = 2. * s->err;
    s->Err_s = 2. * s->err;

    for (;;) {
        double ee = 0.0, es = 0.0;
        int c;
        if (s->opt.fixed_iters > 0) { if (iter >= s->opt.fixed_iters) break; }
        else if (!((iter < s->itermax) && (fmax(fabs(s->Err_s), fabs(s->Err_e)) > s->err))) break;
        iter++;
        /* forward FFT of the 9 stress components (evolution.c:1157-1168) */
        for (c = 0; c < 9; c++) {
            int a = c / 3, b = c % 3;
            if (a > b) continue;
            for (p = 0; p < N; p++) { s->work[2 * p] = s->sig[9 * p + c]; s->work[2 * p + 1] = 0.0; }
            orc_fft3d(s->work, n0, n1, n2, -1, 0);
            for (p = 0; p < N; p++) {
                s->ksig[18 * p + 2 * c] = s->work[2 * p];
                s->ksig[18 * p + 2 * c + 1] = s->work[2 * p + 1];
                s->ksig[18 * p + 2 * (3 * b + a)] = s->work[2 * p];
                s->ksig[18 * p + 2 * (3 * b + a) + 1] = s->work[2 * p + 1];
            }
        }
        /* Green contraction (evolution.c:1169-1183) */
        #pragma omp parallel for schedule(static)
        for (p = 0; p < N; p++) {
            double du[18];
            green_point(s->C0, s->nvec + 3 * p, p == 0, s->dk.ElastBC, s->S0, s->ksig + 18 * p, du);
            memcpy(s->ksig + 18 * p, du, sizeof(du));
        }
        /* inverse FFT and Eq. 15 (evolution.c:1189-1201) */
        for (c = 0; c < 9; c++) {
            int a = c / 3, b = c % 3;
            for (p = 0; p < N; p++) { s->work[2 * p] = s->ksig[18 * p + 2 * c]; s->work[2 * p + 1] = s->ksig[18 * p + 2 * c + 1]; }
            orc_fft3d(s->work, n0, n1, n2, +1, 0);
            for (p = 0; p < N; p++) {
                s->disgrad[9 * p + c] += s->dDisGradAvg[a][b] + s->work[2 * p] / N;
                s->fluct[9 * p + c] += s->work[2 * p] / N;
            }
        }
        /* UpdateStress (evolution.c:1224) */
        s->ws_use = iter > 1;
        #pragma omp parallel for schedule(dynamic, 64)
        for (p = 0; p < N; p++) s->nrit[p] = nr_voxel(s, p, s->conv_s + p, s->conv_e + p);
        for (p = 0; p < N; p++) {
            /* the reference asserts on a NaN slip rate (constitutive.c:660,683) */
            if (s->sig[9 * p] != s->sig[9 * p] || s->sig[9 * p + 8] != s->sig[9 * p + 8]) return 3001;
            ee += s->conv_e[p];
            es += s->conv_s[p];
            nrtot += (long)s->nrit[p];
            if ((int)s->nrit[p] > nrmax) nrmax = (int)s->nrit[p];
        }
        for (p = 0; p < N; p += 64) {
            double gm = 0.0;
            long q;
            for (q = p; q < N && q < p + 64; q++) if (s->nrit[q] > gm) gm = s->nrit[q];
            nrwave += gm;
        }
        s->Err_e = ee * s->wgt;
        s->Err_s = es * s->wgt;
        get_smacro(s);
        if (fabs(s->e_vm) < 1E-5) s->Err_e = 0.0; else s->Err_e /= s->e_vm;
        s->Err_s /= s->s_vm;
        if (s->opt.verbose) printf("iter %d Err_e %e Err_s %e\n", iter, s->Err_e, s->Err_s);
    }
    r.converged = !(fmax(fabs(s->Err_s), fabs(s->Err_e)) > s->err);

    /* evolution.c:1268-1301 */
    {
        double act[3][3], vga[3][3];
        for (p = 0; p < N; p++) for (i = 0; i < 9; i++) s->velgrad[9 * p + i] = (s->disgrad[9 * p + i] - s->velgrad[9 * p + i]) / dt;
        for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) {
            act[i][j] = s->DisGradAvg[i][j] + s->dDisGradAvg_acum[i][j];
            vga[i][j] = (act[i][j] - s->DisGradAvg_t[i][j]) / dt;
        }
        s->e_vm = von_mises(act);
        s->d_vm = von_mises(vga);
        s->TimeTot += dt;
        memcpy(s->DisGradAvg_t, act, sizeof(act));
        memcpy(r.disgrad_avg_actual, act, sizeof(act));
        memcpy(r.velgrad_avg, vga, sizeof(vga));
    }
    for (p = 0; p < 9 * N; p++) s->eps[p] += s->edot[p] * dt;
    {
        double ea[9] = {0}, da[9] = {0}, evmp = 0.0, dvmp = 0.0;
        for (p = 0; p < N; p++) for (i = 0; i < 9; i++) { ea[i] += s->eps[9 * p + i] * s->wgt; da[i] += s->edot[9 * p + i] * s->wgt; }
        for (i = 0; i < 9; i++) { evmp += ea[i] * ea[i]; dvmp += da[i] * da[i]; }
        r.evmp = sqrt(2. / 3. * evmp);
        r.dvmp = sqrt(2. / 3. * dvmp);
        memcpy(r.eps_avg, ea, sizeof(ea));
        memcpy(r.edot_avg, da, sizeof(da));
    }
    if (s->dk.Update_Flag == 1 && s->istep > 1) update_orient(s);
    if (s->opt.gnd && s->dk.Hard_Flag == 1 && s->istep > 1) gradient_shear_rate(s);   /* evolution.c:1346-1353 */
    dislocation_evolution(s, s->istep);
    r.rho_F_avg = avg12(s, s->rho_F);
    r.rho_P_avg = avg12(s, s->rho_P);
    r.rho_s_avg = avg12(s, s->rho_s);
    r.rho_m_avg = avg12(s, s->rho_m);
    r.rho_g_avg = 0.0;
    if (s->opt.gnd) {   /* AvgDislocation, GNDs (constitutive.c:1745-1757) */
        double acc = 0.0;
        for (p = 0; p < N; p++) {
            int jph = s->phase[p];
            if (!is_xtal(s, jph)) continue;
            for (i = 0; i < s->ph[jph - 1].nsys; i++) {
                const long q = 12 * p + i;
                acc += sqrt(s->rho_g1[q] * s->rho_g1[q] + s->rho_g2[q] * s->rho_g2[q] + s->rho_g3[q] * s->rho_g3[q]);
            }
        }
        r.rho_g_avg = acc * (1.0 / N);
    }
    r.rho_dipole_avg = avg12(s, s->rho_dipole);
    r.rho_forest_avg = avg12(s, s->rho_forest);
    r.rho_inplane_avg = avg12(s, s->rho_inplane);
    if (s->fdd) {
        fprintf(s->fdd, "%e\t%e\t%e\t%e\t%e\t%e\t%e\n", s->TimeTot, s->e_vm, r.rho_s_avg, r.rho_m_avg,
                r.rho_dipole_avg, r.rho_inplane_avg, r.rho_forest_avg);
        fflush(s->fdd);
    }
    r.istep = s->istep;
    r.iters = iter;
    r.nr_iters_max = nrmax;
    r.nr_iters_total = nrtot;
    r.nr_wave_iters = nrwave;
    r.err_s = s->Err_s;
    r.err_e = s->Err_e;
    r.time_tot = s->TimeTot;
    r.e_vm = s->e_vm;
    r.d_vm = s->d_vm;
    r.s_vm = s->s_vm;
    r.s_vm1 = s->s_vm1;
    memcpy(r.sig_avg, s->SigAvg, sizeof(r.sig_avg));
    memcpy(r.sig_avg1, s->SigAvg1, sizeof(r.sig_avg1));
    if (s->fvm) {
        fprintf(s->fvm, "%e\t%e\t%e\t%e\t%e\t%e\t%e\n", s->TimeTot, s->e_vm, r.evmp, s->d_vm, r.dvmp, s->s_vm, s->s_vm1);
        fflush(s->fvm);
    }
    if (s->ferr) fflush(s->ferr);
    /* evolution.c:1513-1515 */
    for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) s->DisGradAvg[i][j] = s->DisGradAvg_t[i][j] + s->Udot[i][j] * dt;
    s->istep++;
    s->last = r;
    if (rep) *rep = r;
    return 0;
}

int orc_run(orc_state *s)
{
    while (s->istep < s->nsteps) {
        int rc = orc_step(s, 0);
        if (rc == 2007) break;
        if (rc) return rc;
    }
    return 0;
}

static double *field_ptr(orc_state *s, int f, int *nc)
{
    switch (f) {
    case ORC_F_SIG: *nc = 9; return s->sig;
    case ORC_F_EPS: *nc = 9; return s->eps;
    case ORC_F_EDOT: *nc = 9; return s->edot;
    case ORC_F_DISGRAD: *nc = 9; return s->disgrad;
    case ORC_F_VELGRAD: *nc = 9; return s->velgrad;
    case ORC_F_TRANMAT: *nc = 9; return s->tran;
    case ORC_F_RHO_S: *nc = 12; return s->rho_s;
    case ORC_F_RHO_M: *nc = 12; return s->rho_m;
    case ORC_F_RHO_F: *nc = 12; return s->rho_F;
    case ORC_F_RHO_P: *nc = 12; return s->rho_P;
    case ORC_F_GAMDOT: *nc = 12; return s->gamdot;
    case ORC_F_GAM_ACUM: *nc = 12; return s->gam_acum;
    case ORC_F_RHO_DIPOLE: *nc = 12; return s->rho_dipole;
    case ORC_F_RHO_FOREST: *nc = 12; return s->rho_forest;
    case ORC_F_RHO_INPLANE: *nc = 12; return s->rho_inplane;
    case ORC_F_NR_ITERS: *nc = 1; return s->nrit;
    case ORC_F_FLUCT: *nc = 9; return s->fluct;
    case ORC_F_TRIAL_GAM_ACUM: *nc = 12; return s->trial_gam_acum;
    case ORC_F_RHO_TOT: *nc = 1; return s->rho_tot;
    case ORC_F_RHO_G1: *nc = 12; return s->rho_g1;
    case ORC_F_RHO_G2: *nc = 12; return s->rho_g2;
    default: return 0;
    }
}

int orc_get_field(orc_state *s, int f, double *out, long n)
{
    int nc = 0;
    long p;
    double *src;
    if (f == ORC_F_PHASE || f == ORC_F_GRAIN) {
        if (n < s->nxyz) return 2008;
        for (p = 0; p < s->nxyz; p++) out[p] = f == ORC_F_PHASE ? s->phase[p] : s->grain[p];
        return 0;
    }
    src = field_ptr(s, f, &nc);
    if (!src) return 2009;
    if (n < nc * s->nxyz) return 2008;
    memcpy(out, src, sizeof(double) * nc * s->nxyz);
    return 0;
}

int orc_set_field(orc_state *s, int f, const double *in, long n)
{
    int nc = 0;
    double *dst = field_ptr(s, f, &nc);
    if (!dst) return 2009;
    if (n < nc * s->nxyz) return 2008;
    memcpy(dst, in, sizeof(double) * nc * s->nxyz);
    if (f == ORC_F_RHO_S) memcpy(s->trial_rho_s, in, sizeof(double) * nc * s->nxyz);
    return 0;
}

int orc_get_setup(orc_state *s, int which, double *out, long n)
{
    int i;
    switch (which) {
    case 0: if (n < 36) return 2008; memcpy(out, s->C066, sizeof(s->C066)); return 0;
    case 1: if (n < 36) return 2008; memcpy(out, s->S066, sizeof(s->S066)); return 0;
    case 2: if (n < 5) return 2008;
        out[0] = s->kT; out[1] = s->ph[0].q_slip; out[2] = s->ph[0].q_bulk; out[3] = s->ph[0].bburg; out[4] = s->ph[0].shear_g;
        return 0;
    case 3: if (n < 60) return 2008; memcpy(out, s->ph[0].schm_xt, sizeof(s->ph[0].schm_xt)); return 0;
    case 4: if (n < 144) return 2008; memcpy(out, s->ph[0].chi, sizeof(s->ph[0].chi)); return 0;
    case 5: if (n < 144) return 2008; memcpy(out, s->ph[0].chi1, sizeof(s->ph[0].chi1)); return 0;
    case 6: if (n < 4) return 2008; out[0] = s->ph[0].c11; out[1] = s->ph[0].c12; out[2] = s->ph[0].c44; out[3] = s->ph[0].nu; return 0;
    case 7: if (n < 9) return 2008; memcpy(out, s->Udot, sizeof(s->Udot)); return 0;
    case 8: if (n < 9) return 2008; memcpy(out, s->Scauchy, sizeof(s->Scauchy)); return 0;
    case 9: if (n < 2) return 2008; out[0] = s->d_vm0; out[1] = s->e_vm0; return 0;
    case 10: if (n < 1) return 2008; out[0] = s->wgt_ph1; return 0;
    case 11: if (n < 81) return 2008; memcpy(out, s->C0, sizeof(s->C0)); return 0;
    case 12: if (n < 81) return 2008; memcpy(out, s->S0, sizeof(s->S0)); return 0;
    case 13: if (n < 36) return 2008; for (i = 0; i < 36; i++) out[i] = s->cgr[i]; return 0;
    default: return 2009;
    }
}

void orc_close(orc_state *s)
{
    if (!s) return;
    if (s->fvm) fclose(s->fvm);
    if (s->fdd) fclose(s->fdd);
    if (s->fdrx) fclose(s->fdrx);
    if (s->ferr) fclose(s->ferr);
    free(s->tran); free(s->cgr); free(s->schm); free(s->sig); free(s->eps); free(s->edot); free(s->disgrad);
    free(s->velgrad); free(s->fluct); free(s->rho_s); free(s->rho_m); free(s->rho_F); free(s->rho_P);
    free(s->trial_rho_s); free(s->gamdot); free(s->gam_acum); free(s->trial_gam_acum); free(s->rho_dipole);
    free(s->rho_forest); free(s->rho_inplane); free(s->rho_tot); free(s->nrit); free(s->work); free(s->ksig);
    free(s->dsig);
    free(s->nvec); free(s->conv_s); free(s->conv_e); free(s->phase); free(s->grain);
    free(s->rho_g1); free(s->rho_g2); free(s->rho_g3); free(s->gadot_grad); free(s->gbc);
    free(s);
}
