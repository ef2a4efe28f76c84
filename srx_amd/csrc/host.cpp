// host.cpp -- libsrx host side: input readers, setup, increment driver, C ABI.
//
// Replaces the reference's step driver (evp.c main -> GetNameList, SetupJob,
// Evolution, DestroyJob).  The grid stays resident on one GPU; the host only
// holds scalars between kernels (the 20 reduced values of each outer
// iteration) and does the macro-stress / mixed-boundary-condition algebra of
// get_smacro (evolution.c:590-664) on them.
#include "srx_dev.hpp"
#include "../../include/srx.h"

#include <hip/hip_runtime.h>
#include <omp.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <atomic>
#include <thread>
#include <vector>

namespace srx {
void launch_fft_real_inverse_test(hipStream_t s, int prec, void *specC, void *specB, int ncomp, double *out,
                                  const int n[3], const FftPlan *pl, void *const *tw, double divisor);
}

using namespace srx;

namespace {

// ---------------------------------------------------------------------------
// constants of evp.h:84-94
// ---------------------------------------------------------------------------
// outer iterations enqueued between two read-backs of the control block
constexpr int ITER_BATCH = 8;
constexpr double PI_REF = 3.14159265359;
constexpr double EV_IN_J = 1.60217733e-19;
constexpr double J_IN_EV = (1.0 / EV_IN_J);
constexpr double MY_J_IN_EV = 0.62415063631;
constexpr double BOLZ = 1.380658e-23;
constexpr double R2 = 0.70710678118654744, R3 = 0.57735026918962584, R6 = 0.40824829046386304;

// B basis (kinematics.c:35-53)
struct Basis {
    double b[6][3][3];
    Basis() {
        std::memset(b, 0, sizeof(b));
        b[0][0][0] = -R2; b[0][1][1] = R2;
        b[1][0][0] = -R6; b[1][1][1] = -R6; b[1][2][2] = 2. * R6;
        b[2][1][2] = R2; b[2][2][1] = R2;
        b[3][0][2] = R2; b[3][2][0] = R2;
        b[4][0][1] = R2; b[4][1][0] = R2;
        b[5][0][0] = R3; b[5][1][1] = R3; b[5][2][2] = R3;
    }
};
const Basis BB;

void t9_to_b6(const double t[9], double v[6]) {
    for (int i = 0; i < 6; ++i) {
        double s = 0.0;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) s += t[3 * a + b] * BB.b[i][a][b];
        v[i] = s;
    }
}
void b6_to_t9(const double v[6], double t[9], int dim = 6) {
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            double s = 0.0;
            for (int i = 0; i < dim; ++i) s += v[i] * BB.b[i][a][b];
            t[3 * a + b] = s;
        }
}
// 6x6 B-basis matrix -> 4th-order tensor (chg_basis opt 3)
void c66_to_c4(const double c[36], double t[81]) {
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            for (int c3 = 0; c3 < 3; ++c3)
                for (int d = 0; d < 3; ++d) {
                    double s = 0.0;
                    for (int i = 0; i < 6; ++i)
                        for (int j = 0; j < 6; ++j) s += c[6 * i + j] * BB.b[i][a][b] * BB.b[j][c3][d];
                    t[((a * 3 + b) * 3 + c3) * 3 + d] = s;
                }
}


// Gauss-Jordan inverse with partial pivoting (n <= 6)
bool invert(double *A, int n) {
    double M[6][12];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < 2 * n; ++j) M[i][j] = j < n ? A[n * i + j] : (double)(j - n == i);
    for (int c = 0; c < n; ++c) {
        int p = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(M[r][c]) > std::fabs(M[p][c])) p = r;
        if (M[p][c] == 0.0) return false;
        if (p != c)
            for (int j = 0; j < 2 * n; ++j) std::swap(M[p][j], M[c][j]);
        const double r = 1.0 / M[c][c];
        for (int j = 0; j < 2 * n; ++j) M[c][j] *= r;
        for (int i = 0; i < n; ++i) {
            if (i == c) continue;
            const double f = M[i][c];
            if (f != 0.0)
                for (int j = 0; j < 2 * n; ++j) M[i][j] -= f * M[c][j];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) A[n * i + j] = M[i][j + n];
    return true;
}

double von_mises(const double t[9]) {   // kinematics.c:465-483
    const double tr = t[0] + t[4] + t[8];
    double vm = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const double d = (t[3 * i + j] + t[3 * j + i]) / 2. - (double)(i == j) * tr / 3.0;
            vm += d * d;
        }
    return std::sqrt(2. / 3. * vm);
}

void voigt_to_full(const double *v, double m[9]) {   // kinematics.c:5-12
    m[0] = v[0]; m[1] = v[5]; m[2] = v[4];
    m[3] = v[5]; m[4] = v[1]; m[5] = v[3];
    m[6] = v[4]; m[7] = v[3]; m[8] = v[2];
}

// opt 1: matrix -> angles in degrees
void matrix_to_euler(const double a[9], double &t1, double &ph, double &t2) {
    double p = std::acos(a[8]), x1, x2;
    if (std::fabs(a[8]) > 0.9999) {
        x2 = 0.;
        x1 = std::atan2(a[1], a[0]);
    } else {
        const double sph = std::sin(p);
        x1 = std::atan2(a[6] / sph, -1.0 * a[7] / sph);
        x2 = std::atan2(a[2] / sph, a[5] / sph);
    }
    ph = p * 180. / PI_REF;
    t1 = x1 * 180. / PI_REF;
    t2 = x2 * 180. / PI_REF;
}

// ---------------------------------------------------------------------------
// input deck: namelist semantics of io.c:96-164
// ---------------------------------------------------------------------------
struct Deck {
    int CellDim[3] = {0, 0, 0}, N_phases = 0, Type_phases[2] = {0, 0}, N_steps = 0, Update_Flag = 0, Hard_Flag = 0;
    int Tex_Flag = 0, IterMax = 0, PrintControl[2] = {0, 0}, VelGrad_BC_Flag[6] = {0}, Stress_BC_Flag[6] = {0};
    int ElastBC = 0, nucleus_radius = 0, CellDim_pf[3] = {0}, RandSeed = 0, Nucl_Static_Flag = 0;
    double TimeStep = 0, Err = 0, VelGrad_BC[6] = {0}, Stress_BC[6] = {0}, ElastConst_PhaseI[4] = {0};
    double ElastConst_PhaseII[4] = {0};
    char Slip_PhaseI[100] = {0}, Slip_PhaseII[100] = {0}, initial_ms[100] = {0};
    double a0[2] = {0}, L0 = 0, GND_ScaleRatio = 0, T__K = 0, Q_slip[2] = {0}, Q_bulk[2] = {0};
    double C_1[2] = {0}, C_2[2] = {0}, C_3[2] = {0}, C_4[2] = {0}, C_5[2] = {0}, C_6[2] = {0}, C_7[2] = {0};
    double C_8[2] = {0};
    double Selfinter0[2] = {0}, Coplanar0[2] = {0}, CrossSlip0[2] = {0}, GlissileJunction0[2] = {0};
    double HirthLock0[2] = {0}, LomerCottrellLock0[2] = {0};
    double rho_SSD_initial = 0, D_0 = 0, E_gb = 0, M_gb = 0, Scale_Fdeform = 0, k_c = 0, alpha = 0, zeta_kc = 0;
    double age_drx = 0, ssdScaler_DRX = 0, gndScaler_DRX = 0, sigScaler_DRX = 0, tau_DRX = 0, DeltaQ_DRX = 0;
    double DeltaTau_DRX = 0, pf_length_scale = 0, pf_time_step = 0;
};

struct Key {
    const char *name;
    void *ptr;
    char type;   // 'i', 'r', 's'
    int len;
};

int read_deck(const char *path, Deck &d) {
    std::vector<Key> keys = {
        {"CellDim", d.CellDim, 'i', 3}, {"N_phases", &d.N_phases, 'i', 1}, {"Type_phases", d.Type_phases, 'i', 2},
        {"TimeStep", &d.TimeStep, 'r', 1}, {"N_steps", &d.N_steps, 'i', 1}, {"Update_Flag", &d.Update_Flag, 'i', 1},
        {"Hard_Flag", &d.Hard_Flag, 'i', 1}, {"Tex_Flag", &d.Tex_Flag, 'i', 1}, {"IterMax", &d.IterMax, 'i', 1},
        {"PrintControl", d.PrintControl, 'i', 2}, {"Err", &d.Err, 'r', 1},
        {"VelGrad_BC_Flag", d.VelGrad_BC_Flag, 'i', 6}, {"VelGrad_BC", d.VelGrad_BC, 'r', 6},
        {"Stress_BC_Flag", d.Stress_BC_Flag, 'i', 6}, {"Stress_BC", d.Stress_BC, 'r', 6},
        {"ElastBC", &d.ElastBC, 'i', 1}, {"ElastConst_PhaseI", d.ElastConst_PhaseI, 'r', 4},
        {"ElastConst_PhaseII", d.ElastConst_PhaseII, 'r', 4}, {"Slip_PhaseI", d.Slip_PhaseI, 's', 1},
        {"Slip_PhaseII", d.Slip_PhaseII, 's', 1}, {"initial_ms", d.initial_ms, 's', 1}, {"a0", d.a0, 'r', 2},
        {"L0", &d.L0, 'r', 1}, {"GND_ScaleRatio", &d.GND_ScaleRatio, 'r', 1}, {"T__K", &d.T__K, 'r', 1},
        {"Q_slip", d.Q_slip, 'r', 2}, {"Q_bulk", d.Q_bulk, 'r', 2}, {"C_1", d.C_1, 'r', 2}, {"C_2", d.C_2, 'r', 2},
        {"C_3", d.C_3, 'r', 2}, {"C_4", d.C_4, 'r', 2}, {"C_5", d.C_5, 'r', 2}, {"C_6", d.C_6, 'r', 2},
        {"C_7", d.C_7, 'r', 2}, {"C_8", d.C_8, 'r', 2}, {"Selfinter0", d.Selfinter0, 'r', 2},
        {"Coplanar0", d.Coplanar0, 'r', 2}, {"CrossSlip0", d.CrossSlip0, 'r', 2},
        {"GlissileJunction0", d.GlissileJunction0, 'r', 2}, {"HirthLock0", d.HirthLock0, 'r', 2},
        {"LomerCottrellLock0", d.LomerCottrellLock0, 'r', 2}, {"rho_SSD_initial", &d.rho_SSD_initial, 'r', 1},
        {"D_0", &d.D_0, 'r', 1}, {"nucleus_radius", &d.nucleus_radius, 'i', 1}, {"CellDim_pf", d.CellDim_pf, 'i', 3},
        {"RandSeed", &d.RandSeed, 'i', 1}, {"E_gb", &d.E_gb, 'r', 1}, {"M_gb", &d.M_gb, 'r', 1},
        {"Scale_Fdeform", &d.Scale_Fdeform, 'r', 1}, {"k_c", &d.k_c, 'r', 1},
        {"Nucl_Static_Flag", &d.Nucl_Static_Flag, 'i', 1}, {"alpha", &d.alpha, 'r', 1},
        {"zeta_kc", &d.zeta_kc, 'r', 1}, {"age_drx", &d.age_drx, 'r', 1}, {"ssdScaler_DRX", &d.ssdScaler_DRX, 'r', 1},
        {"gndScaler_DRX", &d.gndScaler_DRX, 'r', 1}, {"sigScaler_DRX", &d.sigScaler_DRX, 'r', 1},
        {"tau_DRX", &d.tau_DRX, 'r', 1}, {"DeltaQ_DRX", &d.DeltaQ_DRX, 'r', 1},
        {"DeltaTau_DRX", &d.DeltaTau_DRX, 'r', 1}, {"pf_length_scale", &d.pf_length_scale, 'r', 1},
        {"pf_time_step", &d.pf_time_step, 'r', 1},
    };
    std::vector<char> seen(keys.size(), 0);
    FILE *fp = std::fopen(path, "r");
    if (!fp) return SRX_E_FILE;
    char buff[200];
    while (std::fgets(buff, 200, fp)) {
        if (std::feof(fp)) break;                 // a last line without a newline is not read
        if (buff[0] == '#' || buff[0] == '\n') continue;
        char *save = nullptr;
        char *tok = strtok_r(buff, " \t\n", &save);
        if (!tok) break;                          // a blank-but-whitespace line ends the deck
        for (size_t k = 0; k < keys.size(); ++k) {
            if (std::strcmp(tok, keys[k].name) != 0) continue;
            if (!seen[k]) {                       // later duplicates are ignored
                seen[k] = 1;
                for (int j = 0; j < keys[k].len; ++j) {
                    char *t = strtok_r(nullptr, ", \t\n", &save);
                    if (!t) break;
                    if (keys[k].type == 'i') static_cast<int *>(keys[k].ptr)[j] = (int)std::atol(t);
                    else if (keys[k].type == 'r') static_cast<double *>(keys[k].ptr)[j] = std::atof(t);
                    else { std::snprintf(static_cast<char *>(keys[k].ptr), 100, "%s", t); }
                }
            }
            break;
        }
    }
    std::fclose(fp);
    return 0;
}

// ---------------------------------------------------------------------------
// slip-mode file: PlasticInit (init.c:52-231)
// ---------------------------------------------------------------------------
struct PhaseHost {
    int xtal = 0, nsys = 0;
    double c11 = 0, c12 = 0, c44 = 0, nu = 0, G = 0;
    double bn[NSYS][3] = {{0}}, bb[NSYS][3] = {{0}}, p[NSYS][5] = {{0}};
    double tau0xf = 0, b = 0;
};

int read_sx(const char *path, PhaseHost &P) {
    FILE *fp = std::fopen(path, "r");
    if (!fp) return SRX_E_FILE;
    char buf[80];
    auto gl = [&]() { if (!std::fgets(buf, 80, fp)) buf[0] = 0; };
    int nmodes_max = 0, nmodes = 0, imode[16] = {0};
    for (int i = 0; i < 7; ++i) gl();
    gl(); std::sscanf(buf, "%d", &nmodes_max);
    gl(); gl();
    gl(); std::sscanf(buf, "%d", &nmodes);
    nmodes = std::min(nmodes, 16);
    gl(); gl();
    for (int i = 0; i < nmodes; ++i) { gl(); std::sscanf(buf, "%d", &imode[i]); }
    gl();
    P.nsys = 0;
    for (int m = 0; m < nmodes_max; ++m) {
        int mx = -1, nsmx = 0, isect = 0;
        double nrsx, gamd0x, twshx, tau0xf = 0, tau0xb, tau1x, thet0 = 0, thet1 = 0, hs, hl;
        gl(); std::sscanf(buf, "%*s %*s %d %*s", &mx);
        bool match = false;
        for (int j = 0; j < nmodes; ++j)
            if (mx == imode[j]) { match = true; break; }
        gl(); gl(); std::sscanf(buf, "%d %lf %lf %lf %d", &nsmx, &nrsx, &gamd0x, &twshx, &isect);
        gl(); gl(); std::sscanf(buf, "%lf %lf %lf %lf %lf", &tau0xf, &tau0xb, &tau1x, &thet0, &thet1);
        gl(); gl(); std::sscanf(buf, "%lf %lf", &hs, &hl);
        gl();
        if (match) {
            if (nsmx > NSYS) { std::fclose(fp); return SRX_E_SYSTEMS; }
            P.nsys = nsmx;
            if (thet0 < thet1) { std::fclose(fp); return 230; }
            P.tau0xf = tau0xf;
            for (int j = 0; j < nsmx; ++j) {
                gl();
                std::sscanf(buf, "%lf %lf %lf %lf %lf %lf", &P.bn[j][0], &P.bn[j][1], &P.bn[j][2], &P.bb[j][0],
                            &P.bb[j][1], &P.bb[j][2]);
            }
        } else {
            for (int j = 0; j < nsmx; ++j) gl();
        }
    }
    std::fclose(fp);
    for (int j = 0; j < P.nsys; ++j) {
        double nn = std::sqrt(std::pow(P.bn[j][0], 2.0) + std::pow(P.bn[j][1], 2.0) + std::pow(P.bn[j][2], 2.0));
        for (int a = 0; a < 3; ++a) P.bn[j][a] /= nn;
        nn = std::sqrt(std::pow(P.bb[j][0], 2.0) + std::pow(P.bb[j][1], 2.0) + std::pow(P.bb[j][2], 2.0));
        for (int a = 0; a < 3; ++a) P.bb[j][a] /= nn;
        const double dot = P.bn[j][0] * P.bb[j][0] + P.bn[j][1] * P.bb[j][1] + P.bn[j][2] * P.bb[j][2];
        if (dot > 1E-3) return 7;
        double s[9], v[6];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) s[3 * a + b] = (P.bn[j][a] * P.bb[j][b] + P.bn[j][b] * P.bb[j][a]) / 2.0;
        t9_to_b6(s, v);
        for (int a = 0; a < 5; ++a) P.p[j][a] = v[a];
    }
    return 0;
}

// Arsenlis-Parks table (init.c:594-703): S self, C coplanar, X cross slip,
// G glissile, H Hirth, L Lomer; the upper triangle row by row.
const char *const ISM[12] = {"SCCHGLXGGHLG", "SCGXGGHLLHG", "SLGHGLHGGX", "SCCHLGXGG", "SCLHGGHL", "SGGXGLH",
                             "SCCHGL",       "SCGXG",       "SLGH",       "SCC",       "SC",       "S"};
void interaction(double self, double cop, double cross, double gli, double hir, double lom, double chi[NSYS][NSYS]) {
    for (int i = 0; i < 12; ++i)
        for (int j = i; j < 12; ++j) {
            const char c = ISM[i][j - i];
            const double v = c == 'S' ? self : c == 'C' ? cop : c == 'X' ? cross : c == 'G' ? gli : c == 'H' ? hir : lom;
            chi[i][j] = chi[j][i] = v;
        }
}

}  // namespace

// ---------------------------------------------------------------------------
// the solver context
// ---------------------------------------------------------------------------
struct srx_ctx {
    srx_options opt;
    Deck dk;
    std::string workdir;
    int n[3] = {0, 0, 0};
    long nvox = 0;
    double wgt = 0, wgt_ph1 = 0;
    PhaseHost ph[NPH];
    DevConst hc;
    DevConst *dc = nullptr;
    Fields f;
    void *specB = nullptr, *specC = nullptr;
    void *tw[3] = {nullptr, nullptr, nullptr};
    FftPlan pl[3];
    double *partials = nullptr, *red = nullptr, *hred = nullptr, *redseg = nullptr;
    IterCtl *dctl = nullptr, *hctl = nullptr;   // outer-iteration control block (device, pinned host copy)
    int iters_last = 0;                        // outer iterations of the last increment
    int nr_blocks = 0, end_blocks = 0;
    hipStream_t stream = nullptr;
    // macro state
    double Udot[9], Udot_s[9], Udot_a[9], Scauchy[9];
    double DisGradAvg[9], DisGradAvg_t[9], dDisGradAvg[9], acum[9], uniform[9];
    double SigAvg[9], SigAvg1[9];
    double C066[36], S066[36];
    double e_vm = 0, d_vm = 0, s_vm = 0, s_vm1 = 0, TimeTot = 0, err = 0, Err_e = 0, Err_s = 0;
    double sig_sum6[6] = {0, 0, 0, 0, 0, 0};
    int itermax = 0, nsteps = 0, istep = 0, creep = 0;
    std::vector<int> grain;
    FILE *fvm = nullptr, *fdd = nullptr, *ferr = nullptr, *fdrx = nullptr;
    // timing
    int timing = 0;   // 0 off, 1 every kernel, 2 the Newton kernel only
    std::vector<hipEvent_t> evpool;
    std::vector<int> evslot, evit;
    size_t evused = 0;
    struct KT { const char *name; double ms; int calls; };
    KT kt[8];
    int nkt = 0;
    // learned Newton start: packed network, and the training-sample recorder
    MlpPacked *mlp = nullptr;
    float *mlp_blob = nullptr;
    float *rec_x = nullptr, *rec_t = nullptr, *rec_nd = nullptr;
    int *rec_count = nullptr;
    int rec_stride = 0, rec_cap = 0;
    // outer-iteration batches captured as HIP graphs: [first batch of an increment][rotated]
    hipGraphExec_t gexec[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    bool graphs = true;
    // rows per block of the five FFT passes (0: the default) and their tuned times
    int tiles[5] = {0, 0, 0, 0, 0};
    double tile_us[5] = {0, 0, 0, 0, 0};
};

namespace {

#define HIPCK(x)                                                                                         \
    do {                                                                                                 \
        hipError_t e_ = (x);                                                                             \
        if (e_ != hipSuccess) {                                                                          \
            std::fprintf(stderr, "srx: %s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
            return e_ == hipErrorOutOfMemory ? SRX_E_NOMEM : SRX_E_HIP;                                  \
        }                                                                                                \
    } while (0)

std::string resolve(const std::string &dir, const char *name) {
    if (name[0] == '/' || dir.empty()) return name;
    return dir + "/" + name;
}

// The passes' limits, as the launchers size them: a y or z pass holds one pencil (padded pitch
// lds_pitch(n)) and the twiddle table in 48 KB of LDS with at most 2 butterflies per thread; the Green x
// pass (LDS kernel, one row per block) holds 9 component rows and the twiddle table, 10 rows of n0
// complex values (16 B FP64, 8 B mixed), with at most 4 butterflies per thread.
bool grid_supported(const FftPlan pl[3], const int n[3], int prec) {
    const long csz = prec == SRX_PREC_MIXED ? 8 : 16;
    for (int ax = 0; ax < 3; ++ax) {
        const FftPlan &p = pl[ax];
        int mr = 1 << 30;
        for (int i = 0; i < p.nst; ++i) mr = std::min(mr, p.rad[i]);
        if (p.nst == 0) mr = 1;
        const int slots = ax == 0 ? 9 : 1;
        const int maxb = ax == 0 ? 4 : 2;   // butterflies per thread of the x pass and of the y/z passes
        if ((long)slots * p.n / std::max(mr, 1) > (long)maxb * 256) return false;
        const long lds = ax == 0 ? 10L * p.n * csz : ((long)lds_pitch(p.n) + p.n) * csz;
        if (lds > 48 * 1024) return false;
    }
    if (n[0] < 1 || n[1] < 1 || n[2] < 1) return false;
    // the per-voxel kernels address an element of a field by a 32-bit byte offset from its component
    return (long long)n[0] * n[1] * n[2] <= SRX_MAX_VOXELS;
}

}  // namespace

bool srx::srx_grid_supported(const FftPlan pl[3], const int n[3], int prec) { return grid_supported(pl, n, prec); }

namespace {

template <typename T>
int upload_twiddles(int n, void **dst, hipStream_t st = nullptr) {
    std::vector<T> h(2 * std::max(n, 1));
    for (int m = 0; m < n; ++m) {
        const long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)n;
        h[2 * m] = (T)std::cos(ang);
        h[2 * m + 1] = (T)std::sin(ang);
    }
    HIPCK(hipMalloc(dst, h.size() * sizeof(T)));
    HIPCK(hipMemcpyAsync(*dst, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
    HIPCK(hipStreamSynchronize(st));
    return 0;
}

// Copies and fills of a context's buffers, ordered on its stream and complete on return.  The
// context streams do not synchronise with the legacy stream, so one context's setup never waits
// for (or breaks the graph capture of) another context working concurrently.
int ctx_copy(hipStream_t st, void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
    HIPCK(hipMemcpyAsync(dst, src, bytes, kind, st));
    HIPCK(hipStreamSynchronize(st));
    return 0;
}
int ctx_zero(hipStream_t st, void *dst, size_t bytes) {
    HIPCK(hipMemsetAsync(dst, 0, bytes, st));
    HIPCK(hipStreamSynchronize(st));
    return 0;
}

int setup(srx_ctx *c) {
    Deck &d = c->dk;
    for (int i = 0; i < 3; ++i) c->n[i] = d.CellDim[i];
    for (int i = 0; i < 3; ++i)
        if (!make_plan(c->n[i], c->pl[i])) return SRX_E_GRID;
    if (!grid_supported(c->pl, c->n, c->opt.precision)) return SRX_E_GRID;
    c->nvox = (long)c->n[0] * c->n[1] * c->n[2];
    c->wgt = 1. / c->nvox;
    const double kT = (d.T__K * BOLZ * J_IN_EV);
    if (d.Update_Flag != 1) return SRX_E_UPDATE_FLAG;
    // istep % PrintControl[1] at evolution.c:1456 divides by zero for a period <= 0
    if (c->opt.write_fields && d.PrintControl[0] == 1 && d.PrintControl[1] <= 0) return SRX_E_ARG;
    // the GND option runs with the reference's Newton start only
    if (c->opt.gnd && c->opt.warm_start != 0) return SRX_E_ARG;

    // phases (init.c:815-856)
    std::memset(&c->hc, 0, sizeof(c->hc));
    for (int p = 0; p < NPH; ++p) c->ph[p].xtal = (p < d.N_phases) && d.Type_phases[p] == 0;
    if (c->ph[0].xtal) {   // Cu constants at T__K override the deck (init.c:822-826)
        d.ElastConst_PhaseI[0] = (1.818 - 0.0004045 * d.T__K) * 100 * 1000;
        d.ElastConst_PhaseI[1] = (1.285 - 0.0002 * d.T__K) * 100 * 1000;
        d.ElastConst_PhaseI[2] = (0.8396 - 0.0002711 * d.T__K) * 100 * 1000;
    }
    for (int p = 0; p < NPH; ++p) {
        PhaseHost &P = c->ph[p];
        DevPhase &D = c->hc.ph[p];
        D.xtal = P.xtal;
        // interaction matrices exist for every phase (init.c:1100-1116)
        double chi[NSYS][NSYS], chi1[NSYS][NSYS];
        interaction(d.Selfinter0[p], d.Coplanar0[p], 0.0, 0.0, 0.0, 0.0, chi);
        interaction(0.0, 0.0, d.CrossSlip0[p], d.GlissileJunction0[p], d.HirthLock0[p], d.LomerCottrellLock0[p], chi1);
        if (!P.xtal) continue;
        const double *E = p == 0 ? d.ElastConst_PhaseI : d.ElastConst_PhaseII;
        P.c11 = E[0]; P.c12 = E[1]; P.c44 = E[2]; P.nu = E[3];
        P.G = 0.1 * (E[0] - E[1] + 3 * E[2]) + (0.5) * (5 * (E[0] - E[1]) * E[2]) / (4 * E[2] + 3 * (E[0] - E[1]));
        const std::string sx = resolve(c->workdir, p == 0 ? d.Slip_PhaseI : d.Slip_PhaseII);
        int rc = read_sx(sx.c_str(), P);
        if (rc) return rc;
        P.b = d.a0[p] / std::sqrt(2.0);
        D.nsys = P.nsys;
        for (int s = 0; s < NSYS; ++s) {
            for (int k = 0; k < 5; ++k) D.p[s][k] = s < P.nsys ? P.p[s][k] : 0.0;
            for (int k = 0; k < 3; ++k) { D.bn[s][k] = P.bn[s][k]; D.bb[s][k] = P.bb[s][k]; }
            for (int j = 0; j < NSYS; ++j) {
                const bool in = s < P.nsys && j < P.nsys;
                D.chi[s][j] = in ? chi[s][j] : 0.0;
                D.chi1[s][j] = in ? chi1[s][j] : 0.0;
            }
        }
        const double cd[6] = {P.c11 - P.c12, P.c11 - P.c12, 2 * P.c44, 2 * P.c44, 2 * P.c44, P.c11 + 2 * P.c12};
        for (int k = 0; k < 6; ++k) D.sinv[k] = 1.0 / cd[k];
        const double c1 = d.C_1[p], c2 = d.C_2[p], c3 = d.C_3[p];
        D.c1Gb = c1 * P.G * P.b;
        D.kTf = kT / MY_J_IN_EV;
        D.c2 = c2; D.c3 = c3; D.b = P.b;
        D.tau0 = c->opt.tau0_mode == 1 ? c->opt.tau0_value : P.tau0xf;
        D.rhom_k = 2.0 * kT / c1 / c2 / c3 / P.G / (std::pow(P.b, 3.0));
        D.expmq = std::exp(-1.0 * (d.Q_slip[p] * MY_J_IN_EV / kT));
        D.G = P.G;
        D.c4 = d.C_4[p]; D.c5 = d.C_5[p]; D.c6 = d.C_6[p];
        D.expmqb = std::exp(-1.0 * (d.Q_bulk[p] * MY_J_IN_EV / kT));
        D.c1Gb3 = c1 * P.G * P.b * 1.0E-3;
        D.k_tcut = (kT / MY_J_IN_EV) / c2 / c3 / P.b / P.b * 0.1;
        D.k_g0 = 1E-3 * 1.0E10 * c2 * P.b * D.expmq;
        // GND option: bp = bb x bn (init.c:200-202) and the projections of init.c:1118-1141
        for (int j = 0; j < NSYS; ++j) {
            const bool in = j < P.nsys;
            D.bp[j][0] = in ? P.bb[j][1] * P.bn[j][2] - P.bb[j][2] * P.bn[j][1] : 0.0;
            D.bp[j][1] = in ? P.bb[j][2] * P.bn[j][0] - P.bb[j][0] * P.bn[j][2] : 0.0;
            D.bp[j][2] = in ? P.bb[j][0] * P.bn[j][1] - P.bb[j][1] * P.bn[j][0] : 0.0;
        }
        for (int i = 0; i < NSYS; ++i)
            for (int j = 0; j < NSYS; ++j) {
                double c1g = 0.0, c2g = 0.0;
                if (i < P.nsys && j < P.nsys) {
                    const double sv[3] = {P.bn[j][1] * P.bb[j][2] - P.bn[j][2] * P.bb[j][1],
                                          P.bn[j][2] * P.bb[j][0] - P.bn[j][0] * P.bb[j][2],
                                          P.bn[j][0] * P.bb[j][1] - P.bn[j][1] * P.bb[j][0]};
                    c2g = std::fabs(P.bn[i][0] * sv[0] + P.bn[i][1] * sv[1] + P.bn[i][2] * sv[2]);
                    c1g = std::fabs(P.bn[i][0] * P.bb[j][0] + P.bn[i][1] * P.bb[j][1] + P.bn[i][2] * P.bb[j][2]);
                }
                D.cos_g1[i][j] = c1g;
                D.sin_g1[i][j] = std::fabs(std::sqrt(std::fabs(1.0 - c1g * c1g)));
                D.cos_g2[i][j] = c2g;
                D.sin_g2[i][j] = std::fabs(std::sqrt(std::fabs(1.0 - c2g * c2g)));
            }
    }
    c->hc.dt = d.TimeStep;
    c->hc.nu = d.ElastConst_PhaseI[3];
    c->hc.T = d.T__K;
    c->hc.D0 = d.D_0;
    c->hc.elast_bc = d.ElastBC;
    c->hc.gnd = c->opt.gnd ? 1 : 0;
    c->hc.L0 = d.L0;
    for (int i = 0; i < 3; ++i) c->hc.n[i] = c->n[i];

    // microstructure (init.c:233-396)
    const long N = c->nvox;
    const int nthr = c->opt.nthreads > 0 ? c->opt.nthreads : omp_get_max_threads();
    std::vector<int> phase(N, 0);
    std::vector<double> eul(3 * N);   // Bunge angles (degrees) per voxel, for the device setup
    c->grain.assign(N, 0);
    {
        // The reference reads N records with fgets(buf, 80) and sscanf (init.c:253-301).  When each
        // of the first N lines is shorter than 79 characters (so fgets returns whole lines, as
        // it does for every valid mesh) the lines are parsed in parallel from one read of the
        // file; otherwise record by record with fgets, as the reference does.
        const std::string ms = resolve(c->workdir, d.initial_ms);
        FILE *fp = std::fopen(ms.c_str(), "rb");
        if (!fp) return SRX_E_FILE;
        std::vector<char> text;
        {
            char chunk[1 << 16];
            size_t got;
            while ((got = std::fread(chunk, 1, sizeof(chunk), fp)) > 0) text.insert(text.end(), chunk, chunk + got);
        }
        std::fclose(fp);
        std::vector<size_t> start;
        start.reserve(N + 1);
        bool whole_lines = true;
        for (size_t pos = 0; (long)start.size() < N && pos < text.size();) {
            const char *nl = static_cast<const char *>(std::memchr(&text[pos], '\n', text.size() - pos));
            const size_t end = nl ? (size_t)(nl - &text[0]) : text.size();
            if (end - pos >= 79) whole_lines = false;
            start.push_back(pos);
            pos = end + 1;
        }
        if ((long)start.size() < N) return 1139;   // fgets hits EOF before N records
        std::vector<int> ijk(3 * N);
        auto parse = [&](const char *line, long v) {
            double t1 = 0, th = 0, om = 0;
            int ii = 0, jj = 0, kk = 0, jgr = 0, jph = 0;
            std::sscanf(line, "%lf %lf %lf %d %d %d %d %d", &t1, &th, &om, &ii, &jj, &kk, &jgr, &jph);
            eul[3 * v] = t1; eul[3 * v + 1] = th; eul[3 * v + 2] = om;
            ijk[3 * v] = ii; ijk[3 * v + 1] = jj; ijk[3 * v + 2] = kk;
            c->grain[v] = jgr;
            phase[v] = jph;
        };
        if (whole_lines) {
#pragma omp parallel for schedule(static) num_threads(nthr)
            for (long v = 0; v < N; ++v) {
                char buf[80];
                const size_t b0 = start[v];
                size_t len = (v + 1 < (long)start.size() ? start[v + 1] : text.size()) - b0;
                if (len > 79) len = 79;
                std::memcpy(buf, &text[b0], len);
                buf[len] = 0;
                parse(buf, v);
            }
        } else {
            FILE *f2 = std::fopen(ms.c_str(), "r");
            if (!f2) return SRX_E_FILE;
            char buf[80];
            for (long v = 0; v < N; ++v) {
                if (!std::fgets(buf, 80, f2)) { std::fclose(f2); return 1139; }
                parse(buf, v);
            }
            std::fclose(f2);
        }
        // record order and phase checks, in voxel order as the reference meets them
        long nph1 = 0;
        for (long v = 0; v < N; ++v) {
            const long idx = ((long)(ijk[3 * v] - 1) * c->n[1] + ijk[3 * v + 1] - 1) * c->n[2] + ijk[3 * v + 2] - 1;
            const int jph = phase[v];
            if (idx != v || jph < 1 || jph > NPH) return 1139;
            if (jph == 1) ++nph1;
        }
        c->wgt_ph1 = (1.0 * (double)nph1) / N;
    }
    // boundary conditions (init.c:865-901)
    {
        if (d.VelGrad_BC_Flag[0] + d.VelGrad_BC_Flag[1] + d.VelGrad_BC_Flag[2] == 2) return 116;
        double vg[6];
        for (int i = 0; i < 6; ++i) vg[i] = d.VelGrad_BC[i] * d.VelGrad_BC_Flag[i];
        voigt_to_full(vg, c->Udot);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                c->Udot_s[3 * i + j] = (c->Udot[3 * i + j] + c->Udot[3 * j + i]) / 2.;
                c->Udot_a[3 * i + j] = (c->Udot[3 * i + j] - c->Udot[3 * j + i]) / 2.;
            }
        double d6[6], dev[9];
        t9_to_b6(c->Udot_s, d6);
        b6_to_t9(d6, dev, 5);
        c->d_vm = 0.0;
        for (int k = 0; k < 9; ++k) c->d_vm += std::pow(dev[k], 2.0);
        c->d_vm = std::sqrt(2. / 3. * c->d_vm);
        for (int i = 0; i < 6; ++i)
            if ((d.Stress_BC_Flag[i] * d.VelGrad_BC_Flag[i] != 0) || (d.Stress_BC_Flag[i] + d.VelGrad_BC_Flag[i] != 1))
                return 114;
        c->creep = 0;
        for (int i = 0; i < 6; ++i) c->creep += d.Stress_BC_Flag[i];
        c->creep /= 6;
        voigt_to_full(d.Stress_BC, c->Scauchy);
        for (int k = 0; k < 9; ++k) c->hc.Udot_a[k] = c->Udot_a[k];
    }
    // InitFields (init.c:444-516): Sig = C_gr : (Udot dt)
    for (int k = 0; k < 9; ++k) {
        c->DisGradAvg[k] = c->Udot[k] * d.TimeStep;
        c->dDisGradAvg[k] = c->acum[k] = c->DisGradAvg_t[k] = c->uniform[k] = 0.0;
    }
    // per voxel on the device after allocation (launch_setup): Sig = C_gr : sym(Udot dt) (+ Scauchy)
    double e9[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) e9[3 * i + j] = (c->DisGradAvg[3 * i + j] + c->DisGradAvg[3 * j + i]) / 2.0;
    c->e_vm = c->d_vm * d.TimeStep;
    c->err = c->opt.err_override > 0 ? c->opt.err_override : d.Err;
    c->itermax = c->opt.itermax_override > 0 ? c->opt.itermax_override : d.IterMax;
    c->nsteps = c->opt.nsteps_override > 0 ? c->opt.nsteps_override : d.N_steps;

    // device allocation
    HIPCK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    hipStream_t st = c->stream;
    HIPCK(hipMalloc(&c->dc, sizeof(DevConst)));
    if (int e = ctx_copy(st, c->dc, &c->hc, sizeof(DevConst), hipMemcpyHostToDevice)) return e;
    auto dalloc = [&](double **p, int comps, double init) -> int {
        HIPCK(hipMalloc(p, sizeof(double) * comps * N));
        if (init == 0.0) return ctx_zero(st, *p, sizeof(double) * comps * N);
        std::vector<double> h((size_t)comps * N, init);
        return ctx_copy(st, *p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice);
    };
    Fields &F = c->f;
    int rc = 0;
    {
        // the Newton constants: 36 components at an even pitch (16-byte aligned component slices)
        const size_t bytes = sizeof(double) * 36 * nrc_pitch(N);
        HIPCK(hipMalloc(&F.nrc, bytes));
        if ((rc = ctx_zero(st, F.nrc, bytes))) return rc;
        // mixed precision keeps them in FP32, in the same memory (36 nrcf_pitch(N) floats fit)
        F.nrcf = c->opt.precision == SRX_PREC_MIXED ? reinterpret_cast<float *>(F.nrc) : nullptr;
    }
    if ((rc = dalloc(&F.sig, 6, 0.0)) || (rc = dalloc(&F.sig_pre, 5, 0.0)) || (rc = dalloc(&F.eps, 5, 0.0)) ||
        (rc = dalloc(&F.edot, 5, 0.0)) || (rc = dalloc(&F.dg, 9, 0.0)) || (rc = dalloc(&F.dg0, 9, 0.0)) ||
        (rc = dalloc(&F.tran, 9, 0.0)) || (rc = dalloc(&F.tran0, 9, 0.0)) ||
        (rc = dalloc(&F.rho_s, 12, d.rho_SSD_initial)) || (rc = dalloc(&F.rho_m, 12, 0.0)) ||
        (rc = dalloc(&F.rho_m_old, 12, 0.0)) || (rc = dalloc(&F.rho_F, 12, 0.0)) || (rc = dalloc(&F.rho_P, 12, 0.0)) ||
        (rc = dalloc(&F.gamdot, 12, 0.0)) || (rc = dalloc(&F.gam_acum, 12, 0.0)) ||
        (rc = dalloc(&F.rho_dipole, 12, 0.0)) || (rc = dalloc(&F.rho_forest, 12, 0.0)) ||
        (rc = dalloc(&F.rho_inplane, 12, 0.0)) || (rc = dalloc(&F.rho_tot, 1, 0.0)) ||
        (rc = dalloc(&F.acr, 15, 0.0)) ||
        (c->creep == 1 && (rc = dalloc(&F.fl0, 9, 0.0))))
        return rc;
    F.acrf = c->opt.precision == SRX_PREC_MIXED ? reinterpret_cast<float *>(F.acr) : nullptr;
    if (c->opt.gnd) {
        // GND densities start at zero (init.c:1077-1087); same-grain flags of Gradient_ExchangeGrainID
        // (constitutive.c:731-880): +x, -x, +y, -y, +z and, as the reference compares, the y- neighbour
        // for the z- side; periodic
        if ((rc = dalloc(&c->hc.rho_g1, 12, 0.0)) || (rc = dalloc(&c->hc.rho_g2, 12, 0.0)) ||
            (rc = dalloc(&c->hc.gnd_fp, 24, 0.0)))
            return rc;
        const int n0 = c->n[0], n1 = c->n[1], n2 = c->n[2];
        std::vector<unsigned char> gb(N);
        const std::vector<int> &gr = c->grain;
        for (long v = 0; v < N; ++v) {
            const int x = (int)(v / ((long)n1 * n2)), y = (int)((v / n2) % n1), z = (int)(v % n2);
            const int xp = x + 1 >= n0 ? 0 : x + 1, xm = x - 1 < 0 ? n0 - 1 : x - 1;
            const int yp = y + 1 >= n1 ? 0 : y + 1, ym = y - 1 < 0 ? n1 - 1 : y - 1, zp = z + 1 >= n2 ? 0 : z + 1;
            auto at = [&](int i, int j, int k) { return gr[((long)i * n1 + j) * n2 + k]; };
            const int g = gr[v];
            gb[v] = (unsigned char)((g == at(xp, y, z)) | (g == at(xm, y, z)) << 1 | (g == at(x, yp, z)) << 2 |
                                    (g == at(x, ym, z)) << 3 | (g == at(x, y, zp)) << 4 | (g == at(x, ym, z)) << 5);
        }
        HIPCK(hipMalloc(&c->hc.gbc, N));
        if ((rc = ctx_copy(st, c->hc.gbc, gb.data(), N, hipMemcpyHostToDevice))) return rc;
    }
    if (c->opt.warm_start == 1 || c->opt.warm_start == 2) {
        HIPCK(hipMalloc(&F.dsig, sizeof(float) * 6 * N));
        if ((rc = ctx_zero(st, F.dsig, sizeof(float) * 6 * N))) return rc;
    }
    if (c->opt.warm_start == 2) {   // the learned start's forcing feature (srx_mlp_record adds it too)
        HIPCK(hipMalloc(&F.ecp, sizeof(float) * 6 * N));
        if ((rc = ctx_zero(st, F.ecp, sizeof(float) * 6 * N))) return rc;
    }
    HIPCK(hipMalloc(&F.phase, sizeof(int) * N));
    HIPCK(hipMalloc(&F.nr_it, sizeof(int) * N));
    if ((rc = ctx_zero(st, F.nr_it, sizeof(int) * N)) ||
        (rc = ctx_copy(st, F.phase, phase.data(), sizeof(int) * N, hipMemcpyHostToDevice)))
        return rc;
    c->nr_blocks = nr_block_count((int)N);
    c->end_blocks = end_block_count((int)N);
    const int pb = std::max(std::max(c->nr_blocks * NR_NPART, c->end_blocks * END_NPART), setup_block_count((int)N) * 42);
    HIPCK(hipMalloc(&c->partials, sizeof(double) * pb));
    HIPCK(hipMalloc(&c->red, sizeof(double) * 64));
    HIPCK(hipMalloc(&c->redseg, sizeof(double) * NR_RED_SEG_DOUBLES));
    HIPCK(hipHostMalloc((void **)&c->hred, sizeof(double) * 64, hipHostMallocDefault));
    // microstructure and InitFields on the device: orientations, initial stress, and the grid sums
    // of C_gr and Sig
    double sums[42];
    {
        double *deul = nullptr;
        HIPCK(hipMalloc(&deul, sizeof(double) * 3 * N));
        rc = ctx_copy(st, deul, eul.data(), sizeof(double) * 3 * N, hipMemcpyHostToDevice);
        if (!rc) {
            launch_setup(st, c->dc, F, (int)N, deul, PI_REF / 180., e9, c->Scauchy, c->creep, c->partials, c->red);
            rc = ctx_copy(st, sums, c->red, sizeof(sums), hipMemcpyDeviceToHost);
        }
        (void)hipFree(deul);
        if (rc) return rc;
    }
    for (int i = 0; i < 6; ++i) c->sig_sum6[i] = sums[36 + i];
    // C066 = <C_gr> over the grid (the device sums, in a fixed order), its inverse and tables
    {
        for (int k = 0; k < 36; ++k) c->C066[k] = sums[k] * c->wgt;
        // exact block structure: the hydrostatic axis of a cubic crystal is isotropic
        for (int k = 0; k < 5; ++k) c->C066[6 * k + 5] = c->C066[30 + k] = 0.0;
        std::memcpy(c->S066, c->C066, sizeof(c->C066));
        if (!invert(c->S066, 6)) return 35;
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) c->hc.S066[i][j] = c->S066[6 * i + j];
        c66_to_c4(c->C066, c->hc.C0);
        c66_to_c4(c->S066, c->hc.S0);
        // G6[(ik)][(jl)] = C0_ijkl (+ C0_ilkj for j != l): A_ik = sum_jl C0_ijkl n_j n_l as a 6x6 product
        static const int PI6[6][2] = {{0, 0}, {1, 1}, {2, 2}, {1, 2}, {0, 2}, {0, 1}};
        for (int a = 0; a < 6; ++a)
            for (int b = 0; b < 6; ++b) {
                const int i = PI6[a][0], k = PI6[a][1], j = PI6[b][0], l = PI6[b][1];
                double v = c->hc.C0[((i * 3 + j) * 3 + k) * 3 + l];
                if (j != l) v += c->hc.C0[((i * 3 + l) * 3 + k) * 3 + j];
                c->hc.G6[6 * a + b] = v;
                c->hc.G6f[6 * a + b] = (float)v;
            }
    }

    if (int e = ctx_copy(st, c->dc, &c->hc, sizeof(DevConst), hipMemcpyHostToDevice)) return e;

    const size_t H = (size_t)(c->n[2] / 2 + 1);
    const size_t spec_elems = 9 * H * c->n[0] * c->n[1];
    const size_t csz = c->opt.precision == SRX_PREC_MIXED ? 2 * sizeof(float) : 2 * sizeof(double);
    HIPCK(hipMalloc(&c->specB, spec_elems * csz));
    HIPCK(hipMalloc(&c->specC, spec_elems * csz));
    for (int ax = 0; ax < 3; ++ax) {
        int r = c->opt.precision == SRX_PREC_MIXED ? upload_twiddles<float>(c->n[ax], &c->tw[ax], st)
                                                   : upload_twiddles<double>(c->n[ax], &c->tw[ax], st);
        if (r) return r;
    }
    HIPCK(hipMalloc(&c->dctl, sizeof(IterCtl)));
    HIPCK(hipHostMalloc((void **)&c->hctl, sizeof(IterCtl), hipHostMallocDefault));
    return 0;
}

void open_files(srx_ctx *c) {
    if (!c->opt.write_files) return;
    auto op = [&](const char *name) { return std::fopen(resolve(c->workdir, name).c_str(), "w+"); };
    c->fvm = op("vm.out");
    c->fdd = op("dd.out");
    c->fdrx = op("drx.out");
    c->ferr = op("err.out");
    // headers of InitFields (init.c:497-505)
    if (c->fvm) std::fprintf(c->fvm, "TIME\tEVM\tEVMP\tDVM\tDVMP\tSVM\tSVMSVM_1\n");
    if (c->fdd) std::fprintf(c->fdd, "TIME\tEVM\tDD_F\tDD_P\tDD_s\tDD_m\tDD_g\n");
    if (c->fdrx) std::fprintf(c->fdrx, "TIME\tEVM\tNEW_IDs\n");
    if (c->ferr) std::fprintf(c->ferr, "IT\tERRE\tERRS\tSVM\n");
}

// get_smacro (evolution.c:590-664) from the reduced sums
// event-timed launch helper: events come from a pool that is reused every increment
struct Timer {
    srx_ctx *c;
    int slot, iter;
    size_t ia = 0;
    bool on;
    // iter: the outer iteration a launch belongs to (0: not an iteration kernel); launches of
    // iterations past the last one run (they return at once) are left out of the totals
    Timer(srx_ctx *cc, int s, int it = 0)
        : c(cc), slot(s), iter(it), on(cc->timing == 1 || (cc->timing == 2 && s == 5)) {
        if (!on) return;
        ia = c->evused;
        if (c->evused + 2 > c->evpool.size()) {
            for (int k = 0; k < 256; ++k) {
                hipEvent_t e;
                (void)hipEventCreate(&e);
                c->evpool.push_back(e);
            }
        }
        c->evused += 2;
        set_launch_events(c->evpool[ia], c->evpool[ia + 1]);
    }
    ~Timer() {
        if (!on) return;
        set_launch_events(nullptr, nullptr);
        c->evslot.push_back(slot);
        c->evit.push_back(iter);
    }
};

const char *KT_NAMES[8] = {"fwd_z", "fwd_y", "green_x", "inv_y", "inv_z", "nr", "end", "reduce"};

void collect_timing(srx_ctx *c) {
    if (!c->timing) return;
    (void)hipStreamSynchronize(c->stream);
    for (size_t i = 0; i < c->evslot.size(); ++i) {
        if (c->evit[i] > c->iters_last) continue;
        float ms = 0;
        (void)hipEventElapsedTime(&ms, c->evpool[2 * i], c->evpool[2 * i + 1]);
        c->kt[c->evslot[i]].ms += ms;
        c->kt[c->evslot[i]].calls += 1;
    }
    c->evslot.clear();
    c->evit.clear();
    c->evused = 0;
}

int sync_read(srx_ctx *c, int nvals) {
    HIPCK(hipMemcpyAsync(c->hred, c->red, sizeof(double) * nvals, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

void srx_default_options(srx_options *o) {
    std::memset(o, 0, sizeof(*o));
}

const char *srx_strerror(int code) {
    switch (code) {
    case 0: return "success";
    case 4: return "chg_basis: wrong option";
    case 7: return "slip system is not orthogonal";
    case 28: return "only rank 0 may write the global files";
    case 35: return "singular 6x6 matrix";
    case 114: return "check boundary conditions on strain rate and stress";
    case 116: return "cannot enforce only two deviatoric components (VelGrad_BC_Flag)";
    case 230: return "initial hardening rate lower than the final one";
    case 445: return "wrong option for Ten4thTransform";
    case 604: return "wrong option for EulerToTransMatrix";
    case 720: return "wrong type of dislocations to write";
    case 1110: return "interaction strength matrix supports FCC only";
    case 1139: return "error in reading the microstructure file (inconsistent index system)";
    case SRX_E_FILE: return "file not found";
    case SRX_E_GRID: return "unsupported grid size";
    case SRX_E_HIP: return "HIP runtime error";
    case SRX_E_SYSTEMS: return "more than 12 slip systems";
    case SRX_E_NOMEM: return "out of memory";
    case SRX_E_SCRATCH: return "problem larger than the static scratch of the allocation-free entry point";
    case SRX_E_DONE: return "all increments done";
    case SRX_E_CREEP_STOP: return "creep run stopped at e_vm > 0.3";
    case SRX_E_SIZE: return "buffer too small";
    case SRX_E_FIELD: return "unknown field id";
    case SRX_E_ARG: return "invalid argument";
    case SRX_E_NAN: return "NaN in the constitutive update";
    case SRX_E_UPDATE_FLAG: return "Update_Flag must be 1";
    default: return "unknown error";
    }
}

// Average time (us) of three launches of FFT pass `pass` (0 forward z, 1 forward y, 2 Green x,
// 3 inverse y, 4 inverse z) at tile `t`, after one warm-up launch, on the context's own buffers with
// its current plans (the inverse z pass writes the scratch DisGrad `scratch`).  With a `flush`
// buffer (2 x FLUSH_BYTES) each launch is timed alone after a streaming copy through it, so the
// pass reads its input from HBM as it does in the solver, where the Newton kernel's traffic has
// evicted it from the last-level cache; else the three launches run back to back.
constexpr size_t FLUSH_BYTES = (size_t)320 << 20;
static int time_pass(srx_ctx *c, int pass, int t, double *scratch, hipEvent_t e0, hipEvent_t e1, double &us,
                     char *flush = nullptr) {
    const int prec = c->opt.precision;
    hipStream_t st = c->stream;
    auto run = [&]() {
        switch (pass) {
        case 0: launch_fft_fwd_z(st, prec, c->f.sig, c->specB, c->n, c->pl[2], c->tw[2], nullptr, t); break;
        case 1: launch_fft_y(st, prec, c->specB, c->specC, 6, -1, 1, c->n, c->pl[1], c->tw[1], nullptr, t); break;
        case 2: launch_green_x(st, prec, c->specC, c->n, c->pl[0], c->tw[0], c->dc, nullptr, t); break;
        case 3: launch_fft_y(st, prec, c->specC, c->specB, 9, +1, 0, c->n, c->pl[1], c->tw[1], nullptr, t); break;
        default:
            launch_fft_inv_z_update(st, prec, c->specB, scratch, c->n, c->pl[2], c->tw[2], nullptr, (double)c->nvox,
                                    t);
        }
    };
    run();
    float ms = 0.0f;
    if (flush) {
        double tot = 0.0;
        for (int r = 0; r < 3; ++r) {
            launch_stream_copy(st, flush, flush + FLUSH_BYTES, FLUSH_BYTES / 16, 1, 2048);
            if (hipEventRecord(e0, st) != hipSuccess) return SRX_E_HIP;
            run();
            if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess || hipGetLastError() != hipSuccess)
                return SRX_E_HIP;
            tot += ms;
        }
        us = 1000.0 * tot / 3.0;
        return 0;
    }
    if (hipEventRecord(e0, st) != hipSuccess) return SRX_E_HIP;
    for (int r = 0; r < 3; ++r) run();
    if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess || hipGetLastError() != hipSuccess)
        return SRX_E_HIP;
    us = 1000.0 * ms / 3.0;
    return 0;
}

static_assert(SRX_FFT_MAX_STAGES == MAX_RADIX_STAGES, "plan stage count of the ABI");

// FFT passes of each plan axis (x: the Green pass; y and z: forward and inverse)
static const int AXIS_PASSES[3][2] = {{2, -1}, {1, 3}, {0, 4}};

// Radix plans chosen by timing (options.reserved[3] = 1), kept per process for the grid shape and
// precision: every context of one shape, e.g. the members of an ensemble, runs the same plans.
static std::mutex g_plan_mu;
static std::map<std::array<int, 5>, FftPlan> g_plan_cache;

// FFT autotuner, at srx_open on the context's own buffers.
// Radix plans (opt-in, options.reserved[3] = 1): for each axis whose passes take a plan (not the
// x axis where the wavefront Green pass runs), every candidate of plan_candidates that the passes
// accept is timed over that axis's passes at their default tiles, each launch alone after a
// streaming copy that evicts its input from the caches, in two rounds (the faster one counts);
// another plan replaces make_plan's only when more than 5% faster, since the timings at
// srx_open vary by a few percent from one context to the next.  Another radix order is another
// rounding order, so results then differ from the default plans in the last bits (within the FFT's
// rounding error), which is why the default keeps make_plan's plans.
// Tiles (unless options.reserved[2] = 1): each pass is timed with every candidate tile (powers of
// two up to the largest valid one, and that one), the default (the largest) first; another must be
// more than 3% faster to replace it, so near-ties do not flip between runs.  Tiles change only how many pencils a block
// transforms, which leaves every result bit-identical; the wavefront x pass (a different operation
// order) is not timed against the LDS kernel but always used where it applies, so results never
// depend on the tile timings (or on what else the GPU runs meanwhile).
static int autotune_fft(srx_ctx *c) {
    const int prec = c->opt.precision;
    double *scratch = nullptr;
    HIPCK(hipMalloc(&scratch, sizeof(double) * 9 * (size_t)c->nvox));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = SRX_E_HIP;
    char *flush = nullptr;
    if (!rc && c->opt.reserved[3] == 1 && hipMalloc(&flush, 2 * FLUSH_BYTES) != hipSuccess) rc = SRX_E_NOMEM;
    for (int ax = 0; ax < 3 && !rc && c->opt.reserved[3] == 1; ++ax) {
        if (ax == 0 && green_wave_ok(c->n)) continue;   // the wavefront pass takes no plan
        const std::array<int, 5> key = {ax, c->n[0], c->n[1], c->n[2], prec};
        {
            std::lock_guard<std::mutex> lk(g_plan_mu);
            auto it = g_plan_cache.find(key);
            if (it != g_plan_cache.end()) {
                c->pl[ax] = it->second;
                continue;
            }
        }
        FftPlan cand[32];
        const int nc = plan_candidates(c->n[ax], cand, 32);
        const FftPlan dflt = c->pl[ax];
        double t_best[32];
        for (int i = 0; i < nc; ++i) t_best[i] = 1e300;
        // two rounds over the candidates, each keeping its faster round, so that a slow spell of
        // the GPU does not favour the candidates timed after it
        for (int round = 0; round < 2 && !rc; ++round)
            for (int i = 0; i < nc && !rc; ++i) {
                FftPlan trial[3] = {c->pl[0], c->pl[1], c->pl[2]};
                trial[ax] = cand[i];
                if (!grid_supported(trial, c->n, prec)) continue;
                c->pl[ax] = cand[i];
                double tot = 0.0;
                for (int k = 0; k < 2 && !rc; ++k) {
                    double us = 0.0;
                    if (AXIS_PASSES[ax][k] >= 0) rc = time_pass(c, AXIS_PASSES[ax][k], 0, scratch, e0, e1, us, flush);
                    tot += us;
                }
                t_best[i] = std::min(t_best[i], tot);
            }
        c->pl[ax] = dflt;
        if (rc) break;
        // make_plan's plan (candidate 0) unless another is more than 5% faster: then the fastest
        int bi = 0;
        for (int i = 1; i < nc; ++i)
            if (t_best[i] < 0.95 * t_best[0] && (bi == 0 || t_best[i] < t_best[bi])) bi = i;
        const FftPlan best_plan = cand[bi];
        std::lock_guard<std::mutex> lk(g_plan_mu);
        c->pl[ax] = g_plan_cache.emplace(key, best_plan).first->second;   // a context that tuned first wins
    }
    for (int pass = 0; pass < 5 && !rc && c->opt.reserved[2] == 0; ++pass) {
        const FftPlan &p = c->pl[pass == 2 ? 0 : (pass == 1 || pass == 3) ? 1 : 2];
        const int tmax = fft_tile_max(pass, c->n, p, prec);
        std::vector<int> cand;
        if (pass == 2 && green_wave_ok(c->n)) cand.push_back(-1);   // one wavefront per row
        else {
            cand.push_back(tmax);
            for (int t = 1; t < tmax; t *= 2) cand.push_back(t);
        }
        double best = 1e300;
        for (int t : cand) {
            double us = 0.0;
            if ((rc = time_pass(c, pass, t, scratch, e0, e1, us))) break;
            if (us < (best < 1e300 ? 0.97 * best : best)) {
                best = us;
                c->tiles[pass] = t;
                c->tile_us[pass] = us;
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(scratch);
    if (flush) (void)hipFree(flush);
    if (c->opt.verbose && !rc) {
        std::printf("fft tiles (fwd z, fwd y, green x, inv y, inv z): %d %d %d %d %d  (%.1f %.1f %.1f %.1f %.1f us)\n",
                    c->tiles[0], c->tiles[1], c->tiles[2], c->tiles[3], c->tiles[4], c->tile_us[0], c->tile_us[1],
                    c->tile_us[2], c->tile_us[3], c->tile_us[4]);
        for (int ax = 0; ax < 3; ++ax) {
            std::printf("fft plan axis %d (n = %d):", ax, c->pl[ax].n);
            for (int s = 0; s < c->pl[ax].nst; ++s) std::printf(" %d", c->pl[ax].rad[s]);
            std::printf("\n");
        }
    }
    return rc;
}

int srx_fft_tiles(srx_ctx *c, int *tiles, double *us) {
    if (!c || !tiles) return SRX_E_ARG;
    for (int k = 0; k < 5; ++k) {
        tiles[k] = c->tiles[k];
        if (us) us[k] = c->tile_us[k];
    }
    return 0;
}

int srx_fft_plan(srx_ctx *c, int axis, int *radices, int *nstages) {
    if (!c || !radices || !nstages || axis < 0 || axis > 2) return SRX_E_ARG;
    *nstages = c->pl[axis].nst;
    for (int s = 0; s < c->pl[axis].nst; ++s) radices[s] = c->pl[axis].rad[s];
    return 0;
}

static void drop_graphs(srx_ctx *c);

int srx_set_fft_plan(srx_ctx *c, int axis, const int *radices, int nstages) {
    if (!c || !radices || axis < 0 || axis > 2 || nstages < 1 || nstages > SRX_FFT_MAX_STAGES) return SRX_E_ARG;
    FftPlan p;
    p.n = c->n[axis];
    p.nst = nstages;
    long prod = 1;
    for (int s = 0; s < nstages; ++s) {
        const int r = radices[s];
        if (r != 2 && r != 3 && r != 4 && r != 5 && r != 8) return SRX_E_ARG;
        p.rad[s] = r;
        prod *= r;
        if (prod > p.n) return SRX_E_ARG;
    }
    if (prod != p.n) return SRX_E_ARG;
    FftPlan trial[3] = {c->pl[0], c->pl[1], c->pl[2]};
    trial[axis] = p;
    if (!grid_supported(trial, c->n, c->opt.precision)) return SRX_E_GRID;
    HIPCK(hipStreamSynchronize(c->stream));
    c->pl[axis] = p;
    for (int k = 0; k < 2; ++k)
        if (AXIS_PASSES[axis][k] >= 0) {
            c->tiles[AXIS_PASSES[axis][k]] = 0;   // the default tile of the new plan
            c->tile_us[AXIS_PASSES[axis][k]] = 0.0;
        }
    drop_graphs(c);   // captured batches hold the previous plan
    return 0;
}

int srx_fft_plan_candidates(int n, int *radices, int max_plans, int *nplans) {
    if (!nplans || max_plans < 0 || (max_plans > 0 && !radices)) return SRX_E_ARG;
    FftPlan cand[32];
    const int nc = plan_candidates(n, cand, 32);
    if (nc == 0) return SRX_E_GRID;
    *nplans = nc;
    for (int i = 0; i < nc && i < max_plans; ++i)
        for (int s = 0; s < SRX_FFT_MAX_STAGES; ++s)
            radices[i * SRX_FFT_MAX_STAGES + s] = s < cand[i].nst ? cand[i].rad[s] : 0;
    return nc > max_plans ? SRX_E_SIZE : 0;
}

int srx_open(const char *deck, const char *workdir, const srx_options *opt, srx_ctx **out) {
    if (!deck || !out) return SRX_E_ARG;
    srx_ctx *c = new srx_ctx();
    if (opt) c->opt = *opt; else srx_default_options(&c->opt);
    if (workdir) c->workdir = workdir;
    else {
        std::string d(deck);
        const size_t p = d.find_last_of('/');
        c->workdir = p == std::string::npos ? std::string(".") : d.substr(0, p);
    }
    c->timing = c->opt.reserved[0] == 2 ? 2 : (c->opt.reserved[0] != 0 ? 1 : 0);
    c->graphs = c->opt.reserved[1] == 0;
    for (int k = 0; k < 8; ++k) c->kt[k] = {KT_NAMES[k], 0.0, 0};
    c->nkt = 8;
    int rc = read_deck(deck, c->dk);
    if (!rc) rc = setup(c);
    if (!rc && (c->opt.reserved[2] == 0 || c->opt.reserved[3] == 1)) rc = autotune_fft(c);
    if (rc) {
        srx_close(c);
        return rc;
    }
    open_files(c);
    *out = c;
    return 0;
}

int srx_dims(srx_ctx *c, int *n3) {
    if (!c || !n3) return SRX_E_ARG;
    for (int i = 0; i < 3; ++i) n3[i] = c->n[i];
    return 0;
}

// ---- field dumps at PrintControl steps (evolution.c:1456-1483) -----------
// Raw native-endian records, one per voxel in the reference's voxel order, with
// the file names and record layouts of io.c.  Record holes are written as zeros
// (MPI-IO leaves them unwritten in a new file).

// EulerToTransMatrix, opt 1 (kinematics.c:196-214): sample->crystal matrix to Bunge
// angles in degrees, with the truncated pi of evp.h:84.
static void matrix_to_euler_deg(const double a[3][3], double &t1, double &ph, double &t2) {
    const double PI_TRUNC = 3.14159265359;
    ph = std::acos(a[2][2]);
    if (std::fabs(a[2][2]) > 0.9999) {
        t2 = 0.;
        t1 = std::atan2(a[0][1], a[0][0]);
    } else {
        const double sph = std::sin(ph);
        t1 = std::atan2(a[2][0] / sph, -1.0 * a[2][1] / sph);
        t2 = std::atan2(a[0][2] / sph, a[1][2] / sph);
    }
    ph *= 180. / PI_TRUNC;
    t1 *= 180. / PI_TRUNC;
    t2 *= 180. / PI_TRUNC;
}

static int write_blob(const srx_ctx *c, const char *name, const void *data, size_t bytes) {
    FILE *fp = std::fopen(resolve(c->workdir, name).c_str(), "wb");
    if (!fp) return SRX_E_FILE;
    const size_t w = std::fwrite(data, 1, bytes, fp);
    std::fclose(fp);
    return w == bytes ? 0 : SRX_E_FILE;
}

// Six symmetric components in the order 11, 22, 33, 23, 13, 12 (io.c:322-437).
static const int IOUT_SYM[6][2] = {{0, 0}, {1, 1}, {2, 2}, {1, 2}, {0, 2}, {0, 1}};

static int write_field_dumps(srx_ctx *c, int step) {
    const long N = c->nvox;
    int rc = 0;
    std::vector<double> sig(9 * N), eps(9 * N), dg(9 * N), tr(9 * N), rs(12 * N), ga(12 * N), fl(9 * N), ph(N);
    if ((rc = srx_get_field(c, SRX_F_SIG, sig.data(), sig.size())) ||
        (rc = srx_get_field(c, SRX_F_EPS, eps.data(), eps.size())) ||
        (rc = srx_get_field(c, SRX_F_DISGRAD, dg.data(), dg.size())) ||
        (rc = srx_get_field(c, SRX_F_TRANMAT, tr.data(), tr.size())) ||
        (rc = srx_get_field(c, SRX_F_RHO_S, rs.data(), rs.size())) ||
        (rc = srx_get_field(c, SRX_F_TRIAL_GAM_ACUM, ga.data(), ga.size())) ||
        (rc = srx_get_field(c, SRX_F_FLUCT, fl.data(), fl.size())) ||
        (rc = srx_get_field(c, SRX_F_PHASE, ph.data(), ph.size())))
        return rc;
    char name[256];
    std::vector<double> buf(N);
    // WriteEpsMPI, WriteElsMPI, WriteSigMPI
    for (int which = 0; which < 3; ++which) {
        const char *stem = which == 0 ? "eps" : which == 1 ? "els" : "sig_afterDRX";
        for (int i = 0; i < 6; ++i) {
            const int a = IOUT_SYM[i][0], b = IOUT_SYM[i][1];
            for (long v = 0; v < N; ++v) {
                if (which == 0) buf[v] = eps[9 * v + 3 * a + b];
                else if (which == 2) buf[v] = sig[9 * v + 3 * a + b];
                else if (i < 3) buf[v] = dg[9 * v + 3 * a + b] - eps[9 * v + 3 * a + b];
                else buf[v] = (dg[9 * v + 3 * a + b] + dg[9 * v + 3 * b + a]) / 2.0 - eps[9 * v + 3 * a + b];
            }
            std::snprintf(name, sizeof(name), "%s_%01d_S%06d.iout", stem, i + 1, step);
            if ((rc = write_blob(c, name, buf.data(), sizeof(double) * N))) return rc;
        }
    }
    // WriteRhoMPI(..., "SSD", ...): sum of rho_s over the phase's systems (io.c:891-941)
    for (long v = 0; v < N; ++v) {
        const int jph = (int)ph[v];
        const int ns = (jph >= 1 && jph <= NPH && c->ph[jph - 1].xtal) ? c->hc.ph[jph - 1].nsys : 0;
        double acc = 0.0;
        for (int i = 0; i < ns; ++i) acc += rs[12 * v + i];
        buf[v] = acc;
    }
    std::snprintf(name, sizeof(name), "rho_afterDRX_SSD_S%06d.iout", step);
    if ((rc = write_blob(c, name, buf.data(), sizeof(double) * N))) return rc;
    // WriteTextureMPI: 48-byte records {double angle[3]; int coord[3], jgr, jph, buff} (io.c:478-541)
    {
        struct TexRec { double angle[3]; int coord[3]; int jgr, jph, buff; };
        static_assert(sizeof(TexRec) == 48, "tex record");
        std::vector<TexRec> rec(N);
        for (long v = 0; v < N; ++v) {
            double sa2xt[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) sa2xt[i][j] = tr[9 * v + 3 * j + i];
            TexRec &r = rec[v];
            matrix_to_euler_deg(sa2xt, r.angle[0], r.angle[1], r.angle[2]);
            const long x = v / ((long)c->n[1] * c->n[2]), y = (v / c->n[2]) % c->n[1], z = v % c->n[2];
            r.coord[0] = (int)x + 1;
            r.coord[1] = (int)y + 1;
            r.coord[2] = (int)z + 1;
            r.jgr = c->grain[v];
            r.jph = (int)ph[v];
            r.buff = 0;
        }
        std::snprintf(name, sizeof(name), "tex_afterDRX_S%06d.iout", step);
        if ((rc = write_blob(c, name, rec.data(), sizeof(TexRec) * N))) return rc;
    }
    // WriteSLIPMPI / WriteDDMPI: 104-byte records {int jgr; 4-byte hole; double v[12]} (io.c:630-719, 798-887)
    {
        struct SysRec { int jgr; int hole; double v[12]; };
        static_assert(sizeof(SysRec) == 104, "slip record");
        std::vector<SysRec> rec(N);
        for (int which = 0; which < 2; ++which) {
            const std::vector<double> &src = which == 0 ? ga : rs;
            for (long v = 0; v < N; ++v) {
                rec[v].jgr = c->grain[v];
                rec[v].hole = 0;
                for (int k = 0; k < 12; ++k) rec[v].v[k] = src[12 * v + k];
            }
            std::snprintf(name, sizeof(name), "%s_S%04d.iout", which == 0 ? "slip_afterDRX" : "dislocation_afterDRX",
                          step);
            if ((rc = write_blob(c, name, rec.data(), sizeof(SysRec) * N))) return rc;
        }
    }
    // WriteNewPositionMPI: 24-byte records (io.c:723-796).  new_position is the affine image
    // of the voxel under DisGrad (evolution.c:1437-1441); displacement_fluct[i] is what the
    // transform block of evolution.c:1401-1428 leaves in the voxel, -fluctuation[i][2] (its
    // out-of-bounds store lies outside the array and is not reproduced).
    {
        std::vector<double> rec(3 * N);
        for (long v = 0; v < N; ++v) {
            const long x = v / ((long)c->n[1] * c->n[2]), y = (v / c->n[2]) % c->n[1], z = v % c->n[2];
            const double *D = &dg[9 * v];
            const double X = (double)x, Y = (double)y, Z = (double)z;
            const double np0 = (1.0 + D[0]) * X + D[1] * Y + D[2] * Z;
            const double np1 = (1.0 + D[4]) * Y + D[3] * X + D[5] * Z;
            const double np2 = (1.0 + D[8]) * Z + D[6] * X + D[7] * Y;
            rec[3 * v + 0] = np0 + (-fl[9 * v + 2]) - X;
            rec[3 * v + 1] = np1 + (-fl[9 * v + 5]) - Y;
            rec[3 * v + 2] = np2 + (-fl[9 * v + 8]) - Z;
        }
        std::snprintf(name, sizeof(name), "new_position_afterDRX_S%04d.iout", step);
        if ((rc = write_blob(c, name, rec.data(), sizeof(double) * 3 * N))) return rc;
    }
    return 0;
}

// One increment (one pass of the istep loop of evolution.c:726-1520) in three phases, so that an
// ensemble can run the outer iterations of all its members together (run_lockstep):
// step_begin: the increment prologue and the control block of its outer iterations;
// step_iterate: the outer iterations of this context alone; step_finish: everything after them.
static int step_begin(srx_ctx *c, EnsIncr *batch = nullptr) {
    const Deck &d = c->dk;
    const double dt = d.TimeStep;
    const int prec = c->opt.precision;
    const int N = (int)c->nvox;
    (void)dt; (void)prec; (void)N; (void)d;
    if (c->ferr && c->nsteps != 1) std::fprintf(c->ferr, "STEP = %d\n", c->istep);

    // evolution.c:1115-1137
    double add[9], set[9];
    for (int k = 0; k < 9; ++k) {
        add[k] = c->Udot[k] * dt;
        set[k] = c->DisGradAvg[k] + c->acum[k];
    }
    if (batch) {   // a lockstep ensemble launches every member's prologue at once
        for (int k = 0; k < 9; ++k) {
            batch->add[k] = add[k];
            batch->set[k] = set[k];
            batch->uold[k] = c->uniform[k];
        }
        batch->creep = c->creep;
    } else {
        launch_prologue(c->stream, c->f, N, add, c->creep, set, c->uniform);
    }
    if (c->creep == 1) for (int k = 0; k < 9; ++k) c->uniform[k] = set[k];
    else for (int k = 0; k < 9; ++k) c->uniform[k] += add[k];
    for (int k = 0; k < 9; ++k) c->dDisGradAvg[k] = c->acum[k] = 0.0;
    // outer fixed point (evolution.c:1150-1256), controlled on the device: the host enqueues
    // iterations in batches and reads the control block once per batch
    IterCtl &ctl = *c->hctl;
    std::memset(&ctl, 0, sizeof(IterCtl));
    ctl.wgt = c->wgt;
    ctl.wgt_ph1 = c->wgt_ph1;
    ctl.err = c->err;
    ctl.e_vm_prev = c->e_vm;
    ctl.svm_stray = c->opt.svm_stray;
    ctl.svm_mode = c->opt.svm_mode;
    ctl.nvox = (double)c->nvox;
    for (int k = 0; k < 9; ++k) ctl.Scauchy[k] = c->Scauchy[k];
    for (int k = 0; k < 6; ++k) {
        ctl.vbc[k] = d.VelGrad_BC_Flag[k];
        ctl.sbc[k] = d.Stress_BC_Flag[k];
    }
    ctl.itermax = c->itermax;
    ctl.fixed_iters = c->opt.fixed_iters;
    ctl.mixed = prec == SRX_PREC_MIXED;
    ctl.Err_s = ctl.Err_e = 2. * c->err;
    for (int k = 0; k < 6; ++k) {
        ctl.sig_sum6[k] = c->sig_sum6[k];
        ctl.mean6[k] = ctl.mixed ? c->sig_sum6[k] / ctl.nvox : 0.0;
        ctl.dcs6[k] = ctl.mixed ? ctl.mean6[k] * ctl.nvox : 0.0;
    }
    for (int k = 0; k < 9; ++k) ctl.uniform[k] = c->uniform[k];
    ctl.done = ctl_stop(ctl);
    HIPCK(hipMemcpyAsync(c->dctl, &ctl, sizeof(IterCtl), hipMemcpyHostToDevice, c->stream));
    return 0;
}

static int step_iterate(srx_ctx *c) {
    const Deck &d = c->dk;
    const double dt = d.TimeStep;
    const int prec = c->opt.precision;
    const int N = (int)c->nvox;
    (void)dt; (void)prec; (void)N; (void)d;
    const int rotated = c->istep > 2 ? 1 : 0;
    IterCtl &ctl = *c->hctl;
    const double divisor = (double)c->nvox;
    const int batch = c->opt.verbose ? 1 : ITER_BATCH;
    const int most = std::max(c->itermax, c->opt.fixed_iters);
    // a batch is replayed from a HIP graph unless kernels are timed or samples recorded (those
    // batches differ from launch to launch)
    const bool use_graph = c->graphs && batch == ITER_BATCH && !c->timing && c->rec_stride == 0;
    int launched = 0;
    while (!ctl.done) {
        if (launched > most) return SRX_E_HIP;   // the control kernel must have stopped the loop
        hipGraphExec_t *gx = use_graph ? &c->gexec[launched == 0 ? 1 : 0][rotated] : nullptr;
        if (gx && *gx) {
            HIPCK(hipGraphLaunch(*gx, c->stream));
            launched += batch;
        } else {
        if (gx) HIPCK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        for (int b = 0; b < batch; ++b) {
            ++launched;
            IterCtl *g = c->dctl;
            const int *tl = c->tiles;
            { Timer t(c, 0, launched); launch_fft_fwd_z(c->stream, prec, c->f.sig, c->specB, c->n, c->pl[2], c->tw[2], g, tl[0]); }
            { Timer t(c, 1, launched); launch_fft_y(c->stream, prec, c->specB, c->specC, 6, -1, 1, c->n, c->pl[1], c->tw[1], g, tl[1]); }
            { Timer t(c, 2, launched); launch_green_x(c->stream, prec, c->specC, c->n, c->pl[0], c->tw[0], c->dc, g, tl[2]); }
            { Timer t(c, 3, launched); launch_fft_y(c->stream, prec, c->specC, c->specB, 9, +1, 0, c->n, c->pl[1], c->tw[1], g, tl[3]); }
            { Timer t(c, 4, launched); launch_fft_inv_z_update(c->stream, prec, c->specB, c->f.dg, c->n, c->pl[2], c->tw[2], g, divisor, tl[4]); }
            const bool rec = c->rec_stride > 0 && launched > 1;
            if (rec)
                launch_mlp_record(c->stream, c->dc, c->f, N, c->rec_stride, 0, c->rec_x, c->rec_t, c->rec_nd,
                                  c->rec_count, c->rec_cap, g);
            { Timer t(c, 5, launched); launch_nr(c->stream, c->dc, c->f, N, launched == 1, rotated, c->partials, g, c->opt.gnd, c->hc.ph[1].xtal); }
            if (rec)
                launch_mlp_record(c->stream, c->dc, c->f, N, c->rec_stride, 1, c->rec_x, c->rec_t, c->rec_nd,
                                  c->rec_count, c->rec_cap, g);
            { Timer t(c, 7, launched); launch_reduce_nr(c->stream, c->partials, c->nr_blocks, c->redseg, g); }
            launch_control(c->stream, c->dc, c->redseg, g);
        }
        if (gx) {
            hipGraph_t graph = nullptr;
            HIPCK(hipStreamEndCapture(c->stream, &graph));
            const hipError_t ie = hipGraphInstantiate(gx, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) {
                *gx = nullptr;
                std::fprintf(stderr, "srx: hipGraphInstantiate failed: %s\n", hipGetErrorString(ie));
                return SRX_E_HIP;
            }
            HIPCK(hipGraphLaunch(*gx, c->stream));
        }
        }
        HIPCK(hipMemcpyAsync(&ctl, c->dctl, sizeof(IterCtl), hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
        if (c->opt.verbose) std::printf("iter %d Err_e %e Err_s %e\n", ctl.iter, ctl.Err_e, ctl.Err_s);
    }
    return 0;
}

// An increment's end in two halves around the end-of-increment kernel, so that a lockstep ensemble
// can launch that kernel and its reduction once for all members: finish_begin takes the outer
// iterations' results from the control block, finish_end everything after the kernel's reduced
// values (c->hred).
struct Finish {
    srx_step_report r;
    int iter, nrmax, do_orient;
    long nrtot;
    double nrwave;
};

static int finish_begin(srx_ctx *c, Finish &F) {
    const Deck &d = c->dk;
    const double dt = d.TimeStep;
    IterCtl &ctl = *c->hctl;
    srx_step_report &r = F.r;
    std::memset(&r, 0, sizeof(r));
    if (ctl.nan) return SRX_E_NAN;
    F.iter = ctl.iter;
    F.nrmax = ctl.nrmax;
    F.nrtot = (long)ctl.nrtot;
    F.nrwave = ctl.nrwave;
    c->iters_last = F.iter;
    c->Err_s = ctl.Err_s;
    c->Err_e = ctl.Err_e;
    c->s_vm = ctl.s_vm;
    c->s_vm1 = ctl.s_vm1;
    for (int k = 0; k < 6; ++k) c->sig_sum6[k] = ctl.sig_sum6[k];
    for (int k = 0; k < 9; ++k) {
        c->dDisGradAvg[k] = ctl.ddg[k];
        c->acum[k] = ctl.acum[k];
        c->uniform[k] = ctl.uniform[k];
        c->SigAvg[k] = ctl.SigAvg[k];
        c->SigAvg1[k] = ctl.SigAvg1[k];
    }
    r.converged = !(std::fmax(std::fabs(c->Err_s), std::fabs(c->Err_e)) > c->err);
    // evolution.c:1268-1290
    double act[9], vga[9];
    for (int k = 0; k < 9; ++k) {
        act[k] = c->DisGradAvg[k] + c->acum[k];
        vga[k] = (act[k] - c->DisGradAvg_t[k]) / dt;
    }
    c->e_vm = von_mises(act);
    c->d_vm = von_mises(vga);
    c->TimeTot += dt;
    std::memcpy(c->DisGradAvg_t, act, sizeof(act));
    std::memcpy(r.disgrad_avg_actual, act, sizeof(act));
    std::memcpy(r.velgrad_avg, vga, sizeof(vga));
    // Edot/Eps, update_orient (istep>1), DislocationEvolution, averages: the end-of-increment kernel
    F.do_orient = (d.Update_Flag == 1 && c->istep > 1) ? 1 : 0;
    return 0;
}

static int finish_end(srx_ctx *c, Finish &F, srx_step_report *rep) {
    const Deck &d = c->dk;
    const double dt = d.TimeStep;
    const int N = (int)c->nvox;
    srx_step_report &r = F.r;
    const int iter = F.iter, nrmax = F.nrmax;
    const long nrtot = F.nrtot;
    const double nrwave = F.nrwave;
    int rc = 0;
    const double *h = c->hred;
    if (h[21] > 0) return SRX_E_NAN;
    {
        double e6[6], d6[6], et[9], dtt[9];
        for (int k = 0; k < 6; ++k) { e6[k] = h[k]; d6[k] = h[6 + k]; }
        b6_to_t9(e6, et);
        b6_to_t9(d6, dtt);
        double evmp = 0.0, dvmp = 0.0;
        for (int k = 0; k < 9; ++k) {
            r.eps_avg[k] = et[k] * c->wgt;
            r.edot_avg[k] = dtt[k] * c->wgt;
            evmp += r.eps_avg[k] * r.eps_avg[k];
            dvmp += r.edot_avg[k] * r.edot_avg[k];
        }
        r.evmp = std::sqrt(2. / 3. * evmp);
        r.dvmp = std::sqrt(2. / 3. * dvmp);
    }
    const double inv = 1.0 / c->nvox;
    r.rho_F_avg = h[12] * inv;
    r.rho_P_avg = h[13] * inv;
    r.rho_s_avg = h[14] * inv;
    r.rho_m_avg = h[15] * inv;
    r.rho_g_avg = 0.0;   // set below with the GND option
    r.rho_dipole_avg = h[16] * inv;
    r.rho_forest_avg = h[17] * inv;
    r.rho_inplane_avg = h[20] * inv;
    if (c->opt.gnd) {
        // GND densities: gradient of this increment's slip rates when Hard_Flag = 1 and istep > 1
        // (evolution.c:1346-1353), then the GND branch of DislocationEvolution and its mean
        launch_gnd(c->stream, c->dc, c->hc, c->f, (int)N, (d.Hard_Flag == 1 && c->istep > 1) ? 1 : 0, c->partials);
        launch_reduce(c->stream, c->partials, gnd_block_count((int)N), 1, c->red);
        if ((rc = sync_read(c, 1))) return rc;
        r.rho_g_avg = c->hred[0] * inv;
    }
    if (c->fdd) {
        std::fprintf(c->fdd, "%e\t%e\t%e\t%e\t%e\t%e\t%e\n", c->TimeTot, c->e_vm, r.rho_s_avg, r.rho_m_avg,
                     r.rho_dipole_avg, r.rho_inplane_avg, r.rho_forest_avg);
        std::fflush(c->fdd);
    }
    r.istep = c->istep;
    r.iters = iter;
    r.nr_iters_max = nrmax;
    r.nr_iters_total = nrtot;
    r.nr_wave_iters = nrwave;
    r.err_s = c->Err_s;
    r.err_e = c->Err_e;
    r.time_tot = c->TimeTot;
    r.e_vm = c->e_vm;
    r.d_vm = c->d_vm;
    r.s_vm = c->s_vm;
    r.s_vm1 = c->s_vm1;
    std::memcpy(r.sig_avg, c->SigAvg, sizeof(r.sig_avg));
    std::memcpy(r.sig_avg1, c->SigAvg1, sizeof(r.sig_avg1));
    if (c->fvm) {
        std::fprintf(c->fvm, "%e\t%e\t%e\t%e\t%e\t%e\t%e\n", c->TimeTot, c->e_vm, r.evmp, c->d_vm, r.dvmp, c->s_vm,
                     c->s_vm1);
        std::fflush(c->fvm);
    }
    int frc = 0;
    if (c->opt.write_fields && d.PrintControl[0] == 1 && c->istep % d.PrintControl[1] == 0)
        frc = write_field_dumps(c, c->istep);
    if (c->ferr) std::fflush(c->ferr);
    // evolution.c:1513-1515
    for (int k = 0; k < 9; ++k) c->DisGradAvg[k] = c->DisGradAvg_t[k] + c->Udot[k] * dt;
    collect_timing(c);
    c->istep++;
    if (rep) *rep = r;
    return frc;
}

static int step_finish(srx_ctx *c, srx_step_report *rep) {
    Finish F;
    int rc = finish_begin(c, F);
    if (rc) return rc;
    const int rotated = c->istep > 2 ? 1 : 0;
    { Timer t(c, 6); launch_end(c->stream, c->dc, c->f, (int)c->nvox, F.iter == 1 ? 1 : 0, rotated, F.do_orient, 1,
                                c->istep, c->partials, c->opt.gnd); }
    { Timer t(c, 7); launch_reduce(c->stream, c->partials, c->end_blocks, END_NPART, c->red); }
    if ((rc = sync_read(c, END_NPART))) return rc;
    return finish_end(c, F, rep);
}

int srx_step(srx_ctx *c, srx_step_report *rep) {
    if (!c) return SRX_E_ARG;
    if (c->istep >= c->nsteps) return SRX_E_DONE;
    if (c->creep == 1 && c->e_vm > 0.3) return SRX_E_CREEP_STOP;
    if (c->opt.warm_start == 2 && !c->mlp) return SRX_E_ARG;   // srx_mlp_set first
    int rc = step_begin(c);
    if (!rc) rc = step_iterate(c);
    if (!rc) rc = step_finish(c, rep);
    return rc;
}

int srx_run(srx_ctx *c) {
    if (!c) return SRX_E_ARG;
    while (c->istep < c->nsteps) {
        const int rc = srx_step(c, nullptr);
        if (rc == SRX_E_CREEP_STOP) break;
        if (rc) return rc;
    }
    return 0;
}

// Lockstep ensemble run: every member advances one increment at a time together.  The increment
// prologue, the Newton solve, its reduction and the iteration control, and the end-of-increment
// kernel and its reduction of all members run as one launch each over member tables; each member's
// FFT passes run on one of up to eight side streams (members spread over them).  The host reads
// back only the members' stop flags, once per batch of ITER_BATCH iterations, and their control
// blocks and reduced end-of-increment values once per increment.
static int run_lockstep(const std::vector<int> &members, std::vector<srx_ctx *> &ctx, srx_ensemble_result *results,
                        int max_side) {
    const int M = (int)members.size();
    if (M == 0) return 0;
    hipStream_t S = ctx[members[0]]->stream;
    const int NS = std::max(1, std::min(max_side, M));
    std::vector<hipStream_t> side(NS, nullptr);
    std::vector<hipEvent_t> ev(NS, nullptr);
    hipEvent_t fork = nullptr;
    EnsMember *dtab = nullptr, *htab = nullptr;
    EnsIncr *dinc = nullptr, *hinc = nullptr;
    int *ddone = nullptr, *hdone = nullptr;
    std::vector<hipStream_t> own(M);
    int rc = 0;
    auto ok = [&](hipError_t e) {
        if (e != hipSuccess && !rc) rc = e == hipErrorOutOfMemory ? SRX_E_NOMEM : SRX_E_HIP;
        return rc == 0;
    };
    for (int q = 0; q < NS && ok(hipStreamCreateWithFlags(&side[q], hipStreamNonBlocking)); ++q)
        ok(hipEventCreateWithFlags(&ev[q], hipEventDisableTiming));
    ok(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    ok(hipMalloc(&dtab, sizeof(EnsMember) * M));
    ok(hipHostMalloc((void **)&htab, sizeof(EnsMember) * M, hipHostMallocDefault));
    ok(hipMalloc(&dinc, sizeof(EnsIncr) * M));
    ok(hipHostMalloc((void **)&hinc, sizeof(EnsIncr) * M, hipHostMallocDefault));
    ok(hipMalloc(&ddone, sizeof(int) * M));
    ok(hipHostMalloc((void **)&hdone, sizeof(int) * M, hipHostMallocDefault));
    // the members order their own copies and kernels on the main stream during the run
    for (int j = 0; j < M; ++j) {
        own[j] = ctx[members[j]]->stream;
        ctx[members[j]]->stream = S;
    }
    while (!rc) {
        std::vector<int> act;
        for (int k : members) {
            srx_ctx *c = ctx[k];
            if (results[k].rc) continue;
            if (c->istep >= c->nsteps || (c->creep == 1 && c->e_vm > 0.3)) continue;
            if (c->opt.warm_start == 2 && !c->mlp) { results[k].rc = SRX_E_ARG; continue; }
            act.push_back(k);
        }
        if (act.empty()) break;
        const int nact = (int)act.size();
        const int istep = ctx[act[0]]->istep;
        for (int k : act)
            if (ctx[k]->istep != istep) rc = SRX_E_ARG;   // members advance together from istep 0
        if (rc) break;
        const int rotated = istep > 2 ? 1 : 0;
        int nb = 0, most = 0, pb = 0;
        for (int j = 0; j < nact && !rc; ++j) {
            srx_ctx *c = ctx[act[j]];
            EnsIncr &inc = hinc[j];
            if (int e = step_begin(c, &inc)) rc = e;
            inc.cst = c->dc;
            inc.f = c->f;
            inc.nvox = (int)c->nvox;
            inc.pblk0 = pb;
            inc.pnblk = prologue_block_count((int)c->nvox);
            pb += inc.pnblk;
            inc.partials = c->partials;
            inc.red = c->red;
            htab[j].cst = c->dc;
            htab[j].f = c->f;
            htab[j].nvox = (int)c->nvox;
            htab[j].blk0 = nb;
            htab[j].nblk = c->nr_blocks;
            htab[j].pad = 0;
            htab[j].partials = c->partials;
            htab[j].redseg = c->redseg;
            htab[j].ctl = c->dctl;
            nb += c->nr_blocks;
            most = std::max(most, std::max(c->itermax, c->opt.fixed_iters));
        }
        if (rc || !ok(hipMemcpyAsync(dtab, htab, sizeof(EnsMember) * nact, hipMemcpyHostToDevice, S)) ||
            !ok(hipMemcpyAsync(dinc, hinc, sizeof(EnsIncr) * nact, hipMemcpyHostToDevice, S)))
            break;
        launch_prologue_ens(S, dinc, nact, pb);
        const srx_options &o = ctx[act[0]]->opt;
        int ph2 = 0;
        for (int k : act) ph2 |= ctx[k]->hc.ph[1].xtal;
        int launched = 0;
        bool all_done = false;
        while (!all_done && !rc) {
            if (launched > most) { rc = SRX_E_HIP; break; }
            for (int b = 0; b < ITER_BATCH && ok(hipEventRecord(fork, S)); ++b) {
                ++launched;
                for (int q = 0; q < NS; ++q) ok(hipStreamWaitEvent(side[q], fork, 0));
                for (int j = 0; j < nact; ++j) {
                    srx_ctx *c = ctx[act[j]];
                    hipStream_t st = side[j % NS];
                    const int prec = c->opt.precision;
                    const int *tl = c->tiles;
                    IterCtl *g = c->dctl;
                    launch_fft_fwd_z(st, prec, c->f.sig, c->specB, c->n, c->pl[2], c->tw[2], g, tl[0]);
                    launch_fft_y(st, prec, c->specB, c->specC, 6, -1, 1, c->n, c->pl[1], c->tw[1], g, tl[1]);
                    launch_green_x(st, prec, c->specC, c->n, c->pl[0], c->tw[0], c->dc, g, tl[2]);
                    launch_fft_y(st, prec, c->specC, c->specB, 9, +1, 0, c->n, c->pl[1], c->tw[1], g, tl[3]);
                    launch_fft_inv_z_update(st, prec, c->specB, c->f.dg, c->n, c->pl[2], c->tw[2], g,
                                            (double)c->nvox, tl[4]);
                }
                for (int q = 0; q < NS; ++q) ok(hipEventRecord(ev[q], side[q])) && ok(hipStreamWaitEvent(S, ev[q], 0));
                launch_nr_ens(S, dtab, nact, nb, launched == 1, rotated, o.warm_start, o.gnd, ph2);
                launch_reduce_nr_ens(S, dtab, nact);
                launch_control_ens(S, dtab, nact, ddone);
            }
            if (!ok(hipMemcpyAsync(hdone, ddone, sizeof(int) * nact, hipMemcpyDeviceToHost, S)) ||
                !ok(hipStreamSynchronize(S)))
                break;
            all_done = true;
            for (int j = 0; j < nact; ++j) all_done = all_done && hdone[j] != 0;
        }
        if (rc) break;
        for (int j = 0; j < nact; ++j)
            ok(hipMemcpyAsync(ctx[act[j]]->hctl, ctx[act[j]]->dctl, sizeof(IterCtl), hipMemcpyDeviceToHost, S));
        if (!ok(hipStreamSynchronize(S))) break;
        // the end of the increment: every member's end-of-increment kernel and its reduction in one
        // launch each (members whose iterations failed drop out of the table)
        std::vector<Finish> fin(nact);
        std::vector<int> fj;
        int eb = 0;
        for (int j = 0; j < nact; ++j) {
            srx_ctx *c = ctx[act[j]];
            if (const int e = finish_begin(c, fin[j])) {
                results[act[j]].rc = e;
                continue;
            }
            EnsIncr &inc = hinc[fj.size()];
            if ((int)fj.size() != j) inc = hinc[j];
            inc.blk0 = eb;
            inc.nblk = c->end_blocks;
            eb += c->end_blocks;
            inc.lagged = fin[j].iter == 1 ? 1 : 0;
            inc.orient = fin[j].do_orient;
            fj.push_back(j);
        }
        const int nfin = (int)fj.size();
        if (nfin) {
            if (!ok(hipMemcpyAsync(dinc, hinc, sizeof(EnsIncr) * nfin, hipMemcpyHostToDevice, S))) break;
            launch_end_ens(S, dinc, nfin, eb, istep, o.gnd);
            launch_reduce_end_ens(S, dinc, nfin);
            for (int j : fj)
                ok(hipMemcpyAsync(ctx[act[j]]->hred, ctx[act[j]]->red, sizeof(double) * END_NPART,
                                  hipMemcpyDeviceToHost, S));
            if (!ok(hipStreamSynchronize(S))) break;
        }
        for (int j : fj) {
            const int e = finish_end(ctx[act[j]], fin[j], &results[act[j]].last);
            if (e) results[act[j]].rc = e;
            else ++results[act[j]].increments;
        }
    }
    for (int j = 0; j < M; ++j) ctx[members[j]]->stream = own[j];
    for (int q = 0; q < NS; ++q) {
        if (side[q]) (void)hipStreamDestroy(side[q]);
        if (ev[q]) (void)hipEventDestroy(ev[q]);
    }
    if (fork) (void)hipEventDestroy(fork);
    if (dtab) (void)hipFree(dtab);
    if (dinc) (void)hipFree(dinc);
    if (ddone) (void)hipFree(ddone);
    if (htab) (void)hipHostFree(htab);
    if (hinc) (void)hipHostFree(hinc);
    if (hdone) (void)hipHostFree(hdone);
    return rc;
}

int srx_run_ensemble(const char *const *decks, const char *const *workdirs, int n, const srx_options *opt,
                     int concurrency, srx_ensemble_result *results) {
    if (!decks || n < 0 || (n > 0 && !results)) return SRX_E_ARG;
    if (n == 0) return 0;
    // largest grids first, so the last workers to finish hold the small ones
    std::vector<long> size(n, 0);
    for (int i = 0; i < n; ++i) {
        Deck d;
        if (decks[i] && read_deck(decks[i], d) == 0) size[i] = (long)d.CellDim[0] * d.CellDim[1] * d.CellDim[2];
    }
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return size[a] > size[b]; });
    const int workers = std::max(1, std::min(concurrency > 0 ? concurrency : 4, n));
    srx_options o;
    if (opt) o = *opt; else srx_default_options(&o);
    if (o.nthreads <= 0) o.nthreads = std::max(1, (int)std::thread::hardware_concurrency() / workers);
    std::vector<srx_ctx *> ctx(n, nullptr);
    for (int i = 0; i < n; ++i) {
        std::memset(&results[i], 0, sizeof(results[i]));
        results[i].index = i;
        results[i].voxels = size[i];
    }
    // `workers` host threads take members from a shared queue in `order`, running `job` on each
    auto pool = [&](auto job) {
        std::atomic<int> next{0};
        std::vector<std::thread> th;
        for (int w = 0; w < workers; ++w)
            th.emplace_back([&]() {
                for (int k; (k = next.fetch_add(1)) < n;) job(order[k]);
            });
        for (auto &t : th) t.join();
    };
    // 1. setup of every member (all stay resident: the run phase below is the increment loop alone)
    pool([&](int i) {
        const auto t0 = std::chrono::steady_clock::now();
        results[i].rc = srx_open(decks[i], workdirs ? workdirs[i] : nullptr, &o, &ctx[i]);
        results[i].setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    });
    // 2. the increments: all members in lockstep (concurrency 0) or `workers` members at a time, each
    // on its own stream
    const auto r0 = std::chrono::steady_clock::now();
    if (concurrency <= 0) {
        // lockstep groups: members dealt round-robin (largest first) into -concurrency groups (0: one),
        // each group in lockstep on its own streams from its own host thread, the groups concurrently
        const int ngroups = std::max(1, std::min(-concurrency, n));
        std::vector<std::vector<int>> groups(ngroups);
        int gi = 0;
        for (int k : order)
            if (results[k].rc == 0) groups[gi++ % ngroups].push_back(k);
        std::vector<int> grc(ngroups, 0);
        std::vector<std::thread> th;
        for (int g = 0; g < ngroups; ++g)
            th.emplace_back([&, g]() { grc[g] = run_lockstep(groups[g], ctx, results, std::max(1, 8 / ngroups)); });
        for (auto &t : th) t.join();
        const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - r0).count();
        for (int g = 0; g < ngroups; ++g)
            for (int k : groups[g]) {
                results[k].run_seconds = t;
                if (grc[g] && !results[k].rc) results[k].rc = grc[g];
            }
    } else
    pool([&](int i) {
        srx_ensemble_result &r = results[i];
        if (r.rc) return;
        const auto t0 = std::chrono::steady_clock::now();
        int rc = 0;
        while (rc == 0) {
            rc = srx_step(ctx[i], &r.last);
            if (rc == 0) ++r.increments;
        }
        r.rc = (rc == SRX_E_DONE || rc == SRX_E_CREEP_STOP) ? 0 : rc;
        r.run_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    });
    const double run_phase = std::chrono::duration<double>(std::chrono::steady_clock::now() - r0).count();
    // 3. teardown
    for (int i = 0; i < n; ++i) {
        results[i].run_phase_seconds = run_phase;
        if (ctx[i]) srx_close(ctx[i]);
    }
    for (int i = 0; i < n; ++i)
        if (results[i].rc) return results[i].rc;
    return 0;
}

int srx_hbm_stream(size_t bytes, int reps, double *best_gbs, double *mean_gbs, int *best_unroll, int *best_blocks) {
    if (!best_gbs || bytes < (1u << 20) || reps < 1) return SRX_E_ARG;
    const size_t n16 = bytes / 16;
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return SRX_E_HIP;
    if (hipMalloc(&a, n16 * 16) != hipSuccess || hipMalloc(&b, n16 * 16) != hipSuccess) rc = SRX_E_NOMEM;
    if (!rc && (hipMemset(a, 0, n16 * 16) != hipSuccess || hipMemset(b, 0, n16 * 16) != hipSuccess ||
                hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = SRX_E_HIP;
    double best = 0.0, mean = 0.0;
    int bu = 0, bb = 0;
    const int unrolls[3] = {1, 2, 4}, per_cu[4] = {4, 8, 16, 32};
    for (int iu = 0; iu < 3 && !rc; ++iu)
        for (int ib = 0; ib < 4 && !rc; ++ib) {
            const int blocks = cus * per_cu[ib];
            launch_stream_copy(nullptr, a, b, n16, unrolls[iu], blocks);
            double sum = 0.0, top = 0.0;
            for (int r = 0; r < reps && !rc; ++r) {
                float ms = 0.0f;
                if (hipEventRecord(e0, nullptr) != hipSuccess) { rc = SRX_E_HIP; break; }
                launch_stream_copy(nullptr, a, b, n16, unrolls[iu], blocks);
                if (hipEventRecord(e1, nullptr) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { rc = SRX_E_HIP; break; }
                const double gbs = 2.0 * (double)n16 * 16.0 / (ms * 1e-3) / 1e9;
                sum += gbs;
                top = std::max(top, gbs);
            }
            if (top > best) {
                best = top;
                mean = sum / reps;
                bu = unrolls[iu];
                bb = per_cu[ib];
            }
        }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (rc) return rc;
    *best_gbs = best;
    if (mean_gbs) *mean_gbs = mean;
    if (best_unroll) *best_unroll = bu;
    if (best_blocks) *best_blocks = bb;
    return 0;
}

int srx_kernel_times(srx_ctx *c, const char **names, double *ms, int *calls, int *count) {
    if (!c || !count) return SRX_E_ARG;
    const int n = std::min(*count, c->nkt);
    for (int i = 0; i < n; ++i) {
        if (names) names[i] = c->kt[i].name;
        if (ms) ms[i] = c->kt[i].ms;
        if (calls) calls[i] = c->kt[i].calls;
    }
    *count = n;
    return 0;
}

static int field_info(int id, int &nc) {
    switch (id) {
    case SRX_F_SIG: case SRX_F_EPS: case SRX_F_EDOT: case SRX_F_DISGRAD: case SRX_F_VELGRAD: case SRX_F_TRANMAT:
    case SRX_F_FLUCT: nc = 9; return 0;
    case SRX_F_RHO_S: case SRX_F_RHO_M: case SRX_F_RHO_F: case SRX_F_RHO_P: case SRX_F_GAMDOT: case SRX_F_GAM_ACUM:
    case SRX_F_RHO_DIPOLE: case SRX_F_RHO_FOREST: case SRX_F_RHO_INPLANE: case SRX_F_TRIAL_GAM_ACUM: nc = 12; return 0;
    case SRX_F_RHO_G1: case SRX_F_RHO_G2: nc = 12; return 0;
    case SRX_F_NR_ITERS: case SRX_F_PHASE: case SRX_F_GRAIN: case SRX_F_RHO_TOT: nc = 1; return 0;
    default: return SRX_E_FIELD;
    }
}

int srx_get_field(srx_ctx *c, int id, double *out, size_t n) {
    if (!c || !out) return SRX_E_ARG;
    int nc = 0;
    if (field_info(id, nc)) return SRX_E_FIELD;
    const long N = c->nvox;
    if (n < (size_t)nc * N) return SRX_E_SIZE;
    HIPCK(hipStreamSynchronize(c->stream));
    auto fetch = [&](const void *src, size_t bytes, void *dst) -> int {
        if (int e = ctx_copy(c->stream, dst, src, bytes, hipMemcpyDeviceToHost)) return e;
        return 0;
    };
    std::vector<double> tmp;
    int rc = 0;
    if (id == SRX_F_SIG || id == SRX_F_EPS || id == SRX_F_EDOT) {
        const int m = id == SRX_F_SIG ? 6 : 5;
        const double *src = id == SRX_F_SIG ? c->f.sig : id == SRX_F_EPS ? c->f.eps : c->f.edot;
        tmp.resize((size_t)m * N);
        if ((rc = fetch(src, sizeof(double) * m * N, tmp.data()))) return rc;
        for (long v = 0; v < N; ++v) {
            double b6[6] = {0, 0, 0, 0, 0, 0};
            for (int k = 0; k < m; ++k) b6[k] = tmp[k * N + v];
            b6_to_t9(b6, out + 9 * v, 6);
        }
        return 0;
    }
    if (id == SRX_F_DISGRAD || id == SRX_F_TRANMAT || id == SRX_F_VELGRAD || id == SRX_F_FLUCT) {
        tmp.resize((size_t)9 * N);
        if ((rc = fetch(id == SRX_F_TRANMAT ? c->f.tran : c->f.dg, sizeof(double) * 9 * N, tmp.data()))) return rc;
        std::vector<double> t0;
        if (id == SRX_F_VELGRAD || (id == SRX_F_FLUCT && c->f.fl0)) {
            t0.resize((size_t)9 * N);
            if ((rc = fetch(id == SRX_F_VELGRAD ? c->f.dg0 : c->f.fl0, sizeof(double) * 9 * N, t0.data()))) return rc;
        }
        for (long v = 0; v < N; ++v)
            for (int k = 0; k < 9; ++k) {
                double x = tmp[k * N + v];
                if (id == SRX_F_VELGRAD) x = (x - t0[k * N + v]) / c->dk.TimeStep;
                if (id == SRX_F_FLUCT) x = (x - c->uniform[k]) + (t0.empty() ? 0.0 : t0[k * N + v]);
                out[9 * v + k] = x;
            }
        return 0;
    }
    if (nc == 12) {
        const double *src = id == SRX_F_RHO_S ? c->f.rho_s : id == SRX_F_RHO_M ? c->f.rho_m : id == SRX_F_RHO_F ? c->f.rho_F
                          : id == SRX_F_RHO_P ? c->f.rho_P : id == SRX_F_GAMDOT ? c->f.gamdot
                          : id == SRX_F_RHO_DIPOLE ? c->f.rho_dipole : id == SRX_F_RHO_FOREST ? c->f.rho_forest
                          : id == SRX_F_RHO_INPLANE ? c->f.rho_inplane : id == SRX_F_RHO_G1 ? c->hc.rho_g1
                          : id == SRX_F_RHO_G2 ? c->hc.rho_g2 : c->f.gam_acum;
        if (!src) return SRX_E_FIELD;   // GND densities without the GND option
        tmp.resize((size_t)12 * N);
        if ((rc = fetch(src, sizeof(double) * 12 * N, tmp.data()))) return rc;
        for (long v = 0; v < N; ++v)
            for (int k = 0; k < 12; ++k) out[12 * v + k] = tmp[k * N + v];
        return 0;
    }
    if (id == SRX_F_NR_ITERS || id == SRX_F_PHASE) {
        std::vector<int> t(N);
        if ((rc = fetch(id == SRX_F_NR_ITERS ? c->f.nr_it : c->f.phase, sizeof(int) * N, t.data()))) return rc;
        for (long v = 0; v < N; ++v) out[v] = t[v];
        return 0;
    }
    if (id == SRX_F_GRAIN) {
        for (long v = 0; v < N; ++v) out[v] = c->grain[v];
        return 0;
    }
    if (id == SRX_F_RHO_TOT) return fetch(c->f.rho_tot, sizeof(double) * N, out);
    return SRX_E_FIELD;
}

int srx_set_field(srx_ctx *c, int id, const double *in, size_t n) {
    if (!c || !in) return SRX_E_ARG;
    int nc = 0;
    if (field_info(id, nc)) return SRX_E_FIELD;
    const long N = c->nvox;
    if (n < (size_t)nc * N) return SRX_E_SIZE;
    HIPCK(hipStreamSynchronize(c->stream));
    std::vector<double> tmp;
    if (id == SRX_F_SIG || id == SRX_F_EPS) {
        const int m = id == SRX_F_SIG ? 6 : 5;
        tmp.resize((size_t)m * N);
        for (long v = 0; v < N; ++v) {
            double b6[6];
            t9_to_b6(in + 9 * v, b6);
            for (int k = 0; k < m; ++k) tmp[k * N + v] = b6[k];
        }
        if (int e = ctx_copy(c->stream, id == SRX_F_SIG ? c->f.sig : c->f.eps, tmp.data(), sizeof(double) * m * N,
                             hipMemcpyHostToDevice))
            return e;
        if (id == SRX_F_SIG) {
            for (int k = 0; k < 6; ++k) {
                double s = 0;
                for (long v = 0; v < N; ++v) s += tmp[k * N + v];
                c->sig_sum6[k] = s;
            }
        }
        return 0;
    }
    if (id == SRX_F_DISGRAD || id == SRX_F_TRANMAT) {
        tmp.resize((size_t)9 * N);
        for (long v = 0; v < N; ++v)
            for (int k = 0; k < 9; ++k) tmp[k * N + v] = in[9 * v + k];
        if (int e = ctx_copy(c->stream, id == SRX_F_DISGRAD ? c->f.dg : c->f.tran, tmp.data(), sizeof(double) * 9 * N,
                             hipMemcpyHostToDevice))
            return e;
        return 0;
    }
    if (id == SRX_F_RHO_S || id == SRX_F_RHO_M || id == SRX_F_GAM_ACUM) {
        tmp.resize((size_t)12 * N);
        for (long v = 0; v < N; ++v)
            for (int k = 0; k < 12; ++k) tmp[k * N + v] = in[12 * v + k];
        double *dst = id == SRX_F_RHO_S ? c->f.rho_s : id == SRX_F_RHO_M ? c->f.rho_m : c->f.gam_acum;
        if (int e = ctx_copy(c->stream, dst, tmp.data(), sizeof(double) * 12 * N, hipMemcpyHostToDevice)) return e;
        return 0;
    }
    return SRX_E_FIELD;
}

static void drop_graphs(srx_ctx *c) {
    for (auto &row : c->gexec)
        for (auto &g : row)
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
}

void srx_close(srx_ctx *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->fvm) std::fclose(c->fvm);
    if (c->fdd) std::fclose(c->fdd);
    if (c->fdrx) std::fclose(c->fdrx);
    if (c->ferr) std::fclose(c->ferr);
    double *ptrs[] = {c->f.sig, c->f.sig_pre, c->f.eps, c->f.edot, c->f.dg, c->f.dg0, c->f.tran, c->f.tran0,
                      c->f.rho_s, c->f.rho_m, c->f.rho_m_old, c->f.rho_F, c->f.rho_P, c->f.gamdot, c->f.gam_acum,
                      c->f.rho_dipole, c->f.rho_forest, c->f.rho_inplane, c->f.rho_tot, c->f.fl0, c->f.acr, c->f.nrc,
                      c->hc.rho_g1, c->hc.rho_g2, c->hc.gnd_fp,
                      c->partials, c->red, c->redseg};
    for (double *p : ptrs)
        if (p) (void)hipFree(p);
    if (c->f.dsig) (void)hipFree(c->f.dsig);
    if (c->f.ecp) (void)hipFree(c->f.ecp);
    if (c->hc.gbc) (void)hipFree(c->hc.gbc);
    if (c->mlp) (void)hipFree(c->mlp);
    if (c->mlp_blob) (void)hipFree(c->mlp_blob);
    for (void *p : {(void *)c->rec_x, (void *)c->rec_t, (void *)c->rec_nd, (void *)c->rec_count})
        if (p) (void)hipFree(p);
    if (c->f.phase) (void)hipFree(c->f.phase);
    if (c->f.nr_it) (void)hipFree(c->f.nr_it);
    if (c->specB) (void)hipFree(c->specB);
    if (c->specC) (void)hipFree(c->specC);
    for (int i = 0; i < 3; ++i)
        if (c->tw[i]) (void)hipFree(c->tw[i]);
    if (c->dc) (void)hipFree(c->dc);
    if (c->hred) (void)hipHostFree(c->hred);
    if (c->dctl) (void)hipFree(c->dctl);
    if (c->hctl) (void)hipHostFree(c->hctl);
    for (hipEvent_t e : c->evpool) (void)hipEventDestroy(e);
    drop_graphs(c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

// ---- learned Newton start --------------------------------------------------
int srx_mlp_set(srx_ctx *c, const float *blob) {
    if (!c || !blob || c->opt.warm_start != 2) return SRX_E_ARG;
    for (int k = 0; k < 6; ++k)
        if (!(blob[SRX_MLP_NPARAM + 48 + 6 + k] != 0.0f)) return SRX_E_ARG;   // target spread
    if (!c->mlp) {
        HIPCK(hipMalloc(&c->mlp, mlp_packed_bytes()));
        HIPCK(hipMalloc(&c->mlp_blob, sizeof(float) * SRX_MLP_BLOB));
    }
    HIPCK(hipMemcpyAsync(c->mlp_blob, blob, sizeof(float) * SRX_MLP_BLOB, hipMemcpyHostToDevice, c->stream));
    mlp_pack(c->stream, c->mlp_blob, c->mlp);
    HIPCK(hipGetLastError());
    HIPCK(hipStreamSynchronize(c->stream));
    c->f.mlp = c->mlp;
    drop_graphs(c);   // captured batches hold the previous kernel arguments
    return 0;
}

static void free_recorder(srx_ctx *c) {
    for (void *p : {(void *)c->rec_x, (void *)c->rec_t, (void *)c->rec_nd, (void *)c->rec_count})
        if (p) (void)hipFree(p);
    c->rec_x = c->rec_t = c->rec_nd = nullptr;
    c->rec_count = nullptr;
    c->rec_stride = c->rec_cap = 0;
}

int srx_mlp_record(srx_ctx *c, int stride, size_t capacity) {
    if (!c || stride < 0) return SRX_E_ARG;
    HIPCK(hipStreamSynchronize(c->stream));
    free_recorder(c);
    if (stride == 0) return 0;
    if (!c->f.dsig || capacity == 0 || capacity > ((size_t)1 << 28)) return SRX_E_ARG;
    if (!c->f.ecp) {
        // the features need the previous Newton call's eps_c; graphs captured without it are dropped
        HIPCK(hipMalloc(&c->f.ecp, sizeof(float) * 6 * (size_t)c->nvox));
        if (int e = ctx_zero(c->stream, c->f.ecp, sizeof(float) * 6 * (size_t)c->nvox)) return e;
        drop_graphs(c);
    }
    const size_t cap = capacity;
    HIPCK(hipMalloc(&c->rec_x, sizeof(float) * 24 * cap));
    HIPCK(hipMalloc(&c->rec_t, sizeof(float) * 6 * cap));
    HIPCK(hipMalloc(&c->rec_nd, sizeof(float) * cap));
    HIPCK(hipMalloc(&c->rec_count, sizeof(int)));
    HIPCK(hipMemsetAsync(c->rec_count, 0, sizeof(int), c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    c->rec_stride = stride;
    c->rec_cap = (int)cap;
    return 0;
}

int srx_mlp_samples(srx_ctx *c, float *x, float *t, size_t max, size_t *n) {
    if (!c || !n) return SRX_E_ARG;
    *n = 0;
    if (!c->rec_count) return 0;
    int cnt = 0;
    HIPCK(hipStreamSynchronize(c->stream));
    if (int e = ctx_copy(c->stream, &cnt, c->rec_count, sizeof(int), hipMemcpyDeviceToHost)) return e;
    std::vector<float> hx((size_t)cnt * 24), ht((size_t)cnt * 6), hn(cnt);
    if (cnt) {
        int e = 0;
        if ((e = ctx_copy(c->stream, hx.data(), c->rec_x, hx.size() * sizeof(float), hipMemcpyDeviceToHost)) ||
            (e = ctx_copy(c->stream, ht.data(), c->rec_t, ht.size() * sizeof(float), hipMemcpyDeviceToHost)) ||
            (e = ctx_copy(c->stream, hn.data(), c->rec_nd, hn.size() * sizeof(float), hipMemcpyDeviceToHost)))
            return e;
    }
    size_t k = 0;
    for (int i = 0; i < cnt; ++i) {
        bool ok = hn[i] > 0.0f && std::isfinite(hn[i]);
        for (int j = 0; j < 24 && ok; ++j) ok = std::isfinite(hx[(size_t)i * 24 + j]);
        for (int j = 0; j < 6 && ok; ++j) ok = std::isfinite(ht[(size_t)i * 6 + j]);
        if (!ok) continue;
        if (x && t) {
            if (k >= max) break;
            std::memcpy(x + k * 24, &hx[(size_t)i * 24], 24 * sizeof(float));
            std::memcpy(t + k * 6, &ht[(size_t)i * 6], 6 * sizeof(float));
        }
        ++k;
    }
    *n = k;
    return 0;
}

// ---- single-kernel entry points -------------------------------------------
static int fft_test_common(int ncomp, int n0, int n1, int n2, int prec, FftPlan pl[3], void *tw[3], void **sB,
                           void **sC, size_t &specn) {
    const int n[3] = {n0, n1, n2};
    for (int i = 0; i < 3; ++i)
        if (!make_plan(n[i], pl[i])) return SRX_E_GRID;
    if (!grid_supported(pl, n, prec) || ncomp < 1 || ncomp > 9) return SRX_E_GRID;
    for (int ax = 0; ax < 3; ++ax) {
        int r = prec == SRX_PREC_MIXED ? upload_twiddles<float>(n[ax], &tw[ax]) : upload_twiddles<double>(n[ax], &tw[ax]);
        if (r) return r;
    }
    const size_t H = n2 / 2 + 1;
    specn = (size_t)ncomp * H * n0 * n1;
    const size_t csz = prec == SRX_PREC_MIXED ? 8 : 16;
    HIPCK(hipMalloc(sB, specn * csz));
    HIPCK(hipMalloc(sC, specn * csz));
    return 0;
}

int srx_fft_forward(const double *in, int ncomp, int n0, int n1, int n2, int prec, double *out) {
    FftPlan pl[3];
    void *tw[3] = {nullptr, nullptr, nullptr}, *sB = nullptr, *sC = nullptr;
    size_t specn = 0;
    int rc = fft_test_common(ncomp, n0, n1, n2, prec, pl, tw, &sB, &sC, specn);
    const size_t N = (size_t)n0 * n1 * n2;
    double *din = nullptr;
    const int n[3] = {n0, n1, n2};
    if (!rc) {
        HIPCK(hipMalloc(&din, sizeof(double) * ncomp * N));
        HIPCK(hipMemcpy(din, in, sizeof(double) * ncomp * N, hipMemcpyHostToDevice));
        launch_fft_real_forward_test(nullptr, prec, din, ncomp, sB, sC, n, pl, tw);
        HIPCK(hipDeviceSynchronize());
        // layout C [c][kz][y][x] -> out [c][x][y][kz]
        const size_t H = n2 / 2 + 1;
        const size_t csz = prec == SRX_PREC_MIXED ? 8 : 16;
        std::vector<char> h(specn * csz);
        HIPCK(hipMemcpy(h.data(), sC, specn * csz, hipMemcpyDeviceToHost));
        for (int c = 0; c < ncomp; ++c)
            for (size_t kz = 0; kz < H; ++kz)
                for (int y = 0; y < n1; ++y)
                    for (int x = 0; x < n0; ++x) {
                        const size_t src = ((c * H + kz) * n1 + y) * n0 + x;
                        const size_t dst = (((size_t)c * n0 + x) * n1 + y) * H + kz;
                        if (prec == SRX_PREC_MIXED) {
                            const float *f = reinterpret_cast<const float *>(h.data()) + 2 * src;
                            out[2 * dst] = f[0];
                            out[2 * dst + 1] = f[1];
                        } else {
                            const double *f = reinterpret_cast<const double *>(h.data()) + 2 * src;
                            out[2 * dst] = f[0];
                            out[2 * dst + 1] = f[1];
                        }
                    }
    }
    if (din) (void)hipFree(din);
    if (sB) (void)hipFree(sB);
    if (sC) (void)hipFree(sC);
    for (int i = 0; i < 3; ++i)
        if (tw[i]) (void)hipFree(tw[i]);
    return rc;
}

int srx_fft_inverse(const double *in, int ncomp, int n0, int n1, int n2, int prec, double scale, double *out) {
    FftPlan pl[3];
    void *tw[3] = {nullptr, nullptr, nullptr}, *sB = nullptr, *sC = nullptr;
    size_t specn = 0;
    int rc = fft_test_common(ncomp, n0, n1, n2, prec, pl, tw, &sB, &sC, specn);
    const size_t N = (size_t)n0 * n1 * n2;
    double *dout = nullptr;
    const int n[3] = {n0, n1, n2};
    if (!rc) {
        const size_t H = n2 / 2 + 1;
        const size_t csz = prec == SRX_PREC_MIXED ? 8 : 16;
        std::vector<char> h(specn * csz);
        for (int c = 0; c < ncomp; ++c)
            for (size_t kz = 0; kz < H; ++kz)
                for (int y = 0; y < n1; ++y)
                    for (int x = 0; x < n0; ++x) {
                        const size_t dst = ((c * H + kz) * n1 + y) * n0 + x;
                        const size_t src = (((size_t)c * n0 + x) * n1 + y) * H + kz;
                        if (prec == SRX_PREC_MIXED) {
                            float *f = reinterpret_cast<float *>(h.data()) + 2 * dst;
                            f[0] = (float)in[2 * src];
                            f[1] = (float)in[2 * src + 1];
                        } else {
                            double *f = reinterpret_cast<double *>(h.data()) + 2 * dst;
                            f[0] = in[2 * src];
                            f[1] = in[2 * src + 1];
                        }
                    }
        HIPCK(hipMemcpy(sC, h.data(), specn * csz, hipMemcpyHostToDevice));
        HIPCK(hipMalloc(&dout, sizeof(double) * ncomp * N));
        launch_fft_real_inverse_test(nullptr, prec, sC, sB, ncomp, dout, n, pl, tw, 1.0 / scale);
        HIPCK(hipDeviceSynchronize());
        HIPCK(hipMemcpy(out, dout, sizeof(double) * ncomp * N, hipMemcpyDeviceToHost));
    }
    if (dout) (void)hipFree(dout);
    if (sB) (void)hipFree(sB);
    if (sC) (void)hipFree(sC);
    for (int i = 0; i < 3; ++i)
        if (tw[i]) (void)hipFree(tw[i]);
    return rc;
}

}  // extern "C"
