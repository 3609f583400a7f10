// voce_ctx.cpp -- host side of the C ABI: setup arithmetic (SetupJob,
// init.c:617-828), device state, the per-increment launch sequence and
// field read-back.  Setup runs once per context on the host in FP64; the
// per-voxel work of every increment runs on the device.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <array>
#include <map>
#include <tuple>
#include <vector>

#include "../../include/voce.h"
#include "../../include/voce_prof.h"
#include "../../include/voce_batch.h"
#include "voce_internal.h"
#include "voce_host.h"

namespace {

const double RSQ2 = 0.70710678118654744;   // kinematics.c:22-24
const double RSQ3 = 0.57735026918962584;
const double RSQ6 = 0.40824829046386304;
const double PI_REF = 3.14159265359;       // evp.h:57

// Lebensohn-Tome basis (kinematics.c:28-54)
struct Basis {
    double b[6][3][3];
    Basis()
    {
        memset(b, 0, sizeof(b));
        b[1][0][0] = -1. * RSQ6; b[1][1][1] = -1. * RSQ6; b[1][2][2] = 2. * RSQ6;
        b[0][0][0] = -1. * RSQ2; b[0][1][1] = RSQ2;
        b[2][1][2] = RSQ2; b[2][2][1] = RSQ2;
        b[3][0][2] = RSQ2; b[3][2][0] = RSQ2;
        b[4][0][1] = RSQ2; b[4][1][0] = RSQ2;
        b[5][0][0] = RSQ3; b[5][1][1] = RSQ3; b[5][2][2] = RSQ3;
    }
};
const Basis BB;

// chg_basis opt 1 / opt 2 (nc = 6 or 5)
void v_to_t(const double *V, double T[3][3], int nc)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            T[i][j] = 0.0;
            for (int a = 0; a < nc; ++a) T[i][j] += V[a] * BB.b[a][i][j];
        }
}
void t_to_v(const double T[3][3], double *V, int nc)
{
    for (int a = 0; a < nc; ++a) {
        V[a] = 0.0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) V[a] += T[i][j] * BB.b[a][i][j];
    }
}
// chg_basis opt 3: 6x6 -> 3x3x3x3
void c66_to_c4(const double C[6][6], double T4[3][3][3][3])
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l) {
                    double s = 0.0;
                    for (int a = 0; a < 6; ++a)
                        for (int b = 0; b < 6; ++b) s += C[a][b] * BB.b[a][i][j] * BB.b[b][k][l];
                    T4[i][j][k][l] = s;
                }
}
// chg_basis opt 4: 3x3x3x3 -> 6x6
void c4_to_c66(const double T4[3][3][3][3], double C[6][6])
{
    for (int a = 0; a < 6; ++a)
        for (int b = 0; b < 6; ++b) {
            double s = 0.0;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    for (int k = 0; k < 3; ++k)
                        for (int l = 0; l < 3; ++l) s += T4[i][j][k][l] * BB.b[a][i][j] * BB.b[b][k][l];
            C[a][b] = s;
        }
}

// Crout LU inverse with implicit scaling (kinematics.c:282-424); 35 on a
// row whose largest magnitude is below 1e-5.
int lu_inv66(double c[6][6])
{
    double a[6][6], vv[6], y[6][6];
    int indx[6];
    memcpy(a, c, sizeof(a));
    for (int i = 0; i < 6; ++i) {
        double big = 0.0;
        for (int j = 0; j < 6; ++j) big = fmax(big, fabs(a[i][j]));
        if (big < 1E-5) return VOCE_ERR_SINGULAR;
        vv[i] = 1.0 / big;
    }
    for (int j = 0; j < 6; ++j) {
        for (int i = 0; i < j; ++i) {
            double sum = a[i][j];
            for (int k = 0; k < i; ++k) sum -= a[i][k] * a[k][j];
            a[i][j] = sum;
        }
        double big = 0.0;
        int imax = j;
        for (int i = j; i < 6; ++i) {
            double sum = a[i][j];
            for (int k = 0; k < j; ++k) sum -= a[i][k] * a[k][j];
            a[i][j] = sum;
            const double dum = vv[i] * fabs(sum);
            if (dum >= big) { big = dum; imax = i; }
        }
        if (j != imax) {
            for (int k = 0; k < 6; ++k) {
                const double d = a[imax][k];
                a[imax][k] = a[j][k];
                a[j][k] = d;
            }
            vv[imax] = vv[j];
        }
        indx[j] = imax;
        if (fabs(a[j][j]) < 1E-5) a[j][j] = 1.0E-20;
        if (j != 5) {
            const double dum = 1.0 / a[j][j];
            for (int i = j + 1; i < 6; ++i) a[i][j] *= dum;
        }
    }
    for (int col = 0; col < 6; ++col) {
        double b[6] = {0, 0, 0, 0, 0, 0};
        b[col] = 1.0;
        int ii = -1;
        for (int i = 0; i < 6; ++i) {
            const int ip = indx[i];
            double sum = b[ip];
            b[ip] = b[i];
            if (ii >= 0)
                for (int j = ii; j <= i - 1; ++j) sum -= a[i][j] * b[j];
            else if (fabs(sum) > 1E-5)
                ii = i;
            b[i] = sum;
        }
        for (int i = 5; i >= 0; --i) {
            double sum = b[i];
            for (int j = i + 1; j < 6; ++j) sum -= a[i][j] * b[j];
            b[i] = sum / a[i][i];
        }
        for (int i = 0; i < 6; ++i) y[i][col] = b[i];
    }
    memcpy(c, y, sizeof(y));
    return 0;
}

// EulerToTransMatrix opt 2 (kinematics.c:218-234), Bunge, sample -> crystal
void euler_to_mat(double t1, double ph, double t2, double a[3][3])
{
    const double ct1 = cos(t1), st1 = sin(t1), cph = cos(ph), sph = sin(ph), ct2 = cos(t2), st2 = sin(t2);
    a[0][0] = ct1 * ct2 - st1 * st2 * cph;
    a[0][1] = st1 * ct2 + ct1 * st2 * cph;
    a[0][2] = st2 * sph;
    a[1][0] = -1.0 * ct1 * st2 - st1 * ct2 * cph;
    a[1][1] = -1.0 * st1 * st2 + ct1 * ct2 * cph;
    a[1][2] = ct2 * sph;
    a[2][0] = st1 * sph;
    a[2][1] = -1.0 * ct1 * sph;
    a[2][2] = cph;
}
// EulerToTransMatrix opt 1 (kinematics.c:198-217): matrix -> degrees
void mat_to_euler(const double a[3][3], double *t1, double *ph, double *t2)
{
    double p = acos(a[2][2]), x1, x2;
    if (fabs(a[2][2]) > 0.9999) {
        x2 = 0.;
        x1 = atan2(a[0][1], a[0][0]);
    } else {
        const double sph = sin(p);
        x1 = atan2(a[2][0] / sph, -1.0 * a[2][1] / sph);
        x2 = atan2(a[0][2] / sph, a[1][2] / sph);
    }
    *ph = p * (180. / PI_REF);
    *t1 = x1 * (180. / PI_REF);
    *t2 = x2 * (180. / PI_REF);
}

// Ten4thTransform opt 2 (kinematics.c:262-273): c1 = a^T-rotation of c2
void ten4_rot_t(const double a[3][3], const double c2[3][3][3][3], double c1[3][3][3][3])
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l) {
                    double s = 0.0;
                    for (int m = 0; m < 3; ++m)
                        for (int n = 0; n < 3; ++n)
                            for (int p = 0; p < 3; ++p)
                                for (int q = 0; q < 3; ++q) s += a[m][i] * a[n][j] * a[p][k] * a[q][l] * c2[m][n][p][q];
                    c1[i][j][k][l] = s;
                }
}

// ElastStiffnessMatrix (init.c:9-49), cubic
void cubic_c4(const double e[3], double c[3][3][3][3])
{
    memset(c, 0, sizeof(double) * 81);
    const double c11 = e[0], c12 = e[1], c44 = e[2];
    for (int i = 0; i < 3; ++i) c[i][i][i][i] = c11;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (i != j) {
                c[i][i][j][j] = c12;
                c[i][j][i][j] = c44;
                c[i][j][j][i] = c44;
            }
}

void voigt_to_full(const double v[6], double m[3][3])
{
    m[0][0] = v[0]; m[0][1] = v[5]; m[0][2] = v[4];
    m[1][0] = v[5]; m[1][1] = v[1]; m[1][2] = v[3];
    m[2][0] = v[4]; m[2][1] = v[3]; m[2][2] = v[2];
}

#define CK(x)                                         \
    do {                                              \
        hipError_t e_ = (x);                          \
        if (e_ != hipSuccess) {                       \
            fprintf(stderr, "voce: %s failed: %s\n", #x, hipGetErrorString(e_)); \
            return VOCE_ERR_DEVICE;                   \
        }                                             \
    } while (0)

}  // namespace

int voce_dev_alloc(voce_ctx *c, void **p, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return VOCE_ERR_NOMEM;
    c->allocs.push_back(*p);
    return 0;
}

// Per-increment workspace of one RVE (pre-update stress, crystal-frame K_c,
// spectra, skew sums, block partials), each part 256-byte aligned.
static size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t voce_ws_bytes(const RveDev &d)
{
    const size_t cb = d.fft_f32 ? 8 : 16, db = sizeof(double);
    return al256(6 * d.V * db) + al256(VOCE_KCV_COMP * voce_kcv_voxels(d.V) * db) + al256(6 * d.nspec * cb) +
           al256(3 * d.nspec * cb) + al256((size_t)d.nblk_vox * 16 * db);
}

void voce_ws_carve(RveDev &d, char *p)
{
    const size_t cb = d.fft_f32 ? 8 : 16, db = sizeof(double);
    d.P6 = (double *)p;
    p += al256(6 * d.V * db);
    d.KCV = (double *)p;
    p += al256(VOCE_KCV_COMP * voce_kcv_voxels(d.V) * db);
    d.spec = p;
    p += al256(6 * d.nspec * cb);
    d.usum = p;
    p += al256(3 * d.nspec * cb);
    d.part = (double *)p;
}

// Push the host descriptor to its device copy and refresh the launch set.
int voce_sync_tab(voce_ctx *c)
{
    const RveDev &d = c->d;
    if (c->on_tab) *c->on_tab = d;
    if (hipMemcpyAsync(c->dtab, &d, sizeof(RveDev), hipMemcpyHostToDevice, c->stream) != hipSuccess)
        return VOCE_ERR_DEVICE;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return VOCE_ERR_DEVICE;
    LaunchSet &L = c->L;
    memset(&L, 0, sizeof(L));
    L.tab = c->dtab;
    L.n_pfwd = d.nx * 6 + 1;   // + the finaliser block
    L.n_plane6 = d.nx * 6;
    L.n_plane3 = d.nx * 3;
    L.n_x = (d.ny * d.nzh + d.tk - 1) / d.tk;
    L.n_vox = d.nblk_vox;
    L.n_one = 1;
    L.plv = d.plv;
    L.p2_n = d.ny;
    L.p2_nx = d.nx;
    L.sp_ny = d.ny;
    L.sp_nz = d.nz;
    L.p2_tk = d.tk;
    L.xv = d.xv;
    L.pl_lds = voce_plane_lds(&d, d.plv);
    L.x_lds = voce_xgreen_lds(&d, d.xv, d.tk);
    if (!L.pl_lds || !L.x_lds) return VOCE_ERR_UNSUPPORTED_GRID;
    L.fft_f32 = d.fft_f32;
    L.uniform_phase = d.uniform_phase;
    L.two_crss = d.two_crss;
    return 0;
}

extern "C" int voce_field_ncomp(int f)
{
    switch (f) {
    case VOCE_FIELD_SIG: case VOCE_FIELD_EPS: case VOCE_FIELD_EDOT: case VOCE_FIELD_ELS: return 6;
    case VOCE_FIELD_DISGRAD: case VOCE_FIELD_TRANMAT: case VOCE_FIELD_VELGRAD: return 9;
    case VOCE_FIELD_CRSS: case VOCE_FIELD_CRSS_BWD: case VOCE_FIELD_GAMTOT: case VOCE_FIELD_GAMDOT:
    case VOCE_FIELD_TAUX: return 12;
    case VOCE_FIELD_GAMACUM: case VOCE_FIELD_SVM: case VOCE_FIELD_EVM: case VOCE_FIELD_WORK_P:
    case VOCE_FIELD_RHO_RSS: case VOCE_FIELD_RHO_CRSS: case VOCE_FIELD_WORK_E: return 1;
    case VOCE_FIELD_NEW_POSITION: return 3;
    case VOCE_FIELD_SCHMID: return 60;
    }
    return -1;
}

extern "C" const char *voce_strerror(int code)
{
    switch (code) {
    case VOCE_OK: return "ok";
    case VOCE_ERR_SLIP_NOT_ORTHOGONAL: return "slip systems are not orthogonal (reference abort 7)";
    case VOCE_ERR_SINGULAR: return "singular matrix in LU decomposition (reference exit 35)";
    case VOCE_ERR_BC: return "inconsistent strain-rate / stress boundary conditions (reference abort 114)";
    case VOCE_ERR_BC_DEVIATORIC: return "cannot enforce only two deviatoric components (reference abort 116)";
    case VOCE_ERR_MESH_INDEX: return "inconsistent index in the microstructure file (reference abort 1139)";
    case VOCE_ERR_ARG: return "invalid argument";
    case VOCE_ERR_FILE: return "cannot open or parse an input file";
    case VOCE_ERR_UNSUPPORTED_GRID: return "grid size not supported (each dimension 2^a 3^b 5^c 7^d, 2..512)";
    case VOCE_ERR_DEVICE: return "HIP runtime error";
    case VOCE_ERR_NOMEM: return "device memory allocation failed";
    case VOCE_ERR_NOT_LOADED: return "HIP library not loaded";
    case VOCE_ERR_SCRATCH: return "RVE workspace larger than the static no-allocation workspace (96 MiB)";
    }
    return "unknown error";
}

extern "C" long voce_ctx_nvox(const voce_ctx *c) { return c ? c->V : 0; }

extern "C" void voce_ctx_destroy(voce_ctx *c)
{
    if (!c) return;
    // teardown: errors here have no caller to go to
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void *p : c->allocs) (void)hipFree(p);
    for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
    if (c->h_ms) (void)hipHostFree(c->h_ms);
    if (c->h_ctl) (void)hipHostFree(c->h_ctl);
    if (c->h_tab) (void)hipHostFree(c->h_tab);
    if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

template <typename T>
static void make_twiddles(int n, std::vector<T> &tw)
{
    tw.resize(2 * n);
    for (int j = 0; j < n; ++j) {
        // exp(-2 pi i j / n) from the exact angle j/n of a turn
        long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)j / (long double)n;
        tw[2 * j] = (T)cosl(a);
        tw[2 * j + 1] = (T)(-sinl(a));
    }
}

static int setup_phase(const voce_phase &in, PhaseConst &pc, double C4[3][3][3][3])
{
    memset(&pc, 0, sizeof(pc));
    pc.gas = in.type != 0;
    cubic_c4(in.elast, C4);
    if (pc.gas) return 0;
    if (in.nsys < 1 || in.nsys > VOCE_NSYS) return VOCE_ERR_ARG;
    pc.nsys = in.nsys;
    pc.nsr = in.nsrs;
    pc.gam0 = in.gam0;
    pc.tau0f = in.tau0f;
    pc.tau0b = in.tau0b;
    pc.tau1 = in.tau1;
    pc.thet0 = in.thet0;
    pc.thet1 = in.thet1;
    if (pc.tau1 < 1.E-6) {   // init.c:141-144: linear hardening
        pc.tau1 = 1.E-6;
        pc.thet0 = pc.thet1;
    }
    pc.two_crss = (in.tau0f != in.tau0b);
    pc.int_n = (in.nsrs == floor(in.nsrs) && in.nsrs >= 1.0 && in.nsrs <= 64.0);
    for (int s = 0; s < pc.nsys; ++s) {
        double bn[3], bb[3], nn = 0, nb = 0, dot = 0;
        for (int m = 0; m < 3; ++m) { nn += pow(in.bn[s][m], 2.0); nb += pow(in.bb[s][m], 2.0); }
        nn = sqrt(nn);
        nb = sqrt(nb);
        for (int m = 0; m < 3; ++m) { bn[m] = in.bn[s][m] / nn; bb[m] = in.bb[s][m] / nb; }
        for (int m = 0; m < 3; ++m) dot += bn[m] * bb[m];
        if (dot > 1E-3) return VOCE_ERR_SLIP_NOT_ORTHOGONAL;   // init.c:213-215
        double a33[3][3], a5[5];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) a33[i][j] = (bn[i] * bb[j] + bn[j] * bb[i]) / 2.0;
        t_to_v(a33, a5, 5);    // Schm_xt (init.c:218-224)
        for (int i = 0; i < 5; ++i) pc.px[s][i] = a5[i];
        double p33[3][3];
        v_to_t(a5, p33, 5);    // as update_schmid expands it (kinematics.c:440-441)
        pc.p3[s][0] = p33[0][0]; pc.p3[s][1] = p33[1][1]; pc.p3[s][2] = p33[2][2];
        pc.p3[s][3] = p33[1][2]; pc.p3[s][4] = p33[0][2]; pc.p3[s][5] = p33[0][1];
        for (int m = 0; m < 3; ++m) {
            pc.bn[s][m] = pc.rn[s][m] = bn[m];
            pc.bb[s][m] = pc.rb[s][m] = bb[m];
        }
    }
    for (int i = 0; i < VOCE_NSYS; ++i)
        for (int j = 0; j < VOCE_NSYS; ++j) pc.hard[i][j] = in.hard[i][j];
    return 0;
}

extern "C" int voce_ctx_create(const voce_setup *su, voce_ctx **out)
{
    return voce_ctx_build(su, nullptr, out);
}

int voce_ctx_build(const voce_setup *su, const voce_build_opts *opt, voce_ctx **out)
{
    if (!out) return VOCE_ERR_ARG;
    *out = nullptr;
    if (!su || !su->euler_deg || !su->grain || !su->phase_id) return VOCE_ERR_ARG;
    const int nx = su->dims[0], ny = su->dims[1], nz = su->dims[2];
    if (nx <= 0 || ny <= 0 || nz <= 0 || su->n_phases < 1 || su->n_phases > 2) return VOCE_ERR_ARG;
    if (!voce_grid_supported(nx) || !voce_grid_supported(ny) || !voce_grid_supported(nz)) return VOCE_ERR_UNSUPPORTED_GRID;
    if (su->iter_max < 0) return VOCE_ERR_ARG;
    const long V = (long)nx * ny * nz;
    const double WGT = 1. / V;

    // materials (init.c:696-703)
    PhaseConst pcs[2];
    double Cph[2][3][3][3][3];
    memset(pcs, 0, sizeof(pcs));
    memset(Cph, 0, sizeof(Cph));
    for (int p = 0; p < su->n_phases; ++p) {
        int rc = setup_phase(su->phase[p], pcs[p], Cph[p]);
        if (rc) return rc;
    }
    if (su->n_phases == 2 && !pcs[0].gas && !pcs[1].gas) {
        // PlasticInit's normalisation loop runs over the mode index, not the
        // phase (init.c:195-224): the second phase's call normalises the
        // first phase's slip vectors once more and builds the second phase's
        // Schmid tensors from them, while the second phase's own vectors stay
        // as read, unnormalised, for update_orient (constitutive.c:202-203).
        // A second phase with more systems than the first would read unset
        // vectors there; it is refused.
        PhaseConst &a = pcs[0], &b = pcs[1];
        const voce_phase &ib = su->phase[1];
        if (b.nsys > a.nsys) return VOCE_ERR_ARG;
        for (int s = 0; s < b.nsys; ++s) {
            double nn = 0, nb = 0, dot = 0, bn[3], bb[3];
            for (int m = 0; m < 3; ++m) { nn += pow(a.bn[s][m], 2.0); nb += pow(a.bb[s][m], 2.0); }
            nn = sqrt(nn);
            nb = sqrt(nb);
            for (int m = 0; m < 3; ++m) { bn[m] = a.bn[s][m] / nn; bb[m] = a.bb[s][m] / nb; }
            for (int m = 0; m < 3; ++m) dot += bn[m] * bb[m];
            if (dot > 1E-3) return VOCE_ERR_SLIP_NOT_ORTHOGONAL;
            double a33[3][3], a5[5], p33[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) a33[i][j] = (bn[i] * bb[j] + bn[j] * bb[i]) / 2.0;
            t_to_v(a33, a5, 5);
            for (int i = 0; i < 5; ++i) b.px[s][i] = a5[i];
            v_to_t(a5, p33, 5);
            b.p3[s][0] = p33[0][0]; b.p3[s][1] = p33[1][1]; b.p3[s][2] = p33[2][2];
            b.p3[s][3] = p33[1][2]; b.p3[s][4] = p33[0][2]; b.p3[s][5] = p33[0][1];
            for (int m = 0; m < 3; ++m) {
                a.rn[s][m] = bn[m];   // the first phase rotates its twice-normalised vectors
                a.rb[s][m] = bb[m];
                b.bn[s][m] = bn[m];
                b.bb[s][m] = bb[m];
                b.rn[s][m] = ib.bn[s][m];
                b.rb[s][m] = ib.bb[s][m];
            }
        }
    }
    if (su->n_phases == 1) { pcs[1] = pcs[0]; pcs[1].gas = 1; pcs[1].nsys = 0; }
    // the Newton kernel evaluates the flow rule of the standard FCC slip
    // systems in a Walsh form whose constants are those of voce_fcc_schmid:
    // taken only when the phase's own Schmid vectors equal them to rounding
    for (int p = 0; p < 2; ++p) {
        PhaseConst &pc = pcs[p];
        int ok = !pc.gas && pc.nsys == 12;
        for (int sy = 0; ok && sy < 12; ++sy)
            for (int i = 0; i < 5; ++i)
                if (fabs(pc.px[sy][i] - voce_fcc_schmid(sy, i)) > 1e-14) ok = 0;
        pc.fcc12 = ok;
    }

    // orientations (InitMicroStruct, init.c:393-483), one table row per
    // distinct (phase, Euler triple)
    std::vector<uint32_t> ori(V);
    std::vector<uint8_t> phv(V);
    std::map<std::tuple<int, double, double, double>, uint32_t> omap;
    std::vector<double> Cgr;   // 36 per orientation
    std::vector<double> Sloc;  // 36 per orientation
    std::vector<std::array<double, 9>> Rtab;
    double c066[6][6];
    memset(c066, 0, sizeof(c066));
    long nph1 = 0;
    std::tuple<int, double, double, double> prev_key;
    uint32_t prev_o = 0;
    bool have_prev = false;
    for (long v = 0; v < V; ++v) {
        const int ph = su->phase_id[v];
        if (ph < 1 || ph > su->n_phases) return VOCE_ERR_MESH_INDEX;
        if (ph == 1) ++nph1;
        phv[v] = (uint8_t)ph;
        const voce_phase &P = su->phase[ph - 1];
        if (P.type != 0) {
            ori[v] = VOCE_GAS;
            continue;
        }
        const double e0 = su->euler_deg[3 * v], e1 = su->euler_deg[3 * v + 1], e2 = su->euler_deg[3 * v + 2];
        auto key = std::make_tuple(ph, e0, e1, e2);
        uint32_t o;
        if (have_prev && key == prev_key) {   // runs of one grain along z
            o = prev_o;
        } else {
        auto it = omap.find(key);
        if (it == omap.end()) {
            o = (uint32_t)Rtab.size();
            omap[key] = o;
            double a[3][3], c4[3][3][3][3], c66[6][6];
            euler_to_mat(e0 * (PI_REF / 180.), e1 * (PI_REF / 180.), e2 * (PI_REF / 180.), a);
            std::array<double, 9> R;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) R[i * 3 + j] = a[j][i];
            Rtab.push_back(R);
            ten4_rot_t(a, Cph[ph - 1], c4);
            c4_to_c66(c4, c66);
            for (int i = 0; i < 36; ++i) Cgr.push_back(c66[i / 6][i % 6]);
            int rc = lu_inv66(c66);   // S_local of UpdateStress (evolution.c:28-31)
            if (rc) return rc;
            for (int i = 0; i < 36; ++i) Sloc.push_back(c66[i / 6][i % 6]);
        } else {
            o = it->second;
        }
        prev_key = key;
        prev_o = o;
        have_prev = true;
        }
        ori[v] = o;
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) c066[i][j] += Cgr[o * 36 + i * 6 + j];
    }
    const long norient = (long)Rtab.size();
    MacroState *ms0 = (MacroState *)calloc(1, sizeof(MacroState));
    if (!ms0) return VOCE_ERR_NOMEM;
    double C066[6][6], S066[6][6], C0[3][3][3][3], S0[3][3][3][3];
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) C066[i][j] = c066[i][j] * WGT;
    memcpy(S066, C066, sizeof(S066));
    if (lu_inv66(S066)) { free(ms0); return VOCE_ERR_SINGULAR; }
    c66_to_c4(C066, C0);
    c66_to_c4(S066, S0);
    // per-orientation table (stride VOCE_OTAB): S_local = C_gr^-1, then
    // K_c = C066^-1 + S_local, the constant part of the compliance-form
    // Newton matrix of UpdateStress (evolution.c:85-93 with C066 factored out).
    // K_c is symmetric up to the rounding of the two inverses; it is stored
    // as the mean of itself and its transpose, so the Newton kernel reads one
    // triangle for both the residual and the LDL^T factorisation.
    std::vector<double> otab(VOCE_OTAB * (norient > 0 ? norient : 1), 0.0);
    for (long o = 0; o < norient; ++o) {
        for (int i = 0; i < 36; ++i) {
            const int a = i / 6, b = i % 6;
            otab[VOCE_OTAB * o + i] = Sloc[36 * o + i];
            otab[VOCE_OTAB * o + 36 + i] = 0.5 * ((S066[a][b] + Sloc[36 * o + i]) + (S066[b][a] + Sloc[36 * o + b * 6 + a]));
        }
        for (int i = 0; i < 9; ++i) otab[VOCE_OTAB * o + 72 + i] = Rtab[o][i];   // initial TranMat
        // Cubic elasticity (ElastStiffnessMatrix, init.c:9-49) keeps the
        // hydrostatic B-basis component an eigenvector of every C_gr, of their
        // mean C066 and of the inverses, so K_c couples it to the deviatoric
        // components by rounding only.  The Newton kernel solves the two parts
        // apart (a 5x5 deviatoric system and one division); a coupling beyond
        // rounding is refused.
        const double *K = &otab[VOCE_OTAB * o + 36];
        for (int j = 0; j < 5; ++j)
            if (!(fabs(K[30 + j]) <= 1e-12 * fabs(K[35]) && fabs(K[6 * j + 5]) <= 1e-12 * fabs(K[35]))) {
                free(ms0);
                return VOCE_ERR_ARG;
            }
    }

    // boundary conditions (init.c:713-745)
    const int *vf = su->velgrad_flag, *sf = su->stress_flag;
    if (vf[0] + vf[1] + vf[2] == 2) { free(ms0); return VOCE_ERR_BC_DEVIATORIC; }
    for (int i = 0; i < 6; ++i)
        if ((sf[i] * vf[i] != 0) || (sf[i] + vf[i] != 1)) { free(ms0); return VOCE_ERR_BC; }
    double Udot[3][3], Us[3][3], Ua[3][3], Sc[3][3];
    if (su->use_full_velgrad) {
        for (int i = 0; i < 9; ++i) Udot[i / 3][i % 3] = su->velgrad_full[i];
    } else {
        voigt_to_full(su->velgrad, Udot);
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            Us[i][j] = (Udot[i][j] + Udot[j][i]) / 2.;
            Ua[i][j] = (Udot[i][j] - Udot[j][i]) / 2.;
        }
    double D6[6], D5[5], devU[3][3];
    t_to_v(Us, D6, 6);
    for (int i = 0; i < 5; ++i) D5[i] = D6[i];
    v_to_t(D5, devU, 5);
    double d_vm = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) d_vm += pow(devU[i][j], 2.0);
    d_vm = sqrt(2. / 3. * d_vm);
    int creep = 0;
    for (int i = 0; i < 6; ++i) creep += sf[i];
    creep /= 6;
    voigt_to_full(su->stress, Sc);
    const double dt = su->time_step;
    for (int p = 0; p < 2; ++p) {
        PhaseConst &pc = pcs[p];
        const double a = 0.5 / sqrt(3.0), k = dt * pc.gam0 * pc.nsr;
        pc.wk[0] = pc.gam0 * a;
        pc.wk[1] = pc.gam0 * 0.5;
        pc.wk[2] = k / 12.0;
        pc.wk[3] = k * (0.5 * a);
        pc.wk[4] = k * 0.25;
    }

    // initial stress per orientation (InitFields, init.c:494-534)
    double DGA[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) DGA[i][j] = Udot[i][j] * dt;
    std::vector<double> sig6o(6 * (norient > 0 ? norient : 1));
    for (long o = 0; o < norient; ++o) {
        double c66[6][6], cg[3][3][3][3], sg[3][3];
        for (int i = 0; i < 36; ++i) c66[i / 6][i % 6] = Cgr[o * 36 + i];
        c66_to_c4(c66, cg);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double s = 0.0;
                for (int k = 0; k < 3; ++k)
                    for (int l = 0; l < 3; ++l) s += cg[i][j][k][l] * DGA[k][l];
                if (creep == 1) s += Sc[i][j];
                sg[i][j] = s;
            }
        t_to_v(sg, &sig6o[6 * o], 6);
    }

    // macro state
    MacroState &M = *ms0;
    M.dt = dt;
    M.err_tol = su->err;
    M.wgt = WGT;
    const double Wgt_ph1 = (1.0 * (double)nph1) / V;
    M.wgt_over_ph1 = WGT / Wgt_ph1;
    for (int i = 0; i < 36; ++i) {
        M.C066[i] = C066[i / 6][i % 6];
        M.S066[i] = S066[i / 6][i % 6];
    }
    memcpy(M.S0, S0, sizeof(M.S0));
    memcpy(M.C0, C0, sizeof(M.C0));
    {
        // K_ik(n) = sum_jl C0_ijkl n_j n_l over m = (n0^2, n1^2, n2^2, n1n2, n0n2, n0n1)
        const int qi[6] = {0, 1, 2, 1, 0, 0}, qk[6] = {0, 1, 2, 2, 2, 1};
        const int pj[6] = {0, 1, 2, 1, 0, 0}, pl[6] = {0, 1, 2, 2, 2, 1};
        for (int q = 0; q < 6; ++q)
            for (int p = 0; p < 6; ++p) {
                const int i = qi[q], k = qk[q], j = pj[p], l = pl[p];
                const double g = (j == l) ? C0[i][j][k][j] : C0[i][j][k][l] + C0[i][l][k][j];
                M.G36[6 * q + p] = g;
                M.G36f[6 * q + p] = (float)g;
            }
    }
    for (int i = 0; i < 9; ++i) {
        M.Udot[i] = Udot[i / 3][i % 3];
        M.Udot_a[i] = Ua[i / 3][i % 3];
        M.Scauchy[i] = Sc[i / 3][i % 3];
        M.DisGradAvg[i] = DGA[i / 3][i % 3];
    }
    for (int i = 0; i < 6; ++i) {
        M.udot_sym6[i] = D6[i];
        M.velgrad_flag[i] = vf[i];
        M.stress_flag[i] = sf[i];
    }
    M.iter_max = su->iter_max;
    M.creep = creep;
    M.update_flag = su->update_flag;
    M.hard_flag = su->hard_flag;
    M.n_phases = su->n_phases;
    M.d_vm = d_vm;
    M.e_vm = d_vm * dt;   // init.c:536
    int two_any = 0;
    for (int p = 0; p < 2; ++p) two_any |= (!pcs[p].gas && pcs[p].two_crss);
    M.two_crss_any = two_any;

    // device state
    voce_ctx *c = new voce_ctx();
    c->dims[0] = nx; c->dims[1] = ny; c->dims[2] = nz;
    c->V = V;
    c->precision = su->precision;
    c->pred_iters = su->iter_max < 8 ? su->iter_max : 8;
    c->update_flag = su->update_flag;
    c->creep = creep;
    c->iter_max = su->iter_max;
    c->two_crss_any = two_any;
    c->grain.assign(su->grain, su->grain + V);
    c->phase_id.assign(su->phase_id, su->phase_id + V);
    int rc = 0;
#define ALLOC(ptr, bytes)                                             \
    do {                                                              \
        void *p_ = nullptr;                                           \
        rc = voce_dev_alloc(c, &p_, (bytes));                              \
        if (rc) { free(ms0); voce_ctx_destroy(c); return rc; }        \
        ptr = (decltype(ptr))p_;                                      \
    } while (0)
    c->log_cap = su->iter_max > 0 ? su->iter_max : 1;
    if (opt && opt->stream) {
        c->stream = opt->stream;
        c->own_stream = false;
    } else {
        if (hipStreamCreate(&c->stream) != hipSuccess) { free(ms0); voce_ctx_destroy(c); return VOCE_ERR_DEVICE; }
        c->own_stream = true;
    }
    if (voce_kernels_prepare() != 0) { free(ms0); voce_ctx_destroy(c); return VOCE_ERR_DEVICE; }
    RveDev &d = c->d;
    memset(&d, 0, sizeof(d));
    d.nx = nx; d.ny = ny; d.nz = nz; d.nzh = nz / 2 + 1;
    d.V = V;
    d.fft_f32 = (su->precision == VOCE_MIXED);
    voce_spec_layout(&d);
    // one crystalline phase, numbered 1: its constants are uniform across a wave
    d.uniform_phase = (!pcs[0].gas) && (su->n_phases == 1 || pcs[1].gas);
    d.two_crss = two_any;
    // default FFT launch choices (a tuned set replaces them)
    d.plv = voce_default_plv(&d);   // line-parallel planes when the grid allows them
    d.xv = 0;
    d.tk = voce_xgreen_tile(nx, 0);
    d.pl_r8 = d.x_r8 = 1;
    const size_t cb = d.fft_f32 ? 8 : 16;
    const size_t db = sizeof(double);
    d.nblk_vox = (int)((V + voce_vox_block() - 1) / voce_vox_block());
    const bool ws = !(opt && opt->no_workspace);
    const bool own_ms = !(opt && opt->ms_dev);
    // one device arena: per-voxel state, then the small tables, then (unless
    // a server assigns them per request) the per-increment workspace
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    const size_t o_E6 = take(6 * V * db), o_S6 = take(6 * V * db), o_EP5 = take(5 * V * db), o_ED5 = take(5 * V * db);
    const size_t o_CR = take(12 * V * db), o_CRB = two_any ? take(12 * V * db) : 0;
    const size_t o_R = take(9 * V * db), o_WG = take(3 * V * db), o_WF = take(3 * V * db), o_GAC = take(V * db);
    const size_t o_GT = take(12 * V * db), o_ORI = take(V * sizeof(uint32_t)), o_PH = take(V);
    const size_t o_SL = take(otab.size() * db), o_PC = take(2 * sizeof(PhaseConst));
    const size_t o_TX = take(nx * cb), o_TY = take(ny * cb), o_TZ = take(nz * cb), o_S6O = take(sig6o.size() * db);
    const size_t o_TAB = take(sizeof(RveDev));
    const size_t o_MS = own_ms ? take(sizeof(MacroState)) : 0, o_LOG = own_ms ? take(3 * (size_t)c->log_cap * db) : 0;
    const size_t o_P6 = ws ? take(6 * V * db) : 0, o_KC = ws ? take(VOCE_KCV_COMP * voce_kcv_voxels(V) * db) : 0;
    const size_t o_SP = ws ? take(6 * d.nspec * cb) : 0;
    const size_t o_US = ws ? take(3 * d.nspec * cb) : 0, o_PT = ws ? take((size_t)d.nblk_vox * 16 * db) : 0;
    char *arena = nullptr;
    ALLOC(arena, off);
    c->arena_bytes = off;
    d.E6 = (double *)(arena + o_E6);
    d.S6 = (double *)(arena + o_S6);
    d.EP5 = (double *)(arena + o_EP5);
    d.ED5 = (double *)(arena + o_ED5);
    d.CRSS = (double *)(arena + o_CR);
    d.CRSSB = two_any ? (double *)(arena + o_CRB) : d.CRSS;
    d.R = (double *)(arena + o_R);
    d.WG = (double *)(arena + o_WG);
    d.WF = (double *)(arena + o_WF);
    d.GAC = (double *)(arena + o_GAC);
    d.GTOT = (double *)(arena + o_GT);
    d.ORI = (uint32_t *)(arena + o_ORI);
    d.PH = (uint8_t *)(arena + o_PH);
    d.SLOC = (double *)(arena + o_SL);
    d.phase = (PhaseConst *)(arena + o_PC);
    d.tw_x = arena + o_TX;
    d.tw_y = arena + o_TY;
    d.tw_z = arena + o_TZ;
    double *sig6o_d = (double *)(arena + o_S6O);
    c->dtab = (RveDev *)(arena + o_TAB);
    if (own_ms) {
        d.ms = (MacroState *)(arena + o_MS);
        c->logs_dev = (double *)(arena + o_LOG);
    } else {
        d.ms = opt->ms_dev;
        c->logs_dev = opt->logs_dev;
    }
    ms0->logs = c->logs_dev;
    ms0->log_cap = c->log_cap;
    if (ws) {
        d.P6 = (double *)(arena + o_P6);
        d.KCV = (double *)(arena + o_KC);
        d.spec = arena + o_SP;
        d.usum = arena + o_US;
        d.part = (double *)(arena + o_PT);
    }
    if (opt) c->on_tab = opt->on_tab;
    c->log_e.reserve(c->log_cap);
    c->log_s.reserve(c->log_cap);
    c->log_v.reserve(c->log_cap);
    if (hipGetDevice(&c->device) != hipSuccess ||
        hipHostMalloc((void **)&c->h_ms, sizeof(MacroState), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&c->h_ctl, 8 * sizeof(int), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&c->h_tab, sizeof(RveDev), hipHostMallocDefault) != hipSuccess) {
        free(ms0);
        voce_ctx_destroy(c);
        return VOCE_ERR_NOMEM;
    }
#undef ALLOC
#define H2D(dst, src, bytes) CK(hipMemcpyAsync((dst), (src), (bytes), hipMemcpyHostToDevice, c->stream))
#define ZERO(dst, bytes) CK(hipMemsetAsync((dst), 0, (bytes), c->stream))
    int e = 0;
    do {
        // per voxel only the orientation index and phase travel; the initial
        // TranMat, CRSS (init.c:789-792) and stress (InitFields) are expanded
        // on the device from the per-orientation and per-phase tables
        H2D(d.ORI, ori.data(), V * sizeof(uint32_t));
        H2D(d.PH, phv.data(), V);
        H2D(d.SLOC, otab.data(), otab.size() * db);
        H2D(d.phase, pcs, 2 * sizeof(PhaseConst));
        H2D(d.ms, ms0, sizeof(MacroState));
        H2D(sig6o_d, sig6o.data(), sig6o.size() * db);
        if (d.usum) ZERO(d.usum, 3 * d.nspec * cb);
        if (d.fft_f32) {
            std::vector<float> tx, ty, tz;
            make_twiddles(nx, tx); make_twiddles(ny, ty); make_twiddles(nz, tz);
            H2D(d.tw_x, tx.data(), nx * cb); H2D(d.tw_y, ty.data(), ny * cb); H2D(d.tw_z, tz.data(), nz * cb);
            CK(hipStreamSynchronize(c->stream));
        } else {
            std::vector<double> tx, ty, tz;
            make_twiddles(nx, tx); make_twiddles(ny, ty); make_twiddles(nz, tz);
            H2D(d.tw_x, tx.data(), nx * cb); H2D(d.tw_y, ty.data(), ny * cb); H2D(d.tw_z, tz.data(), nz * cb);
            CK(hipStreamSynchronize(c->stream));
        }
        e = voce_sync_tab(c);
        if (e) break;
        e = voce_launch_init_fields(c->dtab, V, sig6o_d, c->stream);
        if (e) break;
        CK(hipStreamSynchronize(c->stream));
    } while (0);
#undef H2D
#undef ZERO
    free(ms0);
    if (e) { voce_ctx_destroy(c); return VOCE_ERR_DEVICE; }
    *out = c;
    return VOCE_OK;
}

static int ensure_diag(voce_ctx *c)
{
    RveDev &d = c->d;
    if (d.GAMDOT) return 0;
    const long V = c->V;
    void *p;
    const size_t db = sizeof(double);
    struct { double **dst; size_t n; } want[] = {{&d.GAMDOT, 12 * V * db}, {&d.TAUX, 12 * V * db}, {&d.SVM, V * db},
                                                {&d.EVM, V * db}, {&d.WORKP, V * db}, {&d.RHOR, V * db},
                                                {&d.RHOC, V * db}, {&d.FLU, 6 * V * db}};
    for (auto &w : want) {
        if (voce_dev_alloc(c, &p, w.n)) return VOCE_ERR_NOMEM;
        if (hipMemsetAsync(p, 0, w.n, c->stream) != hipSuccess) return VOCE_ERR_DEVICE;
        *w.dst = (double *)p;
    }
    return voce_sync_tab(c);
}

extern "C" int voce_set_diagnostics(voce_ctx *c, int on)
{
    if (!c) return VOCE_ERR_ARG;
    if (on) {
        int rc = ensure_diag(c);
        if (rc) return rc;
    }
    c->diag = on ? 1 : 0;
    int v = c->diag;
    CK(hipMemcpyAsync((char *)c->d.ms + offsetof(MacroState, diag), &v, sizeof(int), hipMemcpyHostToDevice, c->stream));
    CK(hipStreamSynchronize(c->stream));
    return VOCE_OK;
}

extern "C" int voce_increment(voce_ctx *c, voce_incr_report *rep)
{
    if (!c) return VOCE_ERR_ARG;
    if (!c->d.spec) return VOCE_ERR_ARG;   // a server's resident RVE: increments go through the server
    return voce_run_increment(c, rep, c->prof);
}

// One increment over the descriptor the context's device table holds (its
// own workspace, or one a caller has put there).  timing: sample kernel
// times every timing-th iteration (0: none; no events are created then, and
// with the log vectors reserved at creation nothing is allocated).
int voce_run_increment(voce_ctx *c, voce_incr_report *rep, int timing)
{
    RveDev &d = c->d;
    int e = voce_launch_begin_increment(&c->L, c->stream);
    if (e) return VOCE_ERR_DEVICE;
    const size_t ctl_off = offsetof(MacroState, istep);
    // enqueue iterations in chunks; kernels past convergence return at once
    int launched = 0;
    const int imax = c->iter_max;
    c->h_ctl[0] = 0;
    c->h_ctl[3] = 0;
    c->h_ctl[4] = 0;
    if (imax > 0) {
        CK(hipMemcpyAsync(c->h_ctl, (char *)d.ms + ctl_off, 5 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        CK(hipStreamSynchronize(c->stream));
    }
    // first chunk: the previous increment's iteration count; past it (or
    // with no history) chunks of 2, 4, 8 and then 16, so a long increment
    // costs few host round trips and at most 15 launches past convergence
    int chunk = c->pred_iters > 0 ? c->pred_iters : 1;
    int next_chunk = 2;
    int done_before = 0;
    while (launched < imax && c->h_ctl[3]) {
        int n = chunk;
        if (launched + n > imax) n = imax - launched;
        if (timing && (int)c->ev.size() < 8 * n) {
            const size_t old = c->ev.size();
            c->ev.resize(8 * n);
            for (size_t i = old; i < c->ev.size(); ++i) CK(hipEventCreate(&c->ev[i]));
        }
        // timed iterations: every prof-th FFT iteration, the residue moving
        // by one per increment (early iterations of an increment run more
        // Newton steps; over prof increments every iteration index is sampled)
        auto timed = [&](int it) { return timing > 0 && (it + c->n_steps_done) % timing == 0; };
        for (int i = 0; i < n; ++i) {
            e = voce_launch_iteration(&c->L, c->stream, timed(launched + i) ? (void *)&c->ev[8 * i] : nullptr);
            if (e) return VOCE_ERR_DEVICE;
        }
        launched += n;
        // the chunk's last Newton step is finalised by a launch of its own
        if (voce_launch_finalize(&c->L, c->stream)) return VOCE_ERR_DEVICE;
        CK(hipMemcpyAsync(c->h_ctl, (char *)d.ms + ctl_off, 5 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        CK(hipStreamSynchronize(c->stream));
        if (c->h_ctl[4]) return c->h_ctl[4];
        if (timing) {
            // only the iterations that ran (the rest returned at the flag check)
            const int ran = c->h_ctl[1] - done_before;
            for (int i = 0; i < ran && i < n; ++i) {
                if (!timed(launched - n + i)) continue;
                for (int k = 0; k < VOCE_PROF_FIN; ++k) {   // start/end pair of kernel k
                    float ms = 0.f;
                    if (hipEventElapsedTime(&ms, c->ev[8 * i + 2 * k], c->ev[8 * i + 2 * k + 1]) == hipSuccess) {
                        c->prof_ms[k] += ms;
                        c->prof_n[k] += 1;
                    }
                }
            }
            done_before = c->h_ctl[1];
        }
        if (!c->h_ctl[3]) break;
        chunk = next_chunk;
        next_chunk = next_chunk < 16 ? 2 * next_chunk : 16;
    }
    const int istep = c->h_ctl[0];
    const int need_skew = (c->update_flag == 1 && istep > 1) || c->creep || c->diag;
    e = voce_launch_closeout(&c->L, c->stream, need_skew);
    if (e) return VOCE_ERR_DEVICE;
    CK(hipMemcpyAsync(c->h_ms, d.ms, sizeof(MacroState), hipMemcpyDeviceToHost, c->stream));
    CK(hipStreamSynchronize(c->stream));
    const MacroState &M = *c->h_ms;
    if (M.error) return M.error;
    c->pred_iters = M.iter > 0 ? M.iter : 1;
    c->n_steps_done = M.istep;
    if (rep) {
        const int n = M.iter < c->log_cap ? M.iter : c->log_cap;
        int lrc = voce_read_logs(c, n);
        if (lrc) return lrc;
        voce_fill_report(c, M, rep);
    }
    return VOCE_OK;
}

int voce_read_logs(voce_ctx *c, int n)
{
    c->log_e.resize(n);
    c->log_s.resize(n);
    c->log_v.resize(n);
    if (n <= 0) return 0;
    const size_t b = (size_t)n * sizeof(double);
    CK(hipMemcpyAsync(c->log_e.data(), c->logs_dev, b, hipMemcpyDeviceToHost, c->stream));
    CK(hipMemcpyAsync(c->log_s.data(), c->logs_dev + c->log_cap, b, hipMemcpyDeviceToHost, c->stream));
    CK(hipMemcpyAsync(c->log_v.data(), c->logs_dev + 2 * c->log_cap, b, hipMemcpyDeviceToHost, c->stream));
    CK(hipStreamSynchronize(c->stream));
    return 0;
}

void voce_fill_report(voce_ctx *c, const MacroState &M, voce_incr_report *rep)
{
    memset(rep, 0, sizeof(*rep));
    rep->istep = M.istep;
    rep->iter = M.iter;
    rep->iter_total = M.iter_total;
    rep->e_vm = M.e_vm; rep->evmp = M.evmp; rep->d_vm = M.d_vm; rep->dvmp = M.dvmp;
    rep->s_vm = M.s_vm; rep->s_vm1 = M.s_vm1;
    memcpy(rep->sig_avg, M.SigAvg, 72);
    memcpy(rep->sig_avg1, M.SigAvg1, 72);
    memcpy(rep->disgrad_avg, M.DisGradAvg_actual, 72);
    memcpy(rep->velgrad_avg, M.VelGradAvg, 72);
    memcpy(rep->eps_avg, M.EpsAvg, 72);
    memcpy(rep->edot_avg, M.EdotAvg, 72);
    rep->err_e = c->log_e.data();
    rep->err_s = c->log_s.data();
    rep->svm_it = c->log_v.data();
}

static void b6_to_tensor6(const double *a, double out[6])
{
    // B basis -> Voigt-ordered tensor components 11,22,33,23,13,12
    double T[3][3];
    v_to_t(a, T, 6);
    out[0] = T[0][0]; out[1] = T[1][1]; out[2] = T[2][2];
    out[3] = T[1][2]; out[4] = T[0][2]; out[5] = T[0][1];
}

extern "C" int voce_get_field(voce_ctx *c, int field, double *host, size_t bytes)
{
    if (!c || !host) return VOCE_ERR_ARG;
    const int nc = voce_field_ncomp(field);
    if (nc < 0) return VOCE_ERR_ARG;
    const long V = c->V;
    if (bytes < (size_t)nc * V * sizeof(double)) return VOCE_ERR_ARG;
    RveDev &d = c->d;
    auto d2h = [&](const void *src, void *dst, size_t n) -> int {
        if (hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, c->stream) != hipSuccess) return VOCE_ERR_DEVICE;
        return hipStreamSynchronize(c->stream) == hipSuccess ? 0 : VOCE_ERR_DEVICE;
    };
    const size_t db = sizeof(double);
    std::vector<double> a, b;
    int rc = 0;
    switch (field) {
    case VOCE_FIELD_SIG: case VOCE_FIELD_EPS: case VOCE_FIELD_EDOT: case VOCE_FIELD_ELS: {
        const int ncs = (field == VOCE_FIELD_SIG) ? 6 : 5;
        const double *src = field == VOCE_FIELD_SIG ? d.S6 : (field == VOCE_FIELD_EDOT ? d.ED5 : d.EP5);
        a.resize((size_t)ncs * V);
        if ((rc = d2h(src, a.data(), ncs * V * db))) return rc;
        if (field == VOCE_FIELD_ELS) {
            b.resize(6 * V);
            if ((rc = d2h(d.E6, b.data(), 6 * V * db))) return rc;
        }
        for (long v = 0; v < V; ++v) {
            double x[6] = {0, 0, 0, 0, 0, 0}, t[6];
            for (int i = 0; i < ncs; ++i) x[i] = a[i * V + v];
            b6_to_tensor6(x, t);
            if (field == VOCE_FIELD_ELS) {
                double y[6], ts[6];
                for (int i = 0; i < 6; ++i) y[i] = b[i * V + v];
                b6_to_tensor6(y, ts);
                for (int i = 0; i < 6; ++i) t[i] = ts[i] - t[i];
            }
            for (int i = 0; i < 6; ++i) host[i * V + v] = t[i];
        }
        return 0;
    }
    case VOCE_FIELD_DISGRAD: {
        a.resize(6 * V);
        b.resize(3 * V);
        if ((rc = d2h(d.E6, a.data(), 6 * V * db))) return rc;
        if ((rc = d2h(d.WG, b.data(), 3 * V * db))) return rc;
        for (long v = 0; v < V; ++v) {
            double x[6], T[3][3];
            for (int i = 0; i < 6; ++i) x[i] = a[i * V + v];
            v_to_t(x, T, 6);
            const double w0 = b[v], w1 = b[V + v], w2 = b[2 * V + v];
            T[2][1] += w0; T[1][2] -= w0;
            T[0][2] += w1; T[2][0] -= w1;
            T[1][0] += w2; T[0][1] -= w2;
            for (int i = 0; i < 9; ++i) host[i * V + v] = T[i / 3][i % 3];
        }
        return 0;
    }
    case VOCE_FIELD_VELGRAD: {
        b.resize(3 * V);
        if ((rc = d2h(d.WF, b.data(), 3 * V * db))) return rc;
        const double dt = c->h_ms->dt > 0 ? c->h_ms->dt : 1.0;
        for (long v = 0; v < V; ++v) {
            const double w0 = b[v] / dt, w1 = b[V + v] / dt, w2 = b[2 * V + v] / dt;
            const double T[9] = {0.0, -w2, w1, w2, 0.0, -w0, -w1, w0, 0.0};
            for (int i = 0; i < 9; ++i) host[i * V + v] = T[i];
        }
        return 0;
    }
    case VOCE_FIELD_WORK_E: {
        // WriteWorkEMPI (io.c:231-278): els in Voigt order (11,22,33,23,13,12)
        // goes through chg_basis(opt 1) as if it were a B-basis vector
        a.resize(6 * V);
        b.resize(6 * V);
        if ((rc = voce_get_field(c, VOCE_FIELD_SIG, a.data(), 6 * V * db))) return rc;
        if ((rc = voce_get_field(c, VOCE_FIELD_ELS, b.data(), 6 * V * db))) return rc;
        for (long v = 0; v < V; ++v) {
            double e6[6], T[3][3], S[3][3];
            for (int i = 0; i < 6; ++i) e6[i] = b[i * V + v];
            v_to_t(e6, T, 6);
            const double s6[6] = {a[v], a[V + v], a[2 * V + v], a[3 * V + v], a[4 * V + v], a[5 * V + v]};
            S[0][0] = s6[0]; S[1][1] = s6[1]; S[2][2] = s6[2];
            S[1][2] = S[2][1] = s6[3];
            S[0][2] = S[2][0] = s6[4];
            S[0][1] = S[1][0] = s6[5];
            double w = 0.0;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) w += 0.5 * S[i][j] * T[i][j];
            host[v] = w;
        }
        return 0;
    }
    case VOCE_FIELD_NEW_POSITION: {
        // evolution.c:532-560 transforms fluctuation[.][i][2] forward and back
        // (its Green step writes out of bounds and is lost): the output is
        // -fluctuation[.][i][2]
        if (!d.FLU) {
            memset(host, 0, 3 * (size_t)V * db);
            return 0;
        }
        a.resize(6 * V);
        if ((rc = d2h(d.FLU, a.data(), 6 * V * db))) return rc;
        for (long v = 0; v < V; ++v) {
            const double x[6] = {0.0, a[v], a[V + v], a[2 * V + v], 0.0, a[3 * V + v]};
            double T[3][3];
            v_to_t(x, T, 6);
            const double w0 = a[4 * V + v], w1 = a[5 * V + v];
            host[v] = -(T[0][2] + w1);
            host[V + v] = -(T[1][2] - w0);
            host[2 * V + v] = -T[2][2];
        }
        return 0;
    }
    case VOCE_FIELD_SCHMID: {
        // update_schmid (kinematics.c:426-459): Schm_xt as a tensor, rotated
        // by TranMat, back to 5 components
        a.resize(9 * V);
        if ((rc = d2h(d.R, a.data(), 9 * V * db))) return rc;
        std::vector<uint8_t> ph(V);
        if ((rc = d2h(d.PH, ph.data(), V))) return rc;
        PhaseConst pcs[2];
        if ((rc = d2h(d.phase, pcs, sizeof(pcs)))) return rc;
        for (long v = 0; v < V; ++v) {
            const PhaseConst &pc = pcs[ph[v] >= 1 && ph[v] <= 2 ? ph[v] - 1 : 0];
            double R[3][3];
            for (int i = 0; i < 9; ++i) R[i / 3][i % 3] = a[i * V + v];
            for (int sy = 0; sy < 12; ++sy) {
                double out[5] = {0, 0, 0, 0, 0};
                if (sy < pc.nsys && !pc.gas) {
                    double P[3][3], M[3][3], Q[3][3];
                    v_to_t(pc.px[sy], P, 5);
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j) {
                            M[i][j] = 0.0;
                            for (int k = 0; k < 3; ++k) M[i][j] += R[i][k] * P[k][j];
                        }
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j) {
                            Q[i][j] = 0.0;
                            for (int k = 0; k < 3; ++k) Q[i][j] += M[i][k] * R[j][k];
                        }
                    t_to_v(Q, out, 5);
                }
                for (int k = 0; k < 5; ++k) host[(sy * 5 + k) * V + v] = out[k];
            }
        }
        return 0;
    }
    case VOCE_FIELD_CRSS: return d2h(d.CRSS, host, 12 * V * db);
    case VOCE_FIELD_CRSS_BWD: {
        if (c->two_crss_any) return d2h(d.CRSSB, host, 12 * V * db);
        // tau0f == tau0b: forward and backward CRSS are the same numbers
        return d2h(d.CRSS, host, 12 * V * db);
    }
    case VOCE_FIELD_GAMTOT: return d2h(d.GTOT, host, 12 * V * db);
    case VOCE_FIELD_TRANMAT: return d2h(d.R, host, 9 * V * db);
    case VOCE_FIELD_GAMACUM: return d2h(d.GAC, host, V * db);
    case VOCE_FIELD_GAMDOT: case VOCE_FIELD_TAUX: case VOCE_FIELD_SVM: case VOCE_FIELD_EVM:
    case VOCE_FIELD_WORK_P: case VOCE_FIELD_RHO_RSS: case VOCE_FIELD_RHO_CRSS: {
        const double *src = field == VOCE_FIELD_GAMDOT ? d.GAMDOT : field == VOCE_FIELD_TAUX ? d.TAUX
                          : field == VOCE_FIELD_SVM ? d.SVM : field == VOCE_FIELD_EVM ? d.EVM
                          : field == VOCE_FIELD_WORK_P ? d.WORKP : field == VOCE_FIELD_RHO_RSS ? d.RHOR : d.RHOC;
        if (!src) {
            memset(host, 0, (size_t)nc * V * db);
            return 0;
        }
        return d2h(src, host, nc * V * db);
    }
    }
    return VOCE_ERR_ARG;
}

// ---------------------------------------------------------------------------
// .iout writers (io.c:193-1002), single rank: the whole grid at offset 0
// ---------------------------------------------------------------------------
static int write_raw(const char *dir, const char *name, const void *p, size_t n)
{
    char path[1200];
    snprintf(path, sizeof(path), "%s/%s", dir, name);
    FILE *f = fopen(path, "wb");
    if (!f) return VOCE_ERR_FILE;
    const size_t w = fwrite(p, 1, n, f);
    fclose(f);
    return w == n ? 0 : VOCE_ERR_FILE;
}

extern "C" int voce_write_fields(voce_ctx *c, const char *dir, int step, int tex_flag, int fin)
{
    if (!c || !dir) return VOCE_ERR_ARG;
    const long V = c->V;
    std::vector<double> buf(12 * V);
    char name[256];
    int rc;
    // eps, sig, els, edot (io.c:193-397)
    struct { const char *pre; const char *pre_fin; int field; int wide; } six[4] = {
        {"eps", "eps_final", VOCE_FIELD_EPS, 0}, {"sig", "sig_final", VOCE_FIELD_SIG, 0},
        {"els", "els_final", VOCE_FIELD_ELS, 1}, {"edot", "edot_final", VOCE_FIELD_EDOT, 0}};
    for (auto &s : six) {
        if ((rc = voce_get_field(c, s.field, buf.data(), buf.size() * sizeof(double)))) return rc;
        for (int i = 0; i < 6; ++i) {
            snprintf(name, sizeof(name), s.wide ? "%s_%01d_S%06d.iout" : "%s_%01d_S%04d.iout", fin ? s.pre_fin : s.pre,
                     i + 1, step);
            if ((rc = write_raw(dir, name, buf.data() + i * V, V * sizeof(double)))) return rc;
        }
    }
    if (!tex_flag) return 0;
    // 104-byte {int grain; 4 bytes padding; 12 doubles} records (io.c:651-936)
    struct { const char *pre; const char *pre_fin; int field; } rec[3] = {
        {"slip_sys", "slip_final", VOCE_FIELD_GAMTOT}, {"taux_sys", "taux_final", VOCE_FIELD_TAUX},
        {"crss_sys", "crss_final", VOCE_FIELD_CRSS}};
    std::vector<unsigned char> r(104 * V);
    for (auto &s : rec) {
        if ((rc = voce_get_field(c, s.field, buf.data(), buf.size() * sizeof(double)))) return rc;
        memset(r.data(), 0, r.size());
        for (long v = 0; v < V; ++v) {
            memcpy(&r[104 * v], &c->grain[v], 4);
            for (int k = 0; k < 12; ++k) memcpy(&r[104 * v + 8 + 8 * k], &buf[k * V + v], 8);
        }
        snprintf(name, sizeof(name), "%s_S%04d.iout", fin ? s.pre_fin : s.pre, step);
        if ((rc = write_raw(dir, name, r.data(), r.size()))) return rc;
    }
    // tex: 48-byte {3 double angles; int i,j,k; int grain, phase, 0} (io.c:939-1002)
    if ((rc = voce_get_field(c, VOCE_FIELD_TRANMAT, buf.data(), buf.size() * sizeof(double)))) return rc;
    std::vector<unsigned char> t(48 * V);
    const int ny = c->dims[1], nz = c->dims[2];
    for (long v = 0; v < V; ++v) {
        double sa2xt[3][3], ang[3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) sa2xt[i][j] = buf[(j * 3 + i) * V + v];
        mat_to_euler(sa2xt, &ang[0], &ang[1], &ang[2]);
        int ci[6] = {(int)(v / ((long)ny * nz)) + 1, (int)((v / nz) % ny) + 1, (int)(v % nz) + 1, c->grain[v],
                     c->phase_id[v], 0};
        memcpy(&t[48 * v], ang, 24);
        memcpy(&t[48 * v + 24], ci, 24);
    }
    snprintf(name, sizeof(name), "%s_S%04d.iout", fin ? "tex_final" : "tex", step);
    if ((rc = write_raw(dir, name, t.data(), t.size()))) return rc;
    struct { const char *pre; const char *pre_fin; int field; int wide; } one[5] = {
        {"rho_rss", "rho_rss_final", VOCE_FIELD_RHO_RSS, 0}, {"rho_crss", "rho_crss_final", VOCE_FIELD_RHO_CRSS, 0},
        {"work_p", "work_p_final", VOCE_FIELD_WORK_P, 1}, {"SVM", "SVM_final", VOCE_FIELD_SVM, 0},
        {"EVM", "EVM_final", VOCE_FIELD_EVM, 0}};
    for (auto &s : one) {
        if ((rc = voce_get_field(c, s.field, buf.data(), buf.size() * sizeof(double)))) return rc;
        snprintf(name, sizeof(name), s.wide ? "%s_S%06d.iout" : "%s_S%04d.iout", fin ? s.pre_fin : s.pre, step);
        if ((rc = write_raw(dir, name, buf.data(), V * sizeof(double)))) return rc;
    }
    // new_position: 24-byte {3 double} records (io.c:805-849)
    if ((rc = voce_get_field(c, VOCE_FIELD_NEW_POSITION, buf.data(), buf.size() * sizeof(double)))) return rc;
    {
        std::vector<double> rec3(3 * V);
        for (long v = 0; v < V; ++v)
            for (int k = 0; k < 3; ++k) rec3[3 * v + k] = buf[k * V + v];
        snprintf(name, sizeof(name), "%s_S%04d.iout", fin ? "new_position_final" : "new_position", step);
        if ((rc = write_raw(dir, name, rec3.data(), rec3.size() * sizeof(double)))) return rc;
    }
    // work_e is dumped only at the end (evolution.c:603; its per-step call is commented out)
    if (fin) {
        if ((rc = voce_get_field(c, VOCE_FIELD_WORK_E, buf.data(), buf.size() * sizeof(double)))) return rc;
        snprintf(name, sizeof(name), "work_e_final_S%06d.iout", step);
        if ((rc = write_raw(dir, name, buf.data(), V * sizeof(double)))) return rc;
    }
    return 0;
}

extern "C" int voce_prof_enable(voce_ctx *c, int on)
{
    if (!c) return VOCE_ERR_ARG;
    c->prof = on > 0 ? on : 0;
    for (int k = 0; k < VOCE_PROF_NK; ++k) {
        c->prof_ms[k] = 0.0;
        c->prof_n[k] = 0;
    }
    return VOCE_OK;
}

extern "C" int voce_prof_read(voce_ctx *c, double *ms_total, long *count)
{
    if (!c || !ms_total || !count) return VOCE_ERR_ARG;
    for (int k = 0; k < VOCE_PROF_NK; ++k) {
        ms_total[k] = c->prof_ms[k];
        count[k] = c->prof_n[k];
    }
    return VOCE_OK;
}

extern "C" int voce_ctx_info(const voce_ctx *c, long *out)
{
    if (!c || !out) return VOCE_ERR_ARG;
    out[0] = c->V;
    out[1] = c->dims[0];
    out[2] = c->dims[1];
    out[3] = c->dims[2];
    out[4] = c->precision;
    out[5] = c->two_crss_any;
    return VOCE_OK;
}

extern "C" int voce_set_velgrad(voce_ctx *c, const double L[9])
{
    if (!c || !L) return VOCE_ERR_ARG;
    struct { double vg[9]; } p;
    memcpy(p.vg, L, sizeof(p.vg));
    const int one = 1;
    CK(hipMemcpyAsync((char *)c->d.ms + offsetof(MacroState, vg_next), p.vg, sizeof(p.vg), hipMemcpyHostToDevice,
                      c->stream));
    CK(hipMemcpyAsync((char *)c->d.ms + offsetof(MacroState, vg_pending), &one, sizeof(int), hipMemcpyHostToDevice,
                      c->stream));
    CK(hipStreamSynchronize(c->stream));
    return VOCE_OK;
}
