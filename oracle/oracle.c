/*
 * oracle.c -- CPU restatement of the mesoOSU/Voce EVP-FFT solver (FP64).
 *
 * TEST INFRASTRUCTURE: the checker for the HIP path, never part of it.
 * See oracle.h for what may use this file.
 *
 * Layout and arithmetic mirror the reference: AoS row-major 3x3 tensors per
 * voxel, voxel index (x*Ny + y)*Nz + z (evp.h:32-33), the Lebensohn-Tome
 * B basis, 9 complex FFTs per direction and the full 81-term Green
 * contraction.  Expression order follows the reference so that, were the
 * reference buildable, the two would round alike.  Citations are
 * reference file:line.
 */
#include "oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NSYSMX 12
#define NPHMAX 2
#define PI_REF 3.14159265359      /* evp.h:57 (truncated, as the reference) */
#define TWO_PI_REF 6.28318530718  /* evp.h:58 */
#define LU_TINY 1.0E-20           /* kinematics.c:3 */

typedef double t2[3][3];
typedef double v6[6];
typedef double m66[6][6];

/* ------------------------------------------------------------------ */
/* B basis, kinematics.c:14-174                                         */
/* ------------------------------------------------------------------ */
static double Bb[6][3][3];
static int Bb_ready = 0;

static void build_B(void)
{
    /* kinematics.c:28-54 */
    const double RSQ2 = 0.70710678118654744;
    const double RSQ3 = 0.57735026918962584;
    const double RSQ6 = 0.40824829046386304;
    memset(Bb, 0, sizeof(Bb));
    Bb[1][0][0] = -1. * RSQ6;
    Bb[1][1][1] = -1. * RSQ6;
    Bb[1][2][2] = 2. * RSQ6;
    Bb[0][0][0] = -1. * RSQ2;
    Bb[0][1][1] = RSQ2;
    Bb[2][1][2] = RSQ2;
    Bb[2][2][1] = RSQ2;
    Bb[3][0][2] = RSQ2;
    Bb[3][2][0] = RSQ2;
    Bb[4][0][1] = RSQ2;
    Bb[4][1][0] = RSQ2;
    Bb[5][0][0] = RSQ3;
    Bb[5][1][1] = RSQ3;
    Bb[5][2][2] = RSQ3;
    Bb_ready = 1;
}

/* chg_basis, kinematics.c:14-93 (ncomp 6) and chg_basis5, 95-174 (ncomp 5) */
static void chg_basis_n(double *V, t2 T2, m66 C2, double T4[3][3][3][3], int opt, int nc)
{
    int i, j, mi, mj, mk, ml;
    if (!Bb_ready) build_B();
    if (opt == 1) {
        for (mi = 0; mi < 3; mi++)
            for (mj = 0; mj < 3; mj++) {
                T2[mi][mj] = 0.0;
                for (i = 0; i < nc; i++) T2[mi][mj] += V[i] * Bb[i][mi][mj];
            }
    } else if (opt == 2) {
        for (i = 0; i < nc; i++) {
            V[i] = 0.0;
            for (mi = 0; mi < 3; mi++)
                for (mj = 0; mj < 3; mj++) V[i] += T2[mi][mj] * Bb[i][mi][mj];
        }
    } else if (opt == 3) {
        for (mi = 0; mi < 3; mi++)
            for (mj = 0; mj < 3; mj++)
                for (mk = 0; mk < 3; mk++)
                    for (ml = 0; ml < 3; ml++) {
                        T4[mi][mj][mk][ml] = 0.0;
                        for (i = 0; i < nc; i++)
                            for (j = 0; j < nc; j++)
                                T4[mi][mj][mk][ml] += C2[i][j] * Bb[i][mi][mj] * Bb[j][mk][ml];
                    }
    } else if (opt == 4) {
        for (i = 0; i < nc; i++)
            for (j = 0; j < nc; j++) {
                C2[i][j] = 0.0;
                for (mi = 0; mi < 3; mi++)
                    for (mj = 0; mj < 3; mj++)
                        for (mk = 0; mk < 3; mk++)
                            for (ml = 0; ml < 3; ml++)
                                C2[i][j] += T4[mi][mj][mk][ml] * Bb[i][mi][mj] * Bb[j][mk][ml];
            }
    }
}
#define chg_basis(V, T, C, T4, o) chg_basis_n((V), (T), (C), (T4), (o), 6)
#define chg_basis5(V, T, C, T4, o) chg_basis_n((V), (T), (C), (T4), (o), 5)

void orc_chg_basis(double V6[6], double T2[9], double C2[36], double T4[81], int opt)
{
    chg_basis_n(V6, (double(*)[3])T2, (double(*)[6])C2, (double(*)[3][3][3])T4, opt, 6);
}
void orc_chg_basis5(double V5[5], double T2[9], double C2[36], double T4[81], int opt)
{
    chg_basis_n(V5, (double(*)[3])T2, (double(*)[6])C2, (double(*)[3][3][3])T4, opt, 5);
}

/* VoigtToFull, kinematics.c:5-12 */
static void voigt_to_full(const double *v, t2 m)
{
    m[0][0] = v[0]; m[0][1] = v[5]; m[0][2] = v[4];
    m[1][0] = v[5]; m[1][1] = v[1]; m[1][2] = v[3];
    m[2][0] = v[4]; m[2][1] = v[3]; m[2][2] = v[2];
}

/* SymAntDecompose, kinematics.c:176-184 */
static void sym_ant(t2 t, t2 s, t2 a)
{
    int mi, mj;
    for (mi = 0; mi < 3; mi++)
        for (mj = 0; mj < 3; mj++) {
            s[mi][mj] = (t[mi][mj] + t[mj][mi]) / 2.;
            a[mi][mj] = (t[mi][mj] - t[mj][mi]) / 2.;
        }
}

/* EulerToTransMatrix, kinematics.c:187-240 (Bunge, sample -> crystal) */
void orc_euler_to_mat(double *p_t1, double *p_ph, double *p_t2, double a9[9], int opt)
{
    double (*a)[3] = (double(*)[3])a9;
    double ct1, st1, cph, sph, ct2, st2, t1 = *p_t1, ph = *p_ph, t2v = *p_t2;
    if (opt == 1) {
        ph = acos(a[2][2]);
        if (fabs(a[2][2]) > 0.9999) {
            t2v = 0.;
            t1 = atan2(a[0][1], a[0][0]);
        } else {
            sph = sin(ph);
            t1 = atan2(a[2][0] / sph, -1.0 * a[2][1] / sph);
            t2v = atan2(a[0][2] / sph, a[1][2] / sph);
        }
        ph *= 180. / PI_REF;
        t1 *= 180. / PI_REF;
        t2v *= 180. / PI_REF;
        *p_ph = ph;
        *p_t1 = t1;
        *p_t2 = t2v;
    } else if (opt == 2) {
        ct1 = cos(t1); st1 = sin(t1);
        cph = cos(ph); sph = sin(ph);
        ct2 = cos(t2v); st2 = sin(t2v);
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
}

/* Ten4thTransform, kinematics.c:243-280 */
void orc_ten4th_transform(double c1f[81], const double af[9], double c2f[81], int opt)
{
    double (*c1)[3][3][3] = (double(*)[3][3][3])c1f;
    double (*c2)[3][3][3] = (double(*)[3][3][3])c2f;
    const double (*a)[3] = (const double(*)[3])af;
    int i, j, k, l, m, n, p, q;
    for (i = 0; i < 3; i++)
        for (j = 0; j < 3; j++)
            for (k = 0; k < 3; k++)
                for (l = 0; l < 3; l++) {
                    if (opt == 1) {
                        c2[i][j][k][l] = 0.0;
                        for (m = 0; m < 3; m++)
                            for (n = 0; n < 3; n++)
                                for (p = 0; p < 3; p++)
                                    for (q = 0; q < 3; q++)
                                        c2[i][j][k][l] += a[i][m] * a[j][n] * a[k][p] * a[l][q] * c1[m][n][p][q];
                    } else {
                        c1[i][j][k][l] = 0.0;
                        for (m = 0; m < 3; m++)
                            for (n = 0; n < 3; n++)
                                for (p = 0; p < 3; p++)
                                    for (q = 0; q < 3; q++)
                                        c1[i][j][k][l] += a[m][i] * a[n][j] * a[p][k] * a[q][l] * c2[m][n][p][q];
                    }
                }
}

/* LU_dcmp / LU_bksb / LU_inv_66, kinematics.c:282-424 (Numerical Recipes
 * Crout LU with implicit row scaling; row max < 1e-5 is fatal (exit 35);
 * a pivot below 1e-5 in magnitude is replaced by 1e-20). */
static int lu_dcmp6(double a[6][6], int indx[6])
{
    int i, imax = 0, j, k;
    double big, dum, sum, temp, vv[6];
    const int n = 6;
    for (i = 1; i <= n; i++) {
        big = 0.0;
        for (j = 1; j <= n; j++)
            if ((temp = fabs(a[i - 1][j - 1])) > big) big = temp;
        if (fabs(big) < 1E-5) return 35;
        vv[i - 1] = 1.0 / big;
    }
    for (j = 1; j <= n; j++) {
        for (i = 1; i < j; i++) {
            sum = a[i - 1][j - 1];
            for (k = 1; k < i; k++) sum -= a[i - 1][k - 1] * a[k - 1][j - 1];
            a[i - 1][j - 1] = sum;
        }
        big = 0.0;
        imax = j;
        for (i = j; i <= n; i++) {
            sum = a[i - 1][j - 1];
            for (k = 1; k < j; k++) sum -= a[i - 1][k - 1] * a[k - 1][j - 1];
            a[i - 1][j - 1] = sum;
            if ((dum = vv[i - 1] * fabs(sum)) >= big) {
                big = dum;
                imax = i;
            }
        }
        if (j != imax) {
            for (k = 1; k <= n; k++) {
                dum = a[imax - 1][k - 1];
                a[imax - 1][k - 1] = a[j - 1][k - 1];
                a[j - 1][k - 1] = dum;
            }
            vv[imax - 1] = vv[j - 1];
        }
        indx[j - 1] = imax;
        if (fabs(a[j - 1][j - 1]) < 1E-5) a[j - 1][j - 1] = LU_TINY;
        if (j != n) {
            dum = 1.0 / (a[j - 1][j - 1]);
            for (i = j + 1; i <= n; i++) a[i - 1][j - 1] *= dum;
        }
    }
    return 0;
}

static void lu_bksb6(double a[6][6], const int indx[6], double b[6])
{
    int i, ii = 0, ip, j;
    double sum;
    const int n = 6;
    for (i = 1; i <= n; i++) {
        ip = indx[i - 1];
        sum = b[ip - 1];
        b[ip - 1] = b[i - 1];
        if (ii)
            for (j = ii; j <= i - 1; j++) sum -= a[i - 1][j - 1] * b[j - 1];
        else if (fabs(sum) > 1E-5)
            ii = i;
        b[i - 1] = sum;
    }
    for (i = n; i >= 1; i--) {
        sum = b[i - 1];
        for (j = i + 1; j <= n; j++) sum -= a[i - 1][j - 1] * b[j - 1];
        b[i - 1] = sum / a[i - 1][i - 1];
    }
}

static int lu_inv66(double c[6][6])
{
    int indx[6], i, j, rc;
    double col[6], y[6][6];
    rc = lu_dcmp6(c, indx);
    if (rc) return rc;
    for (j = 0; j < 6; j++) {
        for (i = 0; i < 6; i++) col[i] = 0.0;
        col[j] = 1.0;
        lu_bksb6(c, indx, col);
        for (i = 0; i < 6; i++) y[i][j] = col[i];
    }
    for (i = 0; i < 6; i++)
        for (j = 0; j < 6; j++) c[i][j] = y[i][j];
    return 0;
}

int orc_lu_inv66(double c[36]) { return lu_inv66((double(*)[6])c); }

/* M3Inverse, init.c:546-568 */
static void m3_inverse(double a[3][3], double ia[3][3])
{
    int i, j;
    double deta = 0.0;
    deta = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1])
         - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0])
         + a[0][2] * (a[1][0] * a[2][1] - a[2][0] * a[1][1]);
    ia[0][0] = a[1][1] * a[2][2] - a[1][2] * a[2][1];
    ia[0][1] = a[0][2] * a[2][1] - a[0][1] * a[2][2];
    ia[0][2] = a[0][1] * a[1][2] - a[0][2] * a[1][1];
    ia[1][0] = a[1][2] * a[2][0] - a[1][0] * a[2][2];
    ia[1][1] = a[0][0] * a[2][2] - a[0][2] * a[2][0];
    ia[1][2] = a[0][2] * a[1][0] - a[0][0] * a[1][2];
    ia[2][0] = a[1][0] * a[2][1] - a[1][1] * a[2][0];
    ia[2][1] = a[0][1] * a[2][0] - a[0][0] * a[2][1];
    ia[2][2] = a[0][0] * a[1][1] - a[0][1] * a[1][0];
    for (i = 0; i < 3; i++)
        for (j = 0; j < 3; j++) ia[i][j] /= deta;
}

/* Unit frequency vector, init.c:660-677: t = i*(1/N), t -= round(t),
 * t *= 2pi; n = t/|t|; n(0) = 0. */
void orc_kvector(int px, int py, int pz, int nx, int ny, int nz, double n[3])
{
    double tmp[3], g2;
    int m;
    tmp[0] = (double)(px) * (double)(1.0 / nx);
    tmp[1] = (double)(py) * (double)(1.0 / ny);
    tmp[2] = (double)(pz) * (double)(1.0 / nz);
    for (m = 0; m < 3; m++) {
        tmp[m] -= round(tmp[m]);
        tmp[m] *= TWO_PI_REF;
    }
    g2 = sqrt(tmp[0] * tmp[0] + tmp[1] * tmp[1] + tmp[2] * tmp[2]);
    if ((px + py + pz) == 0) {
        g2 = 1.0;
        tmp[0] = tmp[1] = tmp[2] = 0.0;
    }
    for (m = 0; m < 3; m++) n[m] = tmp[m] * (1. / g2);
}

/* CalcGAMMA body, init.c:595-610: Gamma_ijkl = -omega_ik n_j n_l with
 * omega = (C0_ijkl n_j n_l)^-1.  Caller handles k = 0 (Gamma = 0). */
void orc_green(const double C0f[81], const double n[3], double Gf[81])
{
    const double (*C0)[3][3][3] = (const double(*)[3][3][3])C0f;
    double (*G)[3][3][3] = (double(*)[3][3][3])Gf;
    double iomega[3][3], omega[3][3];
    int i, j, k, l, mi, mj, mip, mjp;
    for (i = 0; i < 3; i++)
        for (k = 0; k < 3; k++) {
            iomega[i][k] = 0.0;
            for (j = 0; j < 3; j++)
                for (l = 0; l < 3; l++) iomega[i][k] += C0[i][j][k][l] * n[j] * n[l];
        }
    m3_inverse(iomega, omega);
    for (mi = 0; mi < 3; mi++)
        for (mj = 0; mj < 3; mj++)
            for (mip = 0; mip < 3; mip++)
                for (mjp = 0; mjp < 3; mjp++)
                    G[mi][mj][mip][mjp] = -1.0 * omega[mi][mip] * n[mj] * n[mjp];
}

/* ------------------------------------------------------------------ */
/* Serial mixed-radix complex DFT (stands in for fftwnd_mpi with        */
/* FFTW_NORMAL_ORDER: forward e^{-i}, unnormalised both ways).          */
/* ------------------------------------------------------------------ */
typedef struct { double re, im; } cpx;

static void dft_rec(int n, int stride, const cpx *in, cpx *out, const cpx *tw, int N0, cpx *scratch)
{
    int p, m, r, k, s, e;
    if (n == 1) { out[0] = in[0]; return; }
    if (n % 4 == 0) p = 4;
    else if (n % 2 == 0) p = 2;
    else if (n % 3 == 0) p = 3;
    else if (n % 5 == 0) p = 5;
    else { for (p = 7; p * p <= n && n % p; p += 2) ; if (n % p) p = n; }
    m = n / p;
    for (r = 0; r < p; r++) dft_rec(m, stride * p, in + r * stride, out + r * m, tw, N0, scratch);
    /* combine: X[k + s m] = sum_r Y_r[k] W_n^{r (k + s m)} */
    for (k = 0; k < m; k++) {
        for (s = 0; s < p; s++) {
            double sr = 0.0, si = 0.0;
            for (r = 0; r < p; r++) {
                cpx y = out[r * m + k];
                cpx w;
                e = (int)(((long)r * (k + s * m)) % n);
                w = tw[(long)e * (N0 / n) % N0];
                sr += y.re * w.re - y.im * w.im;
                si += y.re * w.im + y.im * w.re;
            }
            scratch[s].re = sr;
            scratch[s].im = si;
        }
        for (s = 0; s < p; s++) out[k + s * m] = scratch[s];
    }
}

static void fft1d(int n, cpx *line, cpx *tmp, const cpx *tw, cpx *scr)
{
    dft_rec(n, 1, line, tmp, tw, n, scr);
    memcpy(line, tmp, sizeof(cpx) * n);
}

static cpx *twiddles(int n, int sign)
{
    cpx *tw = (cpx *)malloc(sizeof(cpx) * n);
    int j;
    for (j = 0; j < n; j++) {
        double ang = 2.0 * M_PI * (double)j / (double)n;
        tw[j].re = cos(ang);
        tw[j].im = sign * sin(ang);
    }
    return tw;
}

static int g_threads = 1;
void orc_set_threads(int n)
{
    g_threads = n > 0 ? n : 1;
#ifdef _OPENMP
    omp_set_num_threads(g_threads);
#endif
}

/* data: V complex values, index (x*ny + y)*nz + z */
static void fft3d_cpx(cpx *data, int nx, int ny, int nz, int sign)
{
    int dims[3] = {nx, ny, nz};
    int ax;
    for (ax = 2; ax >= 0; ax--) {
        int n = dims[ax];
        long stride = (ax == 2) ? 1 : (ax == 1 ? nz : (long)ny * nz);
        long nlines = ((long)nx * ny * nz) / n;
        cpx *tw = twiddles(n, sign);
        long L;
#pragma omp parallel
        {
            cpx *line = (cpx *)malloc(sizeof(cpx) * n);
            cpx *tmp = (cpx *)malloc(sizeof(cpx) * n);
            cpx *scr = (cpx *)malloc(sizeof(cpx) * n);
#pragma omp for schedule(static)
            for (L = 0; L < nlines; L++) {
                long base;
                int q;
                if (ax == 2) base = L * nz;
                else if (ax == 1) base = (L / nz) * (long)ny * nz + (L % nz);
                else base = L;
                for (q = 0; q < n; q++) line[q] = data[base + q * stride];
                fft1d(n, line, tmp, tw, scr);
                for (q = 0; q < n; q++) data[base + q * stride] = line[q];
            }
            free(line); free(tmp); free(scr);
        }
        free(tw);
    }
}

void orc_fft3d(double *re, double *im, int nx, int ny, int nz, int sign)
{
    long V = (long)nx * ny * nz, i;
    cpx *d = (cpx *)malloc(sizeof(cpx) * V);
    for (i = 0; i < V; i++) { d[i].re = re[i]; d[i].im = im[i]; }
    fft3d_cpx(d, nx, ny, nz, sign);
    for (i = 0; i < V; i++) { re[i] = d[i].re; im[i] = d[i].im; }
    free(d);
}

/* ------------------------------------------------------------------ */
/* Constitutive law                                                     */
/* ------------------------------------------------------------------ */

/* StrainRate_eval, constitutive.c:4-72 */
static void strain_rate_eval(const double stress[6], const double sc[NSYSMX][5],
                             const double tau_t[NSYSMX][2], const double *xkinaux, int nsys,
                             const double *nsr, const double *gam0,
                             double edot[6], double d_edot[6][6],
                             double *gamdot, double *taux)
{
    int i, j, k, isign;
    double rss[NSYSMX], rss1[NSYSMX], rss2[NSYSMX];
    for (i = 0; i < nsys; i++) {
        rss[i] = sc[i][0] * stress[0] + sc[i][1] * stress[1] + sc[i][2] * stress[2] +
                 sc[i][3] * stress[3] + sc[i][4] * stress[4];
        isign = ((rss[i] - xkinaux[i]) < 0) ? 1 : 0;
        rss[i] = (rss[i] - xkinaux[i]) / tau_t[i][isign];
        rss1[i] = gam0[i] * nsr[i] * fabs(pow(rss[i], nsr[i] - 1.)) / tau_t[i][isign];
        rss2[i] = gam0[i] * fabs(pow(rss[i], nsr[i])) * ((double)(2 * (rss[i] > 0) - 1));
        if (gamdot) gamdot[i] = rss2[i];
        if (taux) taux[i] = rss1[i];
    }
    for (i = 0; i < 5; i++) {
        edot[i] = 0.;
        for (j = 0; j < nsys; j++) edot[i] += sc[j][i] * rss2[j];
    }
    edot[5] = 0.0;
    for (i = 0; i < 5; i++)
        for (j = 0; j < 5; j++) {
            d_edot[i][j] = 0.0;
            for (k = 0; k < nsys; k++) d_edot[i][j] += sc[k][i] * sc[k][j] * rss1[k];
        }
    for (i = 0; i < 6; i++) {
        d_edot[i][5] = 0.0;
        d_edot[5][i] = 0.0;
    }
}

void orc_strain_rate_eval(const double sig6[6], const double schm[60], const double tau_t[24],
                          const double xkin[12], int nsys, const double nsr[12],
                          const double gam0[12], double edot[6], double d_edot[36],
                          double gamdot[12], double taux[12])
{
    strain_rate_eval(sig6, (const double(*)[5])schm, (const double(*)[2])tau_t, xkin, nsys, nsr,
                     gam0, edot, (double(*)[6])d_edot, gamdot, taux);
}

/* Orientation, constitutive.c:117-169 */
static void orientation(t2 tran, t2 rot)
{
    int i, mi, mj;
    double w[3], w_norm, w_tan;
    t2 omega, omega2, tran_new;
    w[0] = rot[2][1];
    w[1] = rot[0][2];
    w[2] = rot[1][0];
    w_norm = 0.0;
    for (i = 0; i < 3; i++) w_norm += w[i] * w[i];
    w_norm = sqrt(w_norm);
    w_tan = tan(w_norm / 2.0);
    if (fabs(w_norm) < 1E-6) w_norm = 1.0;
    for (i = 0; i < 3; i++) w[i] *= w_tan / w_norm;
    omega[0][0] = 0.; omega[0][1] = -1. * w[2]; omega[0][2] = w[1];
    omega[1][0] = w[2]; omega[1][1] = 0.; omega[1][2] = -1. * w[0];
    omega[2][0] = -1. * w[1]; omega[2][1] = w[0]; omega[2][2] = 0.;
    for (mi = 0; mi < 3; mi++)
        for (mj = 0; mj < 3; mj++) {
            omega2[mi][mj] = 0.0;
            for (i = 0; i < 3; i++) omega2[mi][mj] += omega[mi][i] * omega[i][mj];
        }
    w_tan = w_tan * w_tan;
    for (mi = 0; mi < 3; mi++)
        for (mj = 0; mj < 3; mj++)
            rot[mi][mj] = (double)(mi == mj) + 2. * (omega[mi][mj] + omega2[mi][mj]) / (1. + w_tan);
    for (mi = 0; mi < 3; mi++)
        for (mj = 0; mj < 3; mj++) {
            tran_new[mi][mj] = 0.0;
            for (i = 0; i < 3; i++) tran_new[mi][mj] += rot[mi][i] * tran[i][mj];
        }
    memcpy(tran, tran_new, sizeof(t2));
}

void orc_orientation(double tran[9], double rot[9])
{
    orientation((double(*)[3])tran, (double(*)[3])rot);
}

/* harden body for one voxel, constitutive.c:264-305 */
static void harden_voxel(double crss[NSYSMX][2], double *gamacum, double *gammatotal,
                         const double *gamdot, int nsys, const double hard[NSYSMX][NSYSMX],
                         const double tau[NSYSMX][3], const double thet[NSYSMX][2], double dt)
{
    int i, j;
    double gamtot, deltgam, deltagamma[NSYSMX], dtau, tau0, tau1, thet0, thet1;
    double voce, fact, exp_ini, exp_del, TINY;
    gamtot = *gamacum;
    for (i = 0; i < nsys; i++) {
        deltagamma[i] = (gamdot[i]) * dt;
        gammatotal[i] = gammatotal[i] + deltagamma[i];
    }
    deltgam = 0.0;
    for (i = 0; i < nsys; i++) deltgam += fabs(gamdot[i]) * dt;
    for (i = 0; i < nsys; i++) {
        dtau = 0.;
        for (j = 0; j < nsys; j++) dtau += hard[i][j] * fabs(gamdot[j]) * dt;
        tau0 = tau[i][0];
        tau1 = tau[i][2];
        thet0 = thet[i][0];
        thet1 = thet[i][1];
        TINY = 1.E-4 * tau0;
        voce = 0.0;
        if (fabs(thet0) > TINY) {
            voce = thet1 * deltgam;
            if (fabs(tau1) > TINY) {
                fact = fabs(thet0 / tau1);
                exp_ini = exp(-1.0 * gamtot * fact);
                exp_del = exp(-1.0 * deltgam * fact);
                voce += -1. * (fact * tau1 - thet1) / fact * exp_ini * (exp_del - 1.0) -
                        thet1 / fact * exp_ini * (exp_del * ((gamtot + deltgam) * fact + 1.0) - (gamtot * fact + 1.0));
            }
        }
        crss[i][0] += dtau * voce / deltgam;
        crss[i][1] += dtau * voce / deltgam;
    }
    *gamacum = gamtot + deltgam;
}

void orc_harden_voxel(double crss[24], double *gamacum, double gamtot[12], const double gamdot[12],
                      int nsys, const double hard[144], const double tau[36], const double thet[24],
                      double dt)
{
    harden_voxel((double(*)[2])crss, gamacum, gamtot, gamdot, nsys, (const double(*)[NSYSMX])hard,
                 (const double(*)[3])tau, (const double(*)[2])thet, dt);
}

/* Latent-hardening table, init.c:283-381 */
void orc_hard_matrix(double self, double coplane, double cross, double glissile, double hirth,
                     double lomer, double Hf[144])
{
    double (*H)[12] = (double(*)[12])Hf;
    int i, j;
    const double S = self, C = coplane, X = cross, G = glissile, Hi = hirth, L = lomer;
    const double up[12][12] = {
        {S, C, C, Hi, G, L, X, G, G, Hi, L, G},
        {0, S, C, G, X, G, G, Hi, L, L, Hi, G},
        {0, 0, S, L, G, Hi, G, L, Hi, G, G, X},
        {0, 0, 0, S, C, C, Hi, L, G, X, G, G},
        {0, 0, 0, 0, S, C, L, Hi, G, G, Hi, L},
        {0, 0, 0, 0, 0, S, G, G, X, G, L, Hi},
        {0, 0, 0, 0, 0, 0, S, C, C, Hi, G, L},
        {0, 0, 0, 0, 0, 0, 0, S, C, G, X, G},
        {0, 0, 0, 0, 0, 0, 0, 0, S, L, G, Hi},
        {0, 0, 0, 0, 0, 0, 0, 0, 0, S, C, C},
        {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, S, C},
        {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, S}};
    for (i = 0; i < 12; i++)
        for (j = 0; j < 12; j++) H[i][j] = (j >= i) ? up[i][j] : up[j][i];
}

/* ------------------------------------------------------------------ */
/* Namelist, io.c:25-122                                                */
/* ------------------------------------------------------------------ */
typedef struct {
    int CellDim[3];
    int N_phases;
    int Type_phases[2];
    double TimeStep;
    int N_steps;
    int Update_Flag, Hard_Flag, Tex_Flag, IterMax;
    int PrintControl[2];
    double Err;
    int VelGrad_BC_Flag[6];
    double VelGrad_BC[6];
    int Stress_BC_Flag[6];
    double Stress_BC[6];
    double ElastConst_PhaseI[4], ElastConst_PhaseII[4];
    char Slip_PhaseI[100], Slip_PhaseII[100], initial_ms[100];
    double self[2], coplane[2], cross[2], glissile[2], hirth[2], lomer[2];
} nml_t;

enum { NI, NR, NS };
typedef struct { const char *name; size_t off; int type; int len; } nml_key;
#define KI(x, n) {#x, offsetof(nml_t, x), NI, n}
#define KR(x, n) {#x, offsetof(nml_t, x), NR, n}
#define KS(x) {#x, offsetof(nml_t, x), NS, 1}
static const nml_key nml_keys[26] = {
    KI(CellDim, 3), KI(N_phases, 1), KI(Type_phases, 2), KR(TimeStep, 1), KI(N_steps, 1),
    KI(Update_Flag, 1), KI(Hard_Flag, 1), KI(Tex_Flag, 1), KI(IterMax, 1), KI(PrintControl, 2),
    KR(Err, 1), KI(VelGrad_BC_Flag, 6), KR(VelGrad_BC, 6), KI(Stress_BC_Flag, 6), KR(Stress_BC, 6),
    KR(ElastConst_PhaseI, 4), KR(ElastConst_PhaseII, 4), KS(Slip_PhaseI), KS(Slip_PhaseII),
    KS(initial_ms), KR(self, 2), KR(coplane, 2), KR(cross, 2), KR(glissile, 2), KR(hirth, 2),
    KR(lomer, 2)};

/* GetNameList, io.c:54-122: 80-byte fgets records, '#' or empty line
 * skipped, a line of only separators ends the parse, EOF seen by the
 * read ends the parse before that (possibly unterminated) line is used,
 * keys case-sensitive, separators ", \t\n" for values. */
static int get_namelist(const char *path, nml_t *nm, int status[26])
{
    char buff[80], *token;
    FILE *fp;
    int j, k, match;
    memset(nm, 0, sizeof(*nm));
    for (k = 0; k < 26; k++) status[k] = 0;
    if ((fp = fopen(path, "r")) == 0) return -1;
    while (1) {
        if (!fgets(buff, 80, fp)) { /* nothing read: buffer keeps old bytes, feof ends */ }
        if (feof(fp)) break;
        if (buff[0] == '#' || buff[0] == '\n') continue;
        token = strtok(buff, " \t\n");
        if (!token) break;
        match = 0;
        for (k = 0; k < 26; k++) {
            if (strcmp(token, nml_keys[k].name) == 0) {
                match = 1;
                if (status[k] == 0) {
                    status[k] = 1;
                    for (j = 0; j < nml_keys[k].len; j++) {
                        token = strtok(NULL, ", \t\n");
                        if (token) {
                            char *base = (char *)nm + nml_keys[k].off;
                            if (nml_keys[k].type == NI) ((int *)base)[j] = (int)atol(token);
                            else if (nml_keys[k].type == NR) ((double *)base)[j] = atof(token);
                            else strcpy(base + j, token);
                        } else {
                            status[k] = 2;
                        }
                    }
                    token = strtok(NULL, ", \t\n");
                    if (token) status[k] = 3;
                    break;
                } else {
                    status[k] = 4;
                }
            }
        }
        (void)match;
    }
    fclose(fp);
    return 0;
}

int orc_parse_namelist(const char *path, int status[26], double *vals, int nvals)
{
    nml_t nm;
    int k, j, cnt = 0, w = 0;
    if (get_namelist(path, &nm, status)) return -1;
    for (k = 0; k < 26; k++) {
        if (status[k] == 1) cnt++;
        if (nml_keys[k].type == NS) continue;
        for (j = 0; j < nml_keys[k].len; j++) {
            const char *base = (const char *)&nm + nml_keys[k].off;
            double v = (nml_keys[k].type == NI) ? (double)((const int *)base)[j] : ((const double *)base)[j];
            if (vals && w < nvals) vals[w] = v;
            w++;
        }
    }
    return cnt;
}

/* ------------------------------------------------------------------ */
/* Job                                                                  */
/* ------------------------------------------------------------------ */
struct orc_job {
    nml_t nm;
    int Nx, Ny, Nz, V;
    double WGT, Wgt_ph1;
    /* per phase material (init.c:9-391) */
    double Cijkl[NPHMAX][3][3][3][3];
    int nSYS[NPHMAX];
    double nSRS[NPHMAX][NSYSMX], gam0[NPHMAX][NSYSMX];
    double tau[NPHMAX][NSYSMX][3], thet[NPHMAX][NSYSMX][2];
    double bn[NPHMAX][NSYSMX][3], bb[NPHMAX][NSYSMX][3];
    double Schm_xt[NPHMAX][NSYSMX][5];
    double Hard[NPHMAX][NSYSMX][NSYSMX];
    double C0[3][3][3][3], S0[3][3][3][3];
    double C066[6][6], S066[6][6];
    /* BC and macro state */
    t2 Udot, Udot_s, Udot_a, Scauchy;
    double d_vm, e_vm, s_vm, s_vm1;
    int CREEP_FLAG;
    t2 DisGradAvg, dDisGradAvg, dDisGradAvg_acum, DisGradAvg_t, DisGradAvg_actual;
    t2 SigAvg, SigAvg1, EpsAvg, EdotAvg, VelGradAvg;
    int istep, iter_total;
    /* fields */
    int *grain_f, *phase_f;
    double (*C_gr)[6][6];
    t2 *TranMat;
    double (*Schm_gr)[NSYSMX][5];
    double (*crss)[NSYSMX][2];
    double (*gamdot)[NSYSMX], (*tau_x)[NSYSMX], (*gammatotal)[NSYSMX];
    double *gamacum, *SVM, *EVM, *work_p, *rho_rss, *rho_crss;
    t2 *Sig, *Eps, *Edot, *DisGrad, *VelGrad, *fluctuation;
    t2 *kSig_r, *kSig_i;
    cpx *fft;
    double *err_e_v, *err_s_v;
    double *nr_iters;     /* Newton iterations of each voxel's last UpdateStress */
    int error;
};

static void join_path(char *out, size_t n, const char *base, const char *name)
{
    if (base && name[0] != '/') snprintf(out, n, "%s/%s", base, name);
    else snprintf(out, n, "%s", name);
}

/* ElastStiffnessMatrix, init.c:9-49 */
static void elast_stiffness(const double *ec, double c[3][3][3][3])
{
    double c11 = ec[0], c12 = ec[1], c44 = ec[2];
    memset(c, 0, sizeof(double) * 81);
    c[0][0][0][0] = c11; c[1][1][1][1] = c11; c[2][2][2][2] = c11;
    c[0][0][1][1] = c12; c[1][1][2][2] = c12; c[2][2][0][0] = c12;
    c[0][1][0][1] = c44; c[1][2][1][2] = c44; c[2][0][2][0] = c44;
    c[1][1][0][0] = c12; c[2][2][1][1] = c12; c[0][0][2][2] = c12;
    c[1][0][1][0] = c44; c[0][1][1][0] = c44; c[1][0][0][1] = c44;
    c[2][1][2][1] = c44; c[2][1][1][2] = c44; c[1][2][2][1] = c44;
    c[0][2][0][2] = c44; c[0][2][2][0] = c44; c[2][0][0][2] = c44;
}

/* PlasticInit, init.c:51-391: positional parse of the .sx file. */
static int plastic_init(orc_job *J, const char *path, int phase)
{
    FILE *fp;
    char buffer[80] = {0};
    int i, j, mx = 0, m_count, matcheck, N_modes_max = 0, N_modes = 0, iMode[64];
    int nsmx = 0, isectwx = 0;
    double nrsx = 0, gamd0x = 0, twshx = 0, tau0xf = 0, tau0xb = 0, tau1x = 0, thet0 = 0, thet1 = 0;
    double hselfx = 0, hlatex2 = 0, norm, aux5[5];
    t2 aux33;
    fp = fopen(path, "r");
    if (!fp) return -1;
#define RD() (void)!fgets(buffer, 80, fp)
    RD();
    RD();
    RD(); RD();
    RD();
    RD(); RD();
    RD(); sscanf(buffer, "%d", &N_modes_max);
    RD(); RD();
    RD(); sscanf(buffer, "%d", &N_modes);
    if (N_modes > 64) N_modes = 64;
    RD(); RD();
    for (i = 0; i < N_modes; i++) { RD(); sscanf(buffer, "%d", &iMode[i]); }
    RD();
    m_count = 0;
    for (i = 0; i < N_modes_max; i++) {
        RD();
        sscanf(buffer, "%*s %*s %d %*s", &mx);
        matcheck = 0;
        for (j = 0; j < N_modes; j++)
            if (mx == iMode[j]) { matcheck = 1; m_count++; break; }
        if (matcheck == 1) {
            RD(); RD();
            sscanf(buffer, "%d\t\t%lf\t\t%lf\t\t%lf\t\t%d", &nsmx, &nrsx, &gamd0x, &twshx, &isectwx);
            if (nsmx > NSYSMX) { fclose(fp); return -2; }
            J->nSYS[phase] = nsmx;
            RD(); RD();
            sscanf(buffer, "%lf\t\t%lf\t\t%lf\t\t%lf\t\t%lf", &tau0xf, &tau0xb, &tau1x, &thet0, &thet1);
            RD(); RD();
            sscanf(buffer, "%lf\t\t%lf", &hselfx, &hlatex2);
            RD();
            if (tau1x < 1.E-6) { tau1x = 1.E-6; thet0 = thet1; }
            for (j = 0; j < nsmx; j++) {
                J->nSRS[phase][j] = nrsx;
                J->gam0[phase][j] = gamd0x;
                J->tau[phase][j][0] = tau0xf;
                J->tau[phase][j][1] = tau0xb;
                J->tau[phase][j][2] = tau1x;
                J->thet[phase][j][0] = thet0;
                J->thet[phase][j][1] = thet1;
                RD();
                sscanf(buffer, "%lf %lf %lf\t\t%lf %lf %lf", &J->bn[phase][j][0], &J->bn[phase][j][1],
                       &J->bn[phase][j][2], &J->bb[phase][j][0], &J->bb[phase][j][1], &J->bb[phase][j][2]);
            }
        } else {
            for (j = 0; j < 24; j++) RD();
        }
    }
#undef RD
    fclose(fp);
    /* init.c:191-226: normalisation indexes bn/bb by mode i (reference quirk) */
    for (i = 0; i < N_modes; i++) {
        for (j = 0; j < nsmx; j++) {
            int mi, mj;
            norm = sqrt(pow(J->bn[i][j][0], 2.0) + pow(J->bn[i][j][1], 2.0) + pow(J->bn[i][j][2], 2.0));
            for (mx = 0; mx < 3; mx++) J->bn[i][j][mx] /= norm;
            norm = sqrt(pow(J->bb[i][j][0], 2.0) + pow(J->bb[i][j][1], 2.0) + pow(J->bb[i][j][2], 2.0));
            for (mx = 0; mx < 3; mx++) J->bb[i][j][mx] /= norm;
            norm = 0.0;
            for (mx = 0; mx < 3; mx++) norm += J->bn[i][j][mx] * J->bb[i][j][mx];
            if (norm > 1E-3) return 7;
            for (mi = 0; mi < 3; mi++)
                for (mj = 0; mj < 3; mj++)
                    aux33[mi][mj] = (J->bn[i][j][mi] * J->bb[i][j][mj] + J->bn[i][j][mj] * J->bb[i][j][mi]) / 2.0;
            chg_basis5(aux5, aux33, NULL, NULL, 2);
            for (mx = 0; mx < 5; mx++) J->Schm_xt[phase][j][mx] = aux5[mx];
        }
    }
    orc_hard_matrix(J->nm.self[phase], J->nm.coplane[phase], J->nm.cross[phase], J->nm.glissile[phase],
                    J->nm.hirth[phase], J->nm.lomer[phase], &J->Hard[phase][0][0]);
    (void)hselfx; (void)hlatex2; (void)twshx; (void)isectwx;
    return 0;
}

/* InitMicroStruct, init.c:393-483 (single rank: lxs = 0). */
static int init_microstruct(orc_job *J, const char *path)
{
    FILE *fp;
    char buffer[80] = {0};
    double ph, th, om, caux3333[3][3][3][3];
    double c066_local[6][6];
    int ii, jj, kk, jgr, jph, nph1 = 0, px, py, pz, mi, mj;
    t2 sa2xt;
    memset(c066_local, 0, sizeof(c066_local));
    fp = fopen(path, "r");
    if (!fp) return -1;
    for (px = 0; px < J->Nx; px++)
        for (py = 0; py < J->Ny; py++)
            for (pz = 0; pz < J->Nz; pz++) {
                int p = (px * J->Ny + py) * J->Nz + pz, idx;
                if (!fgets(buffer, 80, fp)) { fclose(fp); return 1139; }
                if (sscanf(buffer, "%lf %lf %lf %d %d %d %d %d", &ph, &th, &om, &ii, &jj, &kk, &jgr, &jph) != 8) {
                    fclose(fp);
                    return 1139;
                }
                idx = ((ii - 1) * J->Ny + jj - 1) * J->Nz + kk - 1;
                if (idx != p) { fclose(fp); return 1139; }
                if (jph < 1 || jph > J->nm.N_phases) { fclose(fp); return 1139; }
                if (jph == 1) nph1++;
                J->grain_f[p] = jgr;
                J->phase_f[p] = jph;
                if (!J->nm.Type_phases[jph - 1]) {
                    ph *= PI_REF / 180.; th *= PI_REF / 180.; om *= PI_REF / 180.;
                    orc_euler_to_mat(&ph, &th, &om, &sa2xt[0][0], 2);
                    for (mi = 0; mi < 3; mi++)
                        for (mj = 0; mj < 3; mj++) J->TranMat[p][mi][mj] = sa2xt[mj][mi];
                    orc_ten4th_transform(&caux3333[0][0][0][0], &sa2xt[0][0], &J->Cijkl[jph - 1][0][0][0][0], 2);
                    chg_basis(NULL, NULL, J->C_gr[p], caux3333, 4);
                    for (mi = 0; mi < 6; mi++)
                        for (mj = 0; mj < 6; mj++) c066_local[mi][mj] += J->C_gr[p][mi][mj];
                }
            }
    fclose(fp);
    J->Wgt_ph1 = (1.0 * (double)nph1) / J->V;
    for (mi = 0; mi < 6; mi++)
        for (mj = 0; mj < 6; mj++) {
            J->C066[mi][mj] = c066_local[mi][mj];
            J->C066[mi][mj] *= J->WGT;
        }
    memcpy(J->S066, J->C066, sizeof(J->S066));
    if (lu_inv66(J->S066)) return 35;
    chg_basis(NULL, NULL, J->C066, J->C0, 3);
    chg_basis(NULL, NULL, J->S066, J->S0, 3);
    return 0;
}

int orc_nvox(const orc_job *J) { return J->V; }
int orc_nsteps(const orc_job *J) { return J->nm.N_steps; }
int orc_dims(const orc_job *J, int d[3]) { d[0] = J->Nx; d[1] = J->Ny; d[2] = J->Nz; return J->V; }

void orc_free(orc_job *J)
{
    if (!J) return;
    free(J->grain_f); free(J->phase_f); free(J->C_gr); free(J->TranMat); free(J->Schm_gr);
    free(J->crss); free(J->gamdot); free(J->tau_x); free(J->gammatotal); free(J->gamacum);
    free(J->SVM); free(J->EVM); free(J->work_p); free(J->rho_rss); free(J->rho_crss);
    free(J->Sig); free(J->Eps); free(J->Edot); free(J->DisGrad); free(J->VelGrad);
    free(J->fluctuation); free(J->kSig_r); free(J->kSig_i); free(J->fft);
    free(J->err_e_v); free(J->err_s_v); free(J->nr_iters);
    free(J);
}

/* SetupJob, init.c:617-828 */
orc_job *orc_setup(const char *input_path, const char *base_dir, int *err)
{
    orc_job *J = (orc_job *)calloc(1, sizeof(orc_job));
    int status[26], i, mi, mj, rc, p;
    char path[1200];
    double D_bar6[6], D_bar5[5];
    t2 dev_Udot_s;
    if (err) *err = 0;
    if (get_namelist(input_path, &J->nm, status)) { if (err) *err = -1; free(J); return NULL; }
    if (!Bb_ready) build_B();
    J->Nx = J->nm.CellDim[0]; J->Ny = J->nm.CellDim[1]; J->Nz = J->nm.CellDim[2];
    J->V = J->Nx * J->Ny * J->Nz;
    if (J->V <= 0 || J->nm.N_phases < 1 || J->nm.N_phases > NPHMAX || J->nm.IterMax > ORC_MAXIT) {
        if (err) *err = -1;
        free(J);
        return NULL;
    }
    J->WGT = 1. / J->V;
    {
        long V = J->V;
        J->grain_f = (int *)calloc(V, sizeof(int));
        J->phase_f = (int *)calloc(V, sizeof(int));
        J->C_gr = calloc(V, sizeof(*J->C_gr));
        J->TranMat = calloc(V, sizeof(t2));
        J->Schm_gr = calloc(V, sizeof(*J->Schm_gr));
        J->crss = calloc(V, sizeof(*J->crss));
        J->gamdot = calloc(V, sizeof(*J->gamdot));
        J->tau_x = calloc(V, sizeof(*J->tau_x));
        J->gammatotal = calloc(V, sizeof(*J->gammatotal));
        J->gamacum = calloc(V, sizeof(double));
        J->SVM = calloc(V, sizeof(double));
        J->EVM = calloc(V, sizeof(double));
        J->work_p = calloc(V, sizeof(double));
        J->rho_rss = calloc(V, sizeof(double));
        J->rho_crss = calloc(V, sizeof(double));
        J->Sig = calloc(V, sizeof(t2));
        J->Eps = calloc(V, sizeof(t2));
        J->Edot = calloc(V, sizeof(t2));
        J->DisGrad = calloc(V, sizeof(t2));
        J->VelGrad = calloc(V, sizeof(t2));
        J->fluctuation = calloc(V, sizeof(t2));
        J->kSig_r = calloc(V, sizeof(t2));
        J->kSig_i = calloc(V, sizeof(t2));
        J->fft = calloc(V, sizeof(cpx));
        J->err_e_v = calloc(V, sizeof(double));
        J->nr_iters = calloc(V, sizeof(double));
        J->err_s_v = calloc(V, sizeof(double));
    }
    /* material, init.c:696-703 */
    if (!J->nm.Type_phases[0]) {
        elast_stiffness(J->nm.ElastConst_PhaseI, J->Cijkl[0]);
        join_path(path, sizeof(path), base_dir, J->nm.Slip_PhaseI);
        rc = plastic_init(J, path, 0);
        if (rc) { if (err) *err = rc; orc_free(J); return NULL; }
    }
    if (J->nm.N_phases == 2 && !J->nm.Type_phases[1]) {
        elast_stiffness(J->nm.ElastConst_PhaseII, J->Cijkl[1]);
        join_path(path, sizeof(path), base_dir, J->nm.Slip_PhaseII);
        rc = plastic_init(J, path, 1);
        if (rc) { if (err) *err = rc; orc_free(J); return NULL; }
    }
    join_path(path, sizeof(path), base_dir, J->nm.initial_ms);
    rc = init_microstruct(J, path);
    if (rc) { if (err) *err = rc; orc_free(J); return NULL; }
    /* boundary conditions, init.c:713-745 */
    if ((J->nm.VelGrad_BC_Flag[0] + J->nm.VelGrad_BC_Flag[1] + J->nm.VelGrad_BC_Flag[2]) == 2) {
        if (err) *err = 116;
        orc_free(J);
        return NULL;
    }
    voigt_to_full(J->nm.VelGrad_BC, J->Udot);
    sym_ant(J->Udot, J->Udot_s, J->Udot_a);
    chg_basis(D_bar6, J->Udot_s, NULL, NULL, 2);
    for (i = 0; i < 5; i++) D_bar5[i] = D_bar6[i];
    chg_basis5(D_bar5, dev_Udot_s, NULL, NULL, 1);
    J->d_vm = 0.0;
    for (mi = 0; mi < 3; mi++)
        for (mj = 0; mj < 3; mj++) J->d_vm += pow(dev_Udot_s[mi][mj], 2.0);
    J->d_vm = sqrt(2. / 3. * J->d_vm);
    for (i = 0; i < 6; i++) {
        if ((J->nm.Stress_BC_Flag[i] * J->nm.VelGrad_BC_Flag[i] != 0) ||
            ((J->nm.Stress_BC_Flag[i] + J->nm.VelGrad_BC_Flag[i]) != 1)) {
            if (err) *err = 114;
            orc_free(J);
            return NULL;
        }
    }
    J->CREEP_FLAG = 0;
    for (i = 0; i < 6; i++) J->CREEP_FLAG += J->nm.Stress_BC_Flag[i];
    J->CREEP_FLAG /= 6;
    voigt_to_full(J->nm.Stress_BC, J->Scauchy);
    /* CRSS, init.c:776-796 */
    for (p = 0; p < J->V; p++) {
        int jph = J->phase_f[p];
        if (J->nm.Type_phases[jph - 1] == 0)
            for (i = 0; i < J->nSYS[jph - 1]; i++) {
                J->crss[p][i][0] = J->tau[jph - 1][i][0];
                J->crss[p][i][1] = J->tau[jph - 1][i][1];
            }
    }
    /* InitFields, init.c:486-544 */
    for (mi = 0; mi < 3; mi++)
        for (mj = 0; mj < 3; mj++) {
            J->DisGradAvg[mi][mj] = J->Udot[mi][mj] * J->nm.TimeStep;
            J->dDisGradAvg[mi][mj] = 0.0;
            J->dDisGradAvg_acum[mi][mj] = 0.0;
        }
    for (p = 0; p < J->V; p++) {
        int jph = J->phase_f[p], mip, mjp;
        if (!J->nm.Type_phases[jph - 1]) {
            double cg[3][3][3][3];
            chg_basis(NULL, NULL, J->C_gr[p], cg, 3);
            for (mi = 0; mi < 3; mi++)
                for (mj = 0; mj < 3; mj++) {
                    J->Sig[p][mi][mj] = 0.0;
                    for (mip = 0; mip < 3; mip++)
                        for (mjp = 0; mjp < 3; mjp++) J->Sig[p][mi][mj] += cg[mi][mj][mip][mjp] * J->DisGradAvg[mip][mjp];
                    if (J->CREEP_FLAG == 1) J->Sig[p][mi][mj] += J->Scauchy[mi][mj];
                }
        }
    }
    J->e_vm = J->d_vm * J->nm.TimeStep;
    J->istep = 0;
    J->iter_total = 0;
    return J;
}

/* update_schmid body for one voxel, kinematics.c:436-456: every system's
 * crystal-frame Schmid vector (5 B-basis components) to a tensor, rotated
 * with the voxel's TranMat (crystal -> sample) as R P R^T, and back. */
void orc_update_schmid_voxel(const double tran[9], const double *schm_xt, int nsys, double *out)
{
    int i, j, mi, mj, mip, mjp;
    double aux5[5];
    t2 aux33, aux33r;
    for (i = 0; i < nsys; i++) {
        for (j = 0; j < 5; j++) aux5[j] = schm_xt[i * 5 + j];
        chg_basis5(aux5, aux33, NULL, NULL, 1);
        for (mi = 0; mi < 3; mi++)
            for (mj = 0; mj < 3; mj++) {
                aux33r[mi][mj] = 0.0;
                for (mip = 0; mip < 3; mip++)
                    for (mjp = 0; mjp < 3; mjp++)
                        aux33r[mi][mj] += tran[mi * 3 + mip] * tran[mj * 3 + mjp] * aux33[mip][mjp];
            }
        chg_basis5(aux5, aux33r, NULL, NULL, 2);
        for (j = 0; j < 5; j++) out[i * 5 + j] = aux5[j];
    }
}

/* update_schmid, kinematics.c:426-459 */
static void update_schmid(orc_job *J)
{
    int p;
#pragma omp parallel for schedule(static)
    for (p = 0; p < J->V; p++) {
        int jph = J->phase_f[p];
        if (J->nm.Type_phases[jph - 1]) continue;
        orc_update_schmid_voxel(&J->TranMat[p][0][0], &J->Schm_xt[jph - 1][0][0], J->nSYS[jph - 1],
                                &J->Schm_gr[p][0][0]);
    }
}

/* Per-voxel body of UpdateStress, evolution.c:28-157.  Returns 0 or 35. */
static int update_stress_voxel(const double C_gr[6][6], const double C066[6][6], t2 Sig, const t2 Eps,
                               const t2 DisGrad, const double sc[NSYSMX][5], const double tau_t[NSYSMX][2],
                               int nsys, const double *nsr, const double *gam0, double dt, t2 Edot,
                               double *err_e, double *err_s, double *SVMp, double *EVMp,
                               double *gamdot, double *taux, double *niter)
{
    double sg66[6][6], jacob_inv[6][6], d_edotp66[6][6];
    t2 xlambda_aux, sig_aux, eps_aux, strain_aux, edotp_aux, tot_e, SigDev, eDev;
    double xlambda6[6], sig6[6], eps6[6], strain6[6], edotp6[6], tot_eps[6], sig6_old[6], res[6];
    double signorm, epsnorm, erroral, erral, conv_NR, conv_e = 0, conv_s = 0, trace, v;
    double xkin0[NSYSMX] = {0};
    int i, j, k, iterl, itmaxal, mi, mj;
    memcpy(sg66, C_gr, sizeof(sg66));
    if (lu_inv66(sg66)) return 35;
    for (mi = 0; mi < 3; mi++)
        for (mj = 0; mj < 3; mj++) {
            xlambda_aux[mi][mj] = Sig[mi][mj];
            sig_aux[mi][mj] = Sig[mi][mj];
            eps_aux[mi][mj] = Eps[mi][mj];
            strain_aux[mi][mj] = (DisGrad[mi][mj] + DisGrad[mj][mi]) / 2.0;
        }
    chg_basis(xlambda6, xlambda_aux, NULL, NULL, 2);
    chg_basis(sig6, sig_aux, NULL, NULL, 2);
    chg_basis(eps6, eps_aux, NULL, NULL, 2);
    chg_basis(strain6, strain_aux, NULL, NULL, 2);
    signorm = 0.0;
    epsnorm = 0.0;
    for (mi = 0; mi < 3; mi++)
        for (mj = 0; mj < 3; mj++) {
            signorm += xlambda_aux[mi][mj] * xlambda_aux[mi][mj];
            epsnorm += strain_aux[mi][mj] * strain_aux[mi][mj];
        }
    signorm = sqrt(signorm);
    (void)epsnorm;
    erroral = 1E-7;
    itmaxal = 100;
    iterl = 0;
    erral = 10 * erroral;
    while ((iterl < itmaxal) && (fabs(erral) > fabs(erroral))) {
        iterl++;
        for (i = 0; i < 6; i++) sig6_old[i] = sig6[i];
        strain_rate_eval(sig6, sc, tau_t, xkin0, nsys, nsr, gam0, edotp6, d_edotp66, gamdot, taux);
        for (i = 0; i < 6; i++) {
            tot_eps[i] = eps6[i] + edotp6[i] * dt;
            for (j = 0; j < 6; j++) tot_eps[i] += sg66[i][j] * sig6[j];
        }
        for (i = 0; i < 6; i++) {
            res[i] = sig6[i] - xlambda6[i];
            for (j = 0; j < 6; j++) res[i] += C066[i][j] * (tot_eps[j] - strain6[j]);
        }
        for (i = 0; i < 6; i++)
            for (j = 0; j < 6; j++) {
                jacob_inv[i][j] = (double)(i == j);
                for (k = 0; k < 6; k++) jacob_inv[i][j] += C066[i][k] * (sg66[k][j] + d_edotp66[k][j] * dt);
            }
        if (lu_inv66(jacob_inv)) return 35;
        for (i = 0; i < 6; i++)
            for (j = 0; j < 6; j++) sig6[i] -= jacob_inv[i][j] * res[j];
        conv_NR = 0.0;
        conv_e = 0.0;
        conv_s = 0.0;
        for (i = 0; i < 6; i++) {
            conv_NR += (sig6[i] - sig6_old[i]) * (sig6[i] - sig6_old[i]);
            conv_e += (tot_eps[i] - strain6[i]) * (tot_eps[i] - strain6[i]);
            conv_s += (sig6[i] - xlambda6[i]) * (sig6[i] - xlambda6[i]);
        }
        erral = sqrt(conv_NR) / signorm;
        conv_e = sqrt(conv_e);
        conv_s = sqrt(conv_s);
    }
    if (niter) *niter = iterl;
    chg_basis(sig6, sig_aux, NULL, NULL, 1);
    chg_basis(edotp6, edotp_aux, NULL, NULL, 1);
    chg_basis(tot_eps, tot_e, NULL, NULL, 1);
    for (mi = 0; mi < 3; mi++)
        for (mj = 0; mj < 3; mj++) {
            Sig[mi][mj] = sig_aux[mi][mj];
            Edot[mi][mj] = edotp_aux[mi][mj];
        }
    trace = sig_aux[0][0] + sig_aux[1][1] + sig_aux[2][2];
    for (mi = 0; mi < 3; mi++)
        for (mj = 0; mj < 3; mj++)
            SigDev[mi][mj] = (sig_aux[mi][mj] + sig_aux[mj][mi]) / 2. - (double)(mi == mj) * trace / 3.0;
    v = 0.0;
    for (mi = 0; mi < 3; mi++)
        for (mj = 0; mj < 3; mj++) v += SigDev[mi][mj] * SigDev[mi][mj];
    *SVMp = sqrt(3. / 2. * v);
    trace = tot_e[0][0] + tot_e[1][1] + tot_e[2][2];
    for (mi = 0; mi < 3; mi++)
        for (mj = 0; mj < 3; mj++)
            eDev[mi][mj] = (tot_e[mi][mj] + tot_e[mj][mi]) / 2. - (double)(mi == mj) * trace / 3.0;
    v = 0.0;
    for (mi = 0; mi < 3; mi++)
        for (mj = 0; mj < 3; mj++) v += eDev[mi][mj] * eDev[mi][mj];
    *EVMp = sqrt(2. / 3. * v);
    *err_s = conv_s;
    *err_e = conv_e;
    return 0;
}

int orc_update_stress_voxel(const double C_gr[36], const double C066[36], const double Sig[9],
                            const double Eps[9], const double DisGrad[9], const double schm[60],
                            const double tau_t[24], int nsys, const double nsr[12], const double gam0[12],
                            double dt, double Sig_out[9], double Edot_out[9], double *err_e,
                            double *err_s, double *svm, double *evm, double gamdot[12], double taux[12])
{
    memcpy(Sig_out, Sig, 9 * sizeof(double));
    return update_stress_voxel((const double(*)[6])C_gr, (const double(*)[6])C066, (double(*)[3])Sig_out,
                               (const double(*)[3])Eps, (const double(*)[3])DisGrad, (const double(*)[5])schm,
                               (const double(*)[2])tau_t, nsys, nsr, gam0, dt, (double(*)[3])Edot_out, err_e,
                               err_s, svm, evm, gamdot, taux, NULL);
}

/* UpdateStress, evolution.c:3-168: voxel loop, then the serial error sum */
static int update_stress(orc_job *J, double *Err_e_local, double *Err_s_local)
{
    int p, fail = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : fail)
    for (p = 0; p < J->V; p++) {
        int jph = J->phase_f[p];
        J->err_e_v[p] = 0.0;
        J->err_s_v[p] = 0.0;
        if (!J->nm.Type_phases[jph - 1]) {
            int rc = update_stress_voxel(J->C_gr[p], J->C066, J->Sig[p], J->Eps[p], J->DisGrad[p], J->Schm_gr[p],
                                         J->crss[p], J->nSYS[jph - 1], J->nSRS[jph - 1], J->gam0[jph - 1],
                                         J->nm.TimeStep, J->Edot[p], &J->err_e_v[p], &J->err_s_v[p], &J->SVM[p],
                                         &J->EVM[p], J->gamdot[p], J->tau_x[p], &J->nr_iters[p]);
            if (rc) fail |= 1;
        } else {
            memset(J->Sig[p], 0, sizeof(t2));
            memset(J->Edot[p], 0, sizeof(t2));
        }
    }
    if (fail) return 35;
    /* evolution.c:156-157: sums in voxel order (gas voxels add nothing) */
    for (p = 0; p < J->V; p++) {
        int jph = J->phase_f[p];
        if (!J->nm.Type_phases[jph - 1]) {
            *Err_s_local += J->err_s_v[p];
            *Err_e_local += J->err_e_v[p];
        }
    }
    return 0;
}

/* Mixed-BC strain correction of get_smacro, evolution.c:199-211: for every
 * Voigt component i that is not velocity-controlled, the entry (IJV[0][i],
 * IJV[1][i]) -- upper triangle only -- of dD gets S0 contracted with the
 * stress misfit of the stress-controlled Voigt components (again only
 * their upper-triangle entries).  Every other entry of dD is left as is,
 * as in the reference (it only writes these six). */
void orc_bc_correction(const double S0[81], const int vflag[6], const int sflag[6], const double Scauchy[9],
                       const double SigAvg[9], double dD[9])
{
    static const int IJV[2][6] = {{0, 1, 2, 1, 0, 0}, {0, 1, 2, 2, 2, 1}};
    int i, k, ii, jj, kk, ll;
    for (i = 0; i < 6; i++) {
        ii = IJV[0][i];
        jj = IJV[1][i];
        dD[ii * 3 + jj] = 0.0;
        if (vflag[i] == 0) {
            for (k = 0; k < 6; k++) {
                kk = IJV[0][k];
                ll = IJV[1][k];
                dD[ii * 3 + jj] += S0[((ii * 3 + jj) * 3 + kk) * 3 + ll] * sflag[k] *
                                   (Scauchy[kk * 3 + ll] - SigAvg[kk * 3 + ll]);
            }
        }
    }
}

/* Error normalisation of Evolution, evolution.c:410-423: the voxel sums
 * times WGT, Err_e divided by e_vm (the previous increment's macroscopic
 * strain, d_vm*dt at the first, init.c:536) only when |e_vm| > 1e-5,
 * Err_s divided by the s_vm get_smacro has just formed. */
void orc_err_normalise(double err_e_sum, double err_s_sum, double wgt, double e_vm, double s_vm, double out[2])
{
    double Err_e = err_e_sum, Err_s = err_s_sum;
    Err_e *= wgt;
    Err_s *= wgt;
    if (fabs(e_vm) > 1E-5) Err_e /= e_vm;
    Err_s /= s_vm;
    out[0] = Err_e;
    out[1] = Err_s;
}

/* get_smacro, evolution.c:170-241 */
static void get_smacro(orc_job *J)
{
    double local_sigavg, local_sigavg1, sav6[6], sav5[5], v;
    t2 SigDevAvg;
    int i, mi, mj, p;
    for (mi = 0; mi < 3; mi++)
        for (mj = 0; mj < 3; mj++) {
            local_sigavg = 0.0;
            local_sigavg1 = 0.0;
            for (p = 0; p < J->V; p++) {
                local_sigavg += J->Sig[p][mi][mj];
                if (J->phase_f[p] == 1) local_sigavg1 += J->Sig[p][mi][mj];
            }
            J->SigAvg[mi][mj] = local_sigavg;
            J->SigAvg1[mi][mj] = local_sigavg1;
            J->SigAvg[mi][mj] *= J->WGT;
            J->SigAvg1[mi][mj] *= J->WGT / J->Wgt_ph1;
        }
    orc_bc_correction(&J->S0[0][0][0][0], J->nm.VelGrad_BC_Flag, J->nm.Stress_BC_Flag, &J->Scauchy[0][0],
                      &J->SigAvg[0][0], &J->dDisGradAvg[0][0]);
    for (mi = 0; mi < 3; mi++)
        for (mj = 0; mj < 3; mj++) J->dDisGradAvg_acum[mi][mj] += J->dDisGradAvg[mi][mj];
    chg_basis(sav6, J->SigAvg, NULL, NULL, 2);
    for (i = 0; i < 5; i++) sav5[i] = sav6[i];
    chg_basis5(sav5, SigDevAvg, NULL, NULL, 1);
    v = 0.0;
    for (mi = 0; mi < 3; mi++)
        for (mj = 0; mj < 3; mj++) v += SigDevAvg[mi][mj] * SigDevAvg[mi][mj];
    J->s_vm = sqrt(3. / 2. * v);
    chg_basis(sav6, J->SigAvg1, NULL, NULL, 2);
    for (i = 0; i < 5; i++) sav5[i] = sav6[i];
    chg_basis5(sav5, SigDevAvg, NULL, NULL, 1);
    v = 0.0;
    for (mi = 0; mi < 3; mi++)
        for (mj = 0; mj < 3; mj++) v += SigDevAvg[mi][mj] * SigDevAvg[mi][mj];
    J->s_vm1 = sqrt(3. / 2. * v);
}

/* VonMises, evolution.c:243-262 */
static double von_mises(t2 t)
{
    double trace, vm = 0.0;
    t2 dt;
    int mi, mj;
    trace = t[0][0] + t[1][1] + t[2][2];
    for (mi = 0; mi < 3; mi++)
        for (mj = 0; mj < 3; mj++) dt[mi][mj] = (t[mi][mj] + t[mj][mi]) / 2. - (double)(mi == mj) * trace / 3.0;
    for (mi = 0; mi < 3; mi++)
        for (mj = 0; mj < 3; mj++) vm += dt[mi][mj] * dt[mi][mj];
    return (sqrt(2. / 3. * vm));
}

/* update_orient body for one voxel, constitutive.c:188-229: local spin
 * skew(VelGrad), plastic spin skew(sum_s (R bb_s) (x) (R bn_s) gamdot_s)
 * with R = TranMat (crystal -> sample), rotation (Udot_a + W_loc - W_p) dt
 * applied by Orientation. */
void orc_update_orient_voxel(double tran[9], const double velgrad[9], const double udot_a[9], const double *bn,
                             const double *bb, const double *gamdot, int nsys, double dt)
{
    int is, i, j, mi, mj;
    t2 tranmat, RotLoc, Lp, RotSlip, Rot;
    double bb_sa[3], bn_sa[3];
    for (mi = 0; mi < 3; mi++)
        for (mj = 0; mj < 3; mj++) {
            RotLoc[mi][mj] = (velgrad[mi * 3 + mj] - velgrad[mj * 3 + mi]) / 2.;
            Lp[mi][mj] = 0.0;
            tranmat[mi][mj] = tran[mi * 3 + mj];
        }
    for (is = 0; is < nsys; is++) {
        for (i = 0; i < 3; i++) {
            bb_sa[i] = 0.0;
            bn_sa[i] = 0.0;
            for (j = 0; j < 3; j++) {
                bb_sa[i] += tranmat[i][j] * bb[is * 3 + j];
                bn_sa[i] += tranmat[i][j] * bn[is * 3 + j];
            }
        }
        for (mi = 0; mi < 3; mi++)
            for (mj = 0; mj < 3; mj++) Lp[mi][mj] += bb_sa[mi] * bn_sa[mj] * gamdot[is];
    }
    for (mi = 0; mi < 3; mi++)
        for (mj = 0; mj < 3; mj++) RotSlip[mi][mj] = (Lp[mi][mj] - Lp[mj][mi]) / 2.;
    for (mi = 0; mi < 3; mi++)
        for (mj = 0; mj < 3; mj++) Rot[mi][mj] = (udot_a[mi * 3 + mj] + RotLoc[mi][mj] - RotSlip[mi][mj]) * dt;
    orientation(tranmat, Rot);
    memcpy(tran, tranmat, sizeof(t2));
}

/* update_orient, constitutive.c:171-248 (diagnostic averages dropped) */
static void update_orient(orc_job *J)
{
    int p;
#pragma omp parallel for schedule(static)
    for (p = 0; p < J->V; p++) {
        int jph = J->phase_f[p];
        if (J->nm.Type_phases[jph - 1]) continue;
        orc_update_orient_voxel(&J->TranMat[p][0][0], &J->VelGrad[p][0][0], &J->Udot_a[0][0],
                                &J->bn[jph - 1][0][0], &J->bb[jph - 1][0][0], J->gamdot[p], J->nSYS[jph - 1],
                                J->nm.TimeStep);
    }
}

/* rho/work diagnostics of one voxel, evolution.c:509-529: with the
 * resolved shear rss_s = Schm_gr_s . chg_basis(Sig) (5 B-basis terms),
 * rho_rss = sum (rss_s/c)^2, rho_crss = sum (crss_s,fwd/c)^2 with
 * c = 0.35*47.565*0.256, work_p = sum |gammatotal_s rss_s|.
 * crss: nsys (forward, backward) pairs.  out: rho_rss, rho_crss, work_p. */
void orc_rho_work_voxel(const double sig[9], const double *schm_gr, const double *crss, const double *gamtot,
                        int nsys, double out[3])
{
    double str[6], rss;
    t2 s33;
    int i;
    memcpy(s33, sig, sizeof(t2));
    out[0] = out[1] = out[2] = 0.0;
    chg_basis(str, s33, NULL, NULL, 2);
    for (i = 0; i < nsys; i++) {
        const double *sc = schm_gr + 5 * i;
        rss = sc[0] * str[0] + sc[1] * str[1] + sc[2] * str[2] + sc[3] * str[3] + sc[4] * str[4];
        out[0] += pow(rss / (0.35 * 47.565 * 0.256), 2);
        out[1] += pow(crss[2 * i] / (0.35 * 47.565 * 0.256), 2);
        out[2] += fabs((gamtot[i]) * rss);
    }
}

/* harden + rho/work diagnostics, constitutive.c:250-310 and evolution.c:509-529 */
static void harden_all(orc_job *J)
{
    int p;
#pragma omp parallel for schedule(static)
    for (p = 0; p < J->V; p++) {
        int jph = J->phase_f[p];
        if (J->nm.Type_phases[jph - 1]) continue;
        harden_voxel(J->crss[p], &J->gamacum[p], J->gammatotal[p], J->gamdot[p], J->nSYS[jph - 1],
                     J->Hard[jph - 1], J->tau[jph - 1], J->thet[jph - 1], J->nm.TimeStep);
        {
            double o3[3];
            orc_rho_work_voxel(&J->Sig[p][0][0], &J->Schm_gr[p][0][0], &J->crss[p][0][0], J->gammatotal[p],
                               J->nSYS[jph - 1], o3);
            J->rho_rss[p] = o3[0];
            J->rho_crss[p] = o3[1];
            J->work_p[p] = o3[2];
        }
    }
}

/* One increment of Evolution, evolution.c:297-530 */
int orc_increment(orc_job *J, orc_report *rep)
{
    int iter, mi, mj, mip, mjp, p, rc;
    double Err_e, Err_s, Err = J->nm.Err, dt = J->nm.TimeStep, Err_e_local, Err_s_local;
    double evmp, dvmp;
    t2 EpsAvg_local, EdotAvg_local;
    const int V = J->V;
    J->istep++;
    if (rep) memset(rep, 0, sizeof(*rep));
    /* predictor, evolution.c:306-318 */
    for (p = 0; p < V; p++)
        for (mi = 0; mi < 3; mi++)
            for (mj = 0; mj < 3; mj++) {
                J->VelGrad[p][mi][mj] = J->DisGrad[p][mi][mj];
                if (J->CREEP_FLAG == 1) J->DisGrad[p][mi][mj] = J->DisGradAvg_actual[mi][mj];
                else J->DisGrad[p][mi][mj] += J->Udot[mi][mj] * dt;
            }
    if (J->istep == 1 || J->nm.Update_Flag == 1) update_schmid(J);
    memset(J->dDisGradAvg, 0, sizeof(t2));
    memset(J->dDisGradAvg_acum, 0, sizeof(t2));
    iter = 0;
    Err_e = 2. * Err;
    Err_s = 2. * Err;
    while ((iter < J->nm.IterMax) && (fmax(fabs(Err_s), fabs(Err_e)) > fabs(Err))) {
        iter++;
        J->iter_total++;
        /* forward FFT of the 9 stress components, evolution.c:350-361 */
        for (mi = 0; mi < 3; mi++)
            for (mj = 0; mj < 3; mj++) {
                for (p = 0; p < V; p++) { J->fft[p].re = J->Sig[p][mi][mj]; J->fft[p].im = 0.0; }
                fft3d_cpx(J->fft, J->Nx, J->Ny, J->Nz, -1);
                for (p = 0; p < V; p++) { J->kSig_r[p][mi][mj] = J->fft[p].re; J->kSig_i[p][mi][mj] = J->fft[p].im; }
            }
        /* Green operator, evolution.c:363-377, with Gamma(k) of init.c:570-615 */
#pragma omp parallel for schedule(static)
        for (p = 0; p < V; p++) {
            int px = p / (J->Ny * J->Nz), py = (p / J->Nz) % J->Ny, pz = p % J->Nz, a, b;
            double G[3][3][3][3], n[3];
            t2 sr, si;
            if (p == 0) memset(G, 0, sizeof(G));
            else {
                orc_kvector(px, py, pz, J->Nx, J->Ny, J->Nz, n);
                orc_green(&J->C0[0][0][0][0], n, &G[0][0][0][0]);
            }
            for (a = 0; a < 3; a++)
                for (b = 0; b < 3; b++) {
                    sr[a][b] = 0.0;
                    si[a][b] = 0.0;
                    for (mip = 0; mip < 3; mip++)
                        for (mjp = 0; mjp < 3; mjp++) {
                            sr[a][b] += G[a][b][mip][mjp] * J->kSig_r[p][mip][mjp];
                            si[a][b] += G[a][b][mip][mjp] * J->kSig_i[p][mip][mjp];
                        }
                }
            memcpy(J->kSig_r[p], sr, sizeof(t2));
            memcpy(J->kSig_i[p], si, sizeof(t2));
        }
        /* inverse FFT and displacement-gradient update, evolution.c:383-394 */
        for (mi = 0; mi < 3; mi++)
            for (mj = 0; mj < 3; mj++) {
                for (p = 0; p < V; p++) { J->fft[p].re = J->kSig_r[p][mi][mj]; J->fft[p].im = J->kSig_i[p][mi][mj]; }
                fft3d_cpx(J->fft, J->Nx, J->Ny, J->Nz, +1);
                for (p = 0; p < V; p++) {
                    J->DisGrad[p][mi][mj] += J->dDisGradAvg[mi][mj] + J->fft[p].re / V;
                    J->fluctuation[p][mi][mj] += J->fft[p].re / V;
                }
            }
        Err_e_local = 0.0;
        Err_s_local = 0.0;
        rc = update_stress(J, &Err_e_local, &Err_s_local);
        if (rc) return rc;
        get_smacro(J);
        {
            double en[2];
            orc_err_normalise(Err_e_local, Err_s_local, J->WGT, J->e_vm, J->s_vm, en);
            Err_e = en[0];
            Err_s = en[1];
        }
        if (rep && iter <= ORC_MAXIT) {
            rep->err_e[iter - 1] = Err_e;
            rep->err_s[iter - 1] = Err_s;
            rep->svm_it[iter - 1] = J->s_vm;
        }
    }
    /* close-out, evolution.c:433-498 */
    for (p = 0; p < V; p++)
        for (mi = 0; mi < 3; mi++)
            for (mj = 0; mj < 3; mj++) J->VelGrad[p][mi][mj] = (J->DisGrad[p][mi][mj] - J->VelGrad[p][mi][mj]) / dt;
    for (mi = 0; mi < 3; mi++)
        for (mj = 0; mj < 3; mj++) {
            J->DisGradAvg_actual[mi][mj] = J->DisGradAvg[mi][mj] + J->dDisGradAvg_acum[mi][mj];
            J->VelGradAvg[mi][mj] = (J->DisGradAvg_actual[mi][mj] - J->DisGradAvg_t[mi][mj]) / dt;
        }
    J->e_vm = von_mises(J->DisGradAvg_actual);
    J->d_vm = von_mises(J->VelGradAvg);
    memcpy(J->DisGradAvg_t, J->DisGradAvg_actual, sizeof(t2));
    for (mi = 0; mi < 3; mi++)
        for (mj = 0; mj < 3; mj++) J->DisGradAvg[mi][mj] = J->DisGradAvg_t[mi][mj] + J->Udot[mi][mj] * dt;
    for (p = 0; p < V; p++)
        for (mi = 0; mi < 3; mi++)
            for (mj = 0; mj < 3; mj++) J->Eps[p][mi][mj] += J->Edot[p][mi][mj] * dt;
    memset(EpsAvg_local, 0, sizeof(t2));
    memset(EdotAvg_local, 0, sizeof(t2));
    for (p = 0; p < V; p++)
        for (mi = 0; mi < 3; mi++)
            for (mj = 0; mj < 3; mj++) {
                EpsAvg_local[mi][mj] += J->Eps[p][mi][mj];
                EdotAvg_local[mi][mj] += J->Edot[p][mi][mj];
            }
    for (mi = 0; mi < 3; mi++)
        for (mj = 0; mj < 3; mj++) {
            J->EpsAvg[mi][mj] = EpsAvg_local[mi][mj] * J->WGT;
            J->EdotAvg[mi][mj] = EdotAvg_local[mi][mj] * J->WGT;
        }
    evmp = 0.0;
    dvmp = 0.0;
    for (mi = 0; mi < 3; mi++)
        for (mj = 0; mj < 3; mj++) {
            evmp += J->EpsAvg[mi][mj] * J->EpsAvg[mi][mj];
            dvmp += J->EdotAvg[mi][mj] * J->EdotAvg[mi][mj];
        }
    evmp = sqrt(2. / 3. * evmp);
    dvmp = sqrt(2. / 3. * dvmp);
    if ((J->nm.Update_Flag == 1) && (J->istep > 1)) update_orient(J);
    if ((J->nm.Hard_Flag == 1) && (J->istep > 2)) harden_all(J);
    if (rep) {
        rep->istep = J->istep;
        rep->iter = iter;
        rep->iter_total = J->iter_total;
        rep->e_vm = J->e_vm; rep->evmp = evmp; rep->d_vm = J->d_vm; rep->dvmp = dvmp;
        rep->s_vm = J->s_vm; rep->s_vm1 = J->s_vm1;
        memcpy(rep->SigAvg, J->SigAvg, sizeof(t2));
        memcpy(rep->SigAvg1, J->SigAvg1, sizeof(t2));
        memcpy(rep->DisGradAvg_actual, J->DisGradAvg_actual, sizeof(t2));
        memcpy(rep->VelGradAvg, J->VelGradAvg, sizeof(t2));
        memcpy(rep->EpsAvg, J->EpsAvg, sizeof(t2));
        memcpy(rep->EdotAvg, J->EdotAvg, sizeof(t2));
    }
    return 0;
}

/* Field ids: include/voce.h VOCE_FIELD_* */
enum {
    F_SIG = 0, F_EPS, F_EDOT, F_ELS, F_DISGRAD, F_CRSS, F_CRSS_BWD, F_GAMTOT, F_GAMDOT, F_TAUX,
    F_TRANMAT, F_GAMACUM, F_SVM, F_EVM, F_WORK_P, F_RHO_RSS, F_RHO_CRSS, F_VELGRAD, F_SCHMID
};

static const int VI[6][2] = {{0, 0}, {1, 1}, {2, 2}, {1, 2}, {0, 2}, {0, 1}};

int orc_get_field(orc_job *J, int field, double *out)
{
    const long V = J->V;
    long p;
    int c;
    switch (field) {
    case F_SIG: case F_EPS: case F_EDOT:
        for (c = 0; c < 6; c++)
            for (p = 0; p < V; p++) {
                t2 *f = field == F_SIG ? J->Sig : (field == F_EPS ? J->Eps : J->Edot);
                out[c * V + p] = f[p][VI[c][0]][VI[c][1]];
            }
        return 6;
    case F_ELS: /* io.c:289-303 */
        for (c = 0; c < 6; c++)
            for (p = 0; p < V; p++) {
                int i = VI[c][0], j = VI[c][1];
                out[c * V + p] = (c < 3) ? J->DisGrad[p][i][i] - J->Eps[p][i][i]
                                         : (J->DisGrad[p][i][j] + J->DisGrad[p][j][i]) / 2.0 - J->Eps[p][i][j];
            }
        return 6;
    case F_DISGRAD: case F_TRANMAT: case F_VELGRAD:
        for (c = 0; c < 9; c++)
            for (p = 0; p < V; p++) {
                t2 *f = field == F_DISGRAD ? J->DisGrad : (field == F_TRANMAT ? J->TranMat : J->VelGrad);
                out[c * V + p] = f[p][c / 3][c % 3];
            }
        return 9;
    case F_CRSS: case F_CRSS_BWD:
        for (c = 0; c < 12; c++)
            for (p = 0; p < V; p++) out[c * V + p] = J->crss[p][c][field == F_CRSS ? 0 : 1];
        return 12;
    case F_GAMTOT: case F_GAMDOT: case F_TAUX:
        for (c = 0; c < 12; c++)
            for (p = 0; p < V; p++)
                out[c * V + p] = field == F_GAMTOT ? J->gammatotal[p][c] : (field == F_GAMDOT ? J->gamdot[p][c] : J->tau_x[p][c]);
        return 12;
    case F_GAMACUM: case F_SVM: case F_EVM: case F_WORK_P: case F_RHO_RSS: case F_RHO_CRSS: {
        double *f = field == F_GAMACUM ? J->gamacum : field == F_SVM ? J->SVM : field == F_EVM ? J->EVM
                  : field == F_WORK_P ? J->work_p : field == F_RHO_RSS ? J->rho_rss : J->rho_crss;
        memcpy(out, f, sizeof(double) * V);
        return 1;
    }
    case 19: /* Newton iterations of the last UpdateStress (diagnostic) */
        memcpy(out, J->nr_iters, sizeof(double) * V);
        return 1;
    case 20: /* work_e, WriteWorkEMPI io.c:231-278: the Voigt-ordered els through chg_basis opt 1 */
        for (p = 0; p < V; p++) {
            double els[6], w = 0.0;
            t2 els33;
            int mi, mj, i;
            for (i = 0; i < 6; i++) {
                if (i < 3) els[i] = J->DisGrad[p][i][i] - J->Eps[p][i][i];
                else if (i == 3) els[i] = (J->DisGrad[p][1][2] + J->DisGrad[p][2][1]) / 2.0 - J->Eps[p][1][2];
                else if (i == 4) els[i] = (J->DisGrad[p][0][2] + J->DisGrad[p][2][0]) / 2.0 - J->Eps[p][0][2];
                else els[i] = (J->DisGrad[p][0][1] + J->DisGrad[p][1][0]) / 2.0 - J->Eps[p][0][1];
            }
            chg_basis(els, els33, NULL, NULL, 1);
            for (mi = 0; mi < 3; mi++)
                for (mj = 0; mj < 3; mj++) w += 0.5 * J->Sig[p][mi][mj] * els33[mi][mj];
            out[p] = w;
        }
        return 1;
    case 21: { /* new_position, evolution.c:532-560: fft_data holds the forward
                  transform of fluctuation[.][mi][2] (the last mj) when the
                  Green step's store lands past the array, so the inverse gives
                  back V * fluctuation[.][mi][2] */
        cpx *d = (cpx *)malloc(sizeof(cpx) * V);
        int mi;
        for (mi = 0; mi < 3; mi++) {
            for (p = 0; p < V; p++) {
                d[p].re = J->fluctuation[p][mi][2];
                d[p].im = 0.0;
            }
            fft3d_cpx(d, J->Nx, J->Ny, J->Nz, -1);
            fft3d_cpx(d, J->Nx, J->Ny, J->Nz, +1);
            for (p = 0; p < V; p++) out[mi * V + p] = -d[p].re / (double)V;
        }
        free(d);
        return 3;
    }
    case 22: case 23: /* per-voxel error partials of the last UpdateStress, evolution.c:103-114
                         (22: |eps_tot - eps|, 23: |sig - lambda|); 0 for gas voxels */
        memcpy(out, field == 22 ? J->err_e_v : J->err_s_v, sizeof(double) * V);
        return 1;
    case F_SCHMID:
        for (c = 0; c < 60; c++)
            for (p = 0; p < V; p++) out[c * V + p] = J->Schm_gr[p][c / 5][c % 5];
        return 60;
    }
    return -1;
}

void orc_get_macro(orc_job *J, double *o)
{
    memcpy(o, J->DisGradAvg, 72);
    memcpy(o + 9, J->dDisGradAvg_acum, 72);
    memcpy(o + 18, J->Udot, 72);
    memcpy(o + 27, J->C066, 288);
    memcpy(o + 63, J->S066, 288);
    o[99] = J->Wgt_ph1;
    o[100] = J->e_vm;
    o[101] = J->d_vm;
}

static int write_doubles(const char *dir, const char *name, const double *v, long n)
{
    char path[1200];
    FILE *f;
    snprintf(path, sizeof(path), "%s/%s", dir, name);
    f = fopen(path, "wb");
    if (!f) return -1;
    fwrite(v, sizeof(double), n, f);
    fclose(f);
    return 0;
}

/* .iout writers, io.c:193-1002 (single rank, file offset 0) */
int orc_write_outputs(orc_job *J, const char *dir, int step, int tex_flag)
{
    const long V = J->V;
    double *buf = (double *)malloc(sizeof(double) * V * 12);
    char name[200];
    int c;
    static const char *pre6[4] = {"eps", "sig", "els", "edot"};
    static const int fid6[4] = {F_EPS, F_SIG, F_ELS, F_EDOT};
    int k;
    for (k = 0; k < 4; k++) {
        orc_get_field(J, fid6[k], buf);
        for (c = 0; c < 6; c++) {
            if (k == 2) snprintf(name, sizeof(name), "%s_%01d_S%06d.iout", pre6[k], c + 1, step);
            else snprintf(name, sizeof(name), "%s_%01d_S%04d.iout", pre6[k], c + 1, step);
            write_doubles(dir, name, buf + c * V, V);
        }
    }
    if (tex_flag) {
        static const char *recn[3] = {"slip_sys", "taux_sys", "crss_sys"};
        static const int recf[3] = {F_GAMTOT, F_TAUX, F_CRSS};
        unsigned char *rec = (unsigned char *)calloc(V, 104);
        long p;
        FILE *f;
        char path[1200];
        for (k = 0; k < 3; k++) {
            orc_get_field(J, recf[k], buf);
            memset(rec, 0, V * 104);
            for (p = 0; p < V; p++) {
                memcpy(rec + p * 104, &J->grain_f[p], 4);
                for (c = 0; c < 12; c++) memcpy(rec + p * 104 + 8 + 8 * c, &buf[c * V + p], 8);
            }
            snprintf(path, sizeof(path), "%s/%s_S%04d.iout", dir, recn[k], step);
            f = fopen(path, "wb");
            if (f) { fwrite(rec, 104, V, f); fclose(f); }
        }
        free(rec);
        /* tex, io.c:939-1002: 48-byte records */
        rec = (unsigned char *)calloc(V, 48);
        for (p = 0; p < V; p++) {
            double ang[3] = {0, 0, 0};
            int ci[6];
            t2 sa2xt;
            int mi, mj;
            for (mi = 0; mi < 3; mi++)
                for (mj = 0; mj < 3; mj++) sa2xt[mi][mj] = J->TranMat[p][mj][mi];
            orc_euler_to_mat(&ang[0], &ang[1], &ang[2], &sa2xt[0][0], 1);
            ci[0] = (int)(p / (J->Ny * J->Nz)) + 1;
            ci[1] = (int)((p / J->Nz) % J->Ny) + 1;
            ci[2] = (int)(p % J->Nz) + 1;
            ci[3] = J->grain_f[p];
            ci[4] = J->phase_f[p];
            ci[5] = 0;
            memcpy(rec + p * 48, ang, 24);
            memcpy(rec + p * 48 + 24, ci, 24);
        }
        snprintf(path, sizeof(path), "%s/tex_S%04d.iout", dir, step);
        f = fopen(path, "wb");
        if (f) { fwrite(rec, 48, V, f); fclose(f); }
        free(rec);
        snprintf(name, sizeof(name), "rho_rss_S%04d.iout", step); write_doubles(dir, name, J->rho_rss, V);
        snprintf(name, sizeof(name), "rho_crss_S%04d.iout", step); write_doubles(dir, name, J->rho_crss, V);
        snprintf(name, sizeof(name), "work_p_S%06d.iout", step); write_doubles(dir, name, J->work_p, V);
        snprintf(name, sizeof(name), "SVM_S%04d.iout", step); write_doubles(dir, name, J->SVM, V);
        snprintf(name, sizeof(name), "EVM_S%04d.iout", step); write_doubles(dir, name, J->EVM, V);
        /* new_position, io.c:805-849: 24-byte {3 double} records */
        {
            double *np3 = (double *)malloc(sizeof(double) * 3 * V), *rec3 = (double *)malloc(sizeof(double) * 3 * V);
            orc_get_field(J, 21, np3);
            for (p = 0; p < V; p++)
                for (c = 0; c < 3; c++) rec3[3 * p + c] = np3[c * V + p];
            snprintf(name, sizeof(name), "new_position_S%04d.iout", step);
            write_doubles(dir, name, rec3, 3 * V);
            free(np3);
            free(rec3);
        }
    }
    free(buf);
    return 0;
}
