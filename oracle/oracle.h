/*
 * oracle.h -- CPU restatement of the mesoOSU/Voce EVP-FFT solver (FP64).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * path is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library never
 * links, calls or falls back to it.
 *
 * The restatement follows the reference's own C sources function by
 * function (file:line citations at every definition in oracle.c):
 *   evolution.c:3-592  UpdateStress / get_smacro / VonMises / Evolution
 *   constitutive.c:4-310  StrainRate_eval / Orientation / update_orient / harden
 *   kinematics.c:5-459  chg_basis(5) / LU_* / Euler / Ten4thTransform / update_schmid
 *   init.c:9-828  stiffness, PlasticInit, InitMicroStruct, InitFields, CalcGAMMA, SetupJob
 *   io.c:54-122  GetNameList
 * The distributed FFT (FFTW 2.x-MPI, fftwnd_mpi) is replaced by a serial
 * mixed-radix complex DFT with the same convention: forward e^{-i}, both
 * directions unnormalised, x-slowest ("normal order") layout.
 *
 * Every voxel-wise loop may run under OpenMP; every sum that the reference
 * performs serially is still performed serially, in the reference's voxel
 * order, so results do not depend on the thread count.
 */
#ifndef VOCE_ORACLE_H
#define VOCE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXIT 1000   /* longest per-increment iteration log kept */

typedef struct orc_job orc_job;

typedef struct {
    int istep;          /* increment number, 1-based */
    int iter;           /* FFT iterations of this increment */
    int iter_total;     /* running total (last column of vm.out) */
    double e_vm, evmp, d_vm, dvmp, s_vm, s_vm1;     /* the vm.out row */
    double SigAvg[9], SigAvg1[9];                   /* row-major 3x3 */
    double DisGradAvg_actual[9], VelGradAvg[9];
    double EpsAvg[9], EdotAvg[9];
    double err_e[ORC_MAXIT], err_s[ORC_MAXIT], svm_it[ORC_MAXIT]; /* err.out rows */
} orc_report;

/* Setup (init.c:617-828).  input_path is a reference namelist; relative
 * file names inside it are opened relative to base_dir when base_dir is
 * non-NULL, else relative to the working directory (as the reference).
 * On failure returns NULL and stores the reference's abort code in *err
 * (7, 114, 116, 1139, 35, or -1 for an unreadable file). */
orc_job *orc_setup(const char *input_path, const char *base_dir, int *err);
void orc_free(orc_job *job);
int orc_nvox(const orc_job *job);
int orc_dims(const orc_job *job, int dims[3]);
int orc_nsteps(const orc_job *job);
void orc_set_threads(int n);

/* One strain increment (evolution.c:297-530).  Returns 0, or the abort code. */
int orc_increment(orc_job *job, orc_report *rep);

/* Copy a field, component-major (comp*V + voxel), voxel in reference
 * order.  Field ids are those of include/voce.h (VOCE_FIELD_*).  Returns
 * the component count, or -1 for an unknown field. */
int orc_get_field(orc_job *job, int field, double *out);
/* Macro quantities: out[0..8] DisGradAvg, [9..17] dDisGradAvg_acum,
 * [18..26] Udot, [27..62] C066, [63..98] S066, [99] Wgt_ph1, [100] e_vm, [101] d_vm */
void orc_get_macro(orc_job *job, double *out);

/* Writers with the reference's file names and byte layouts (io.c:193-1002). */
int orc_write_outputs(orc_job *job, const char *dir, int step, int tex_flag);

/* ---- unit-level entry points (same arithmetic as the job path) ---- */
void orc_chg_basis(double V6[6], double T2[9], double C2[36], double T4[81], int opt);
void orc_chg_basis5(double V5[5], double T2[9], double C2[36], double T4[81], int opt);
int orc_lu_inv66(double c[36]);                 /* 0, or 35 on a singular row */
void orc_euler_to_mat(double *t1, double *ph, double *t2, double a[9], int opt);
void orc_ten4th_transform(double c1[81], const double a[9], double c2[81], int opt);
void orc_orientation(double tran[9], double rot[9]);
void orc_strain_rate_eval(const double sig6[6], const double schm[60],
                          const double tau_t[24], const double xkin[12],
                          int nsys, const double nsr[12], const double gam0[12],
                          double edot[6], double d_edot[36],
                          double gamdot[12], double taux[12]);
/* Per-voxel Newton-Raphson of UpdateStress (evolution.c:28-157).  Inputs
 * are the reference's per-voxel arrays; outputs the post-NR Sig tensor,
 * Edot tensor, the two error partials, SVM, EVM and the NR count. */
int orc_update_stress_voxel(const double C_gr[36], const double C066[36],
                            const double Sig[9], const double Eps[9],
                            const double DisGrad[9], const double schm[60],
                            const double tau_t[24], int nsys, const double nsr[12],
                            const double gam0[12], double dt,
                            double Sig_out[9], double Edot_out[9],
                            double *err_e, double *err_s, double *svm, double *evm,
                            double gamdot[12], double taux[12]);
void orc_green(const double C0[81], const double n[3], double G[81]);
void orc_kvector(int i, int j, int k, int nx, int ny, int nz, double n[3]);
/* In-place unnormalised 3D complex DFT, sign -1 forward / +1 inverse. */
void orc_fft3d(double *re, double *im, int nx, int ny, int nz, int sign);
void orc_harden_voxel(double crss[24], double *gamacum, double gamtot[12],
                      const double gamdot[12], int nsys, const double hard[144],
                      const double tau[36], const double thet[24], double dt);
/* Latent-hardening matrix exactly as PlasticInit builds it (init.c:283-381). */
void orc_hard_matrix(double self, double coplane, double cross, double glissile,
                     double hirth, double lomer, double H[144]);
/* update_orient for one voxel (constitutive.c:188-229): tran updated in place. */
void orc_update_orient_voxel(double tran[9], const double velgrad[9], const double udot_a[9],
                             const double *bn, const double *bb, const double *gamdot, int nsys, double dt);
/* update_schmid for one voxel (kinematics.c:436-456): out = nsys x 5. */
void orc_update_schmid_voxel(const double tran[9], const double *schm_xt, int nsys, double *out);
/* get_smacro's mixed-BC strain correction (evolution.c:199-211). */
void orc_bc_correction(const double S0[81], const int vflag[6], const int sflag[6],
                       const double Scauchy[9], const double SigAvg[9], double dD[9]);
/* Err_e, Err_s from the voxel sums (evolution.c:410-423). */
void orc_err_normalise(double err_e_sum, double err_s_sum, double wgt, double e_vm, double s_vm,
                       double out[2]);
/* rho_rss, rho_crss, work_p of one voxel (evolution.c:509-529). */
void orc_rho_work_voxel(const double sig[9], const double *schm_gr, const double *crss,
                        const double *gamtot, int nsys, double out[3]);
/* Namelist parse only (io.c:54-122); returns number of keys with status 1,
 * and fills status[26] in the reference's key order. */
int orc_parse_namelist(const char *path, int status[26], double *vals_out, int nvals);

#ifdef __cplusplus
}
#endif
#endif
