/*
 * voce.h -- C ABI of the MI355X EVP-FFT crystal-plasticity solver.
 *
 * The reference (mesoOSU/Voce) has no library surface: main() calls three
 * void(void) functions over global state, SetupJob (init.c:617-828),
 * Evolution (evolution.c:265-622) and DestroyJob (init.c:830-930).  This
 * header replaces that boundary with a context handle:
 *
 *   voce_ctx_create   <- SetupJob minus file parsing (init.c:617-828)
 *   voce_increment    <- one `istep` body of Evolution (evolution.c:297-530)
 *   voce_get_field    <- the data gathering inside the Write*MPI writers
 *                        (io.c:193-1002)
 *   voce_ctx_destroy  <- DestroyJob (init.c:830-930), without its double free
 *
 * and keeps the reference's file formats on the host side:
 *
 *   voce_load_case    <- GetNameList (io.c:54-122) + PlasticInit's .sx parse
 *                        (init.c:51-187) + InitMicroStruct's mesh read
 *                        (init.c:393-423)
 *   voce_write_fields <- WriteEps/Sig/Els/Edot/SLIP/taux/CRSS/Texture/...MPI
 *                        (io.c:193-1002), same names and byte layouts
 *
 * Conventions.  Plain C types only.  The host owns configuration and output
 * buffers; the library owns device memory.  One context is used by one host
 * thread at a time and owns one HIP stream.  Failures are returned as error
 * codes (never exit/abort); positive codes are the reference's own abort
 * codes (7, 35, 114, 116, 1139), negative codes are this library's.
 * Units are those of the input (MPa, s).  Voxel order is the reference's:
 * index (x*Ny + y)*Nz + z, x slowest (evp.h:32-33).
 */
#ifndef VOCE_H
#define VOCE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VOCE_NSYS 12     /* evpFlags.h:5 NSYSMX */
#define VOCE_NPHASE 2    /* evpFlags.h:4 NPHMAX */

/* error codes */
#define VOCE_OK 0
#define VOCE_ERR_SLIP_NOT_ORTHOGONAL 7   /* init.c:213-215 */
#define VOCE_ERR_SINGULAR 35             /* kinematics.c:306-309 */
#define VOCE_ERR_BC 114                  /* init.c:732-735 */
#define VOCE_ERR_BC_DEVIATORIC 116       /* init.c:713-715 */
#define VOCE_ERR_MESH_INDEX 1139         /* init.c:420-423 */
#define VOCE_ERR_ARG (-1)
#define VOCE_ERR_FILE (-2)
#define VOCE_ERR_UNSUPPORTED_GRID (-3)   /* a dimension not 2^a 3^b 5^c 7^d in 2..512 */
#define VOCE_ERR_DEVICE (-4)
#define VOCE_ERR_NOMEM (-5)
#define VOCE_ERR_NOT_LOADED (-6)
#define VOCE_ERR_SCRATCH (-7)   /* RVE too large for voce_increment_noalloc's workspace */

/* precision modes */
#define VOCE_FP64 0   /* everything FP64 (drop-in path) */
#define VOCE_MIXED 1  /* FFT + Green operator FP32; Newton-Raphson, Voce, norms FP64 */

/* Per-phase material, as PlasticInit + ElastStiffnessMatrix hold it
 * (init.c:9-391).  Slip vectors are given as read from the .sx file; the
 * library normalises them and checks orthogonality like init.c:191-215. */
typedef struct voce_phase {
    int type;                   /* Type_phases: 0 crystal, 1 gas */
    double elast[3];            /* C11, C12, C44 */
    int nsys;                   /* <= VOCE_NSYS */
    double bn[VOCE_NSYS][3];    /* slip-plane normals (crystal frame) */
    double bb[VOCE_NSYS][3];    /* slip directions (crystal frame) */
    double nsrs;                /* rate-sensitivity exponent n */
    double gam0;                /* reference shear rate */
    double tau0f, tau0b, tau1, thet0, thet1;   /* Voce parameters */
    double hard[VOCE_NSYS][VOCE_NSYS];          /* latent-hardening matrix */
} voce_phase;

/* Everything SetupJob needs after parsing (init.c:617-828). */
typedef struct voce_setup {
    int dims[3];                /* CellDim */
    int n_phases;               /* N_phases (1 or 2) */
    voce_phase phase[VOCE_NPHASE];
    double time_step;           /* TimeStep */
    double err;                 /* Err */
    int iter_max;               /* IterMax */
    int update_flag;            /* Update_Flag: lattice rotation */
    int hard_flag;              /* Hard_Flag: Voce hardening */
    int velgrad_flag[6];        /* VelGrad_BC_Flag, Voigt 11,22,33,23,13,12 */
    double velgrad[6];          /* VelGrad_BC */
    int stress_flag[6];         /* Stress_BC_Flag */
    double stress[6];           /* Stress_BC */
    int use_full_velgrad;       /* 1: take velgrad_full (row-major 3x3, may be
                                   non-symmetric) instead of VoigtToFull(velgrad);
                                   no reference counterpart (kinematics.c:5-12
                                   can only build a symmetric L) */
    double velgrad_full[9];
    const double *euler_deg;    /* V x 3 Bunge angles (deg), mesh order */
    const int *grain;           /* V grain ids */
    const int *phase_id;        /* V phase ids, 1-based */
    int precision;              /* VOCE_FP64 or VOCE_MIXED */
} voce_setup;

/* One increment's results: the vm.out row (evolution.c:563-564), the
 * err.out rows (evolution.c:427) and the macro tensors (row-major 3x3). */
typedef struct voce_incr_report {
    int istep;                  /* 1-based increment number */
    int iter;                   /* FFT iterations of this increment */
    int iter_total;             /* running total, last vm.out column */
    double e_vm, evmp, d_vm, dvmp, s_vm, s_vm1;
    double sig_avg[9], sig_avg1[9];
    double disgrad_avg[9], velgrad_avg[9];
    double eps_avg[9], edot_avg[9];
    const double *err_e;        /* [iter] per-iteration Err_e, Err_s, s_vm; */
    const double *err_s;        /* owned by the context, valid until the   */
    const double *svm_it;       /* next voce_increment on it               */
} voce_incr_report;

/* Field ids for voce_get_field.  Buffers are component-major: element
 * (c, v) at host[c*V + v], voxel v in reference order. */
enum voce_field {
    VOCE_FIELD_SIG = 0,      /* 6: Sig 11,22,33,23,13,12 */
    VOCE_FIELD_EPS = 1,      /* 6: Eps (plastic strain) */
    VOCE_FIELD_EDOT = 2,     /* 6: Edot */
    VOCE_FIELD_ELS = 3,      /* 6: sym(DisGrad) - Eps (io.c:289-303) */
    VOCE_FIELD_DISGRAD = 4,  /* 9: DisGrad row-major */
    VOCE_FIELD_CRSS = 5,     /* 12: crss[.][s][0] */
    VOCE_FIELD_CRSS_BWD = 6, /* 12: crss[.][s][1] */
    VOCE_FIELD_GAMTOT = 7,   /* 12: gammatotal */
    VOCE_FIELD_GAMDOT = 8,   /* 12: gamdot of the final NR evaluation */
    VOCE_FIELD_TAUX = 9,     /* 12: tau_x of the final NR evaluation */
    VOCE_FIELD_TRANMAT = 10, /* 9: TranMat_xt2sa (crystal -> sample) */
    VOCE_FIELD_GAMACUM = 11, /* 1: accumulated shear */
    VOCE_FIELD_SVM = 12,     /* 1 */
    VOCE_FIELD_EVM = 13,     /* 1 */
    VOCE_FIELD_WORK_P = 14,  /* 1 (evolution.c:527) */
    VOCE_FIELD_RHO_RSS = 15, /* 1 (evolution.c:523) */
    VOCE_FIELD_RHO_CRSS = 16,/* 1 (evolution.c:524) */
    VOCE_FIELD_VELGRAD = 17, /* 9: skew part only (the part the solver uses) */
    VOCE_FIELD_SCHMID = 18,  /* 60: Schm_gr[.][s][0..4] of the current TranMat
                                (update_schmid, kinematics.c:426-459): the
                                reference's Schm_gr after the next increment's
                                refresh (evolution.c:320-322) */
    VOCE_FIELD_WORK_E = 19,  /* 1: 0.5 Sig : els33, els in the order above read as
                                B-basis coefficients (WriteWorkEMPI, io.c:231-278) */
    VOCE_FIELD_NEW_POSITION = 20, /* 3: -fluctuation[.][i][2], i = 0..2: what the
                                displacement block writes as new_position
                                (evolution.c:532-560); tracked while diagnostics
                                are on, 0 otherwise */
    VOCE_FIELD_COUNT = 21
};

typedef struct voce_ctx voce_ctx;

/* Number of components of a field (or -1). */
int voce_field_ncomp(int field);

/* SetupJob (init.c:617-828) without the parsing: material constants,
 * orientations, C0/S0, Schmid tensors, initial stress and CRSS, all placed
 * in device memory.  Returns VOCE_OK or an error code; *out is NULL on
 * failure. */
int voce_ctx_create(const voce_setup *setup, voce_ctx **out);

/* One strain increment, evolution.c:297-530: predictor, FFT fixed-point
 * iterations to max(Err_s, Err_e) <= Err or IterMax, close-out (plastic
 * strain, lattice rotation from step 2, hardening from step 3).  All on
 * the device; the report is copied back at the end. */
int voce_increment(voce_ctx *ctx, voce_incr_report *report);

/* voce_increment without allocating: no counterpart in the reference.  For
 * callers that cannot allocate device memory during a solve (tight serving
 * loops, code that runs between HIP graph captures, harnesses that forbid
 * hipMalloc): the increment's workspace (pre-update stress, crystal-frame
 * K_c, spectra, skew sums, block partials) comes from one static device
 * array of VOCE_NOALLOC_MAX_BYTES per device instead of the context's own,
 * and the call makes no hipMalloc/hipFree (it only needs the context, which
 * voce_ctx_create has built).  Results are bit-identical to voce_increment.
 * Limits: a context of voce_ctx_create (not a batch member or server id)
 * with IterMax >= 1, whose workspace fits the array (one 64^3 RVE in FP64,
 * about 61 MiB; larger RVEs return VOCE_ERR_SCRATCH).  Every call on one
 * device shares that device's array: calls that overlap on a device race. */
#define VOCE_NOALLOC_MAX_BYTES (96u << 20)
int voce_increment_noalloc(voce_ctx *ctx, voce_incr_report *report);

/* Copy one field to host memory (bytes >= ncomp*V*sizeof(double)). */
int voce_get_field(voce_ctx *ctx, int field, double *host, size_t bytes);

/* Request the per-voxel diagnostics of evolution.c:134-155 and 509-529
 * (SVM, EVM, gamdot, tau_x, rho_rss, rho_crss, work_p) to be produced by
 * every following increment (0 turns them off; default off). */
int voce_set_diagnostics(voce_ctx *ctx, int on);

/* Number of voxels / grid dims of a context. */
long voce_ctx_nvox(const voce_ctx *ctx);

/* DestroyJob (init.c:830-930).  Safe on NULL. */
void voce_ctx_destroy(voce_ctx *ctx);

/* Human-readable text for an error code. */
const char *voce_strerror(int code);

/* ---- host-side file formats (no device work) ---- */

/* A parsed case: the namelist, its material files and its mesh. */
typedef struct voce_case {
    voce_setup setup;           /* euler/grain/phase point into the arrays below */
    int n_steps;                /* N_steps */
    int print_control[2];       /* PrintControl */
    int tex_flag;               /* Tex_Flag */
    int namelist_status[26];    /* per-key status as GetNameList leaves it
                                   (0 missing, 1 ok, 2 missing data,
                                    3 extra data, 4 multiply defined) */
    double *euler_deg;          /* V x 3 (owned) */
    int *grain;                 /* V (owned) */
    int *phase_id;              /* V (owned) */
} voce_case;

/* Read namelist + .sx + mesh (io.c:54-122, init.c:51-187, init.c:393-423).
 * Relative file names in the namelist are opened relative to base_dir when
 * it is non-NULL, else relative to the working directory (as the
 * reference).  Fills *out; free with voce_case_free. */
int voce_load_case(const char *input_path, const char *base_dir, voce_case *out);
void voce_case_free(voce_case *c);

/* Write the field dumps of evolution.c:569-590 / 594-611 into dir with the
 * reference's names and byte layouts: eps_/sig_/edot_%01d_S%04d.iout,
 * els_%01d_S%06d.iout and, with tex_flag, slip_sys/taux_sys/crss_sys
 * (104-byte records), tex (48-byte records), rho_rss, rho_crss, work_p,
 * SVM, EVM.  `prefix_final` selects the *_final names used after the last
 * increment (evolution.c:594-611). */
int voce_write_fields(voce_ctx *ctx, const char *dir, int step, int tex_flag, int prefix_final);

#ifdef __cplusplus
}
#endif
#endif
