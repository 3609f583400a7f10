// voce_internal.h -- device data layout and kernel launchers shared by the
// host context code (voce_ctx.cpp) and the HIP kernels (voce_kernels.hip).
//
// Layout in HBM for one RVE with V = Nx*Ny*Nz voxels, voxel index
// v = (x*Ny + y)*Nz + z (the reference's order, evp.h:32-33).  Every field
// is structure-of-arrays, component-major (component c at base + c*V):
//
//   E6     6 x V f64   sym(DisGrad) in the Lebensohn-Tome B basis
//   S6     6 x V f64   Cauchy stress, B basis
//   P6     6 x V f64   stress at which the last NR flow-rule evaluation ran
//                      (the "pre-update" stress; Edot and gamdot follow from
//                      it exactly, evolution.c:67,125-133), in the crystal
//                      frame of TranMat R
//   EP5    5 x V f64   plastic strain, B basis (the 6th component is 0)
//   ED5    5 x V f64   plastic strain rate of the increment, B basis
//   CRSS  12 x V f64   forward CRSS; a second 12 x V block holds the
//                      backward CRSS only when tau0f != tau0b
//   R      9 x V f64   TranMat_xt2sa (crystal -> sample), row-major
//   WG     3 x V f64   skew(DisGrad): components (2,1), (0,2), (1,0)
//   WF     3 x V f64   skew(DisGrad increment) of the last increment (its
//                      Green-operator part until the close-out completes it)
//   GAC      V f64     accumulated shear gamacum;  GTOT 12 x V gammatotal
//   ORI      V u32     index of the voxel's initial orientation in the
//                      per-orientation table (S_local, K_c = C066^-1 +
//                      S_local, initial TranMat; 0xffffffff = gas)
//   PH       V u8      phase id (1-based)
//
// FFT work buffers (half spectrum along z, NZH = Nz/2 + 1, complex of the
// FFT type T: double2 in FP64 mode, float2 in MIXED mode).  A (ky, kz)
// pair is a column col = ky*NZH + kz; columns are grouped in blocks of
// TB = 128 bytes (8 double2 or 16 float2), NCB = ceil(Ny*NZH / TB):
//   SPEC  6 x NSPEC   component c, plane kx, column col at
//                     c*NSPEC + ((col/TB)*Nx + kx)*TB + col%TB,
//                     NSPEC = NCB*TB*Nx (whole cache lines for the plane
//                     kernels, which walk columns, and for the x pass,
//                     which walks kx)
//   USUM  3 x NSPEC   same layout: sum over the increment's iterations of
//                     the skew part of the Green update, components
//                     (2,1), (0,2), (1,0) (lattice-spin recovery)
#pragma once
#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime.h>   // int2

#define VOCE_GAS 0xffffffffu
#define VOCE_OTAB 81    // doubles per orientation-table row

struct PhaseConst {
    int nsys;
    int gas;
    int two_crss;   // 1 when forward and backward CRSS are stored separately
    int int_n;      // 1 when nsrs is an integer in [1, 64]
    int fcc12;      // 1 when the twelve Schmid vectors are those of the {111}<110>
                    // systems in the order and signs of Cu_p.sx (voce_fcc_schmid)
    int pad_;
    double nsr, gam0, tau0f, tau0b, tau1, thet0, thet1;
    double px[12][5];      // crystal-frame Schmid vectors Schm_xt (init.c:217-224)
    double p3[12][6];      // the same as symmetric tensors: 00 11 22 12 02 01
    double bn[12][3], bb[12][3];   // slip normal and direction whose product is px
    double rn[12][3], rb[12][3];   // the slip vectors update_orient rotates (constitutive.c:202-203)
    double hard[12][12];
    // factors of the Walsh form of the flow rule (fcc12 only): edot scale
    // gam0/(2 sqrt 3), gam0/2; Newton-matrix scale dt gam0 n times 1/12,
    // 1/(4 sqrt 3), 1/4
    double wk[5];
};

// Macro (RVE-wide) state, resident in device memory.  Updated by the
// single-block finaliser kernels; read with uniform loads by the rest.
struct MacroState {
    // constants
    double dt, err_tol, wgt, wgt_over_ph1;   // wgt_over_ph1 = WGT/Wgt_ph1
    double C066[36];
    double S066[36];        // C066^-1 (B basis)
    double S0[81];
    double C0[81];
    double G36[36];         // K(n) = C0_ijkl n_j n_l as a quadratic form (Green operator)
    float G36f[36];
    double Udot[9], Udot_a[9], Scauchy[9];
    double udot_sym6[6];    // B basis of sym(Udot)
    int velgrad_flag[6], stress_flag[6];
    int iter_max, creep, update_flag, hard_flag;
    int n_phases, diag, two_crss_any, pad0;
    // evolving
    double DisGradAvg[9], DisGradAvg_t[9], DisGradAvg_actual[9];
    double dD[9], acum[9], applied[9];
    double SigAvg[9], SigAvg1[9];
    double EpsAvg[9], EdotAvg[9], VelGradAvg[9];
    double e_vm, d_vm, s_vm, s_vm1, Err_e, Err_s, evmp, dvmp;
    double shift6[6];        // uniform B-basis strain added by the next E update
    double creep_sym6[6];    // creep mode: uniform sym(DisGrad) at predictor
    double creep_skew[3];    // creep mode: uniform skew(DisGrad) at predictor
    int istep, iter, iter_total, active;
    int error, fin_pending, log_cap, vg_pending;   // fin_pending: Newton partials not yet finalised
    // per-iteration Err_e, Err_s, s_vm of the running increment: 3 x log_cap
    double *logs;
    // velocity gradient for the next increment (taken by k_begin_increment
    // when vg_pending is set; a request-driven caller sets it per increment)
    double vg_next[9];
};

// Device pointers of one RVE.
struct RveDev {
    int nx, ny, nz, nzh;
    long V, nspec;               // voxels, complex elements of one spectrum component (padded)
    double *E6, *S6, *P6, *EP5, *ED5, *CRSS, *CRSSB, *R, *WG, *WF, *GAC, *GTOT;
    double *GAMDOT, *TAUX, *SVM, *EVM, *WORKP, *RHOR, *RHOC;   // diagnostics (may be null)
    // diagnostics: running sums of the Green-operator updates (the reference's
    // `fluctuation`, evolution.c:392) that new_position reads: B-basis
    // components 2, 3, 4, 6 of the symmetric part, then skew (2,1) and (0,2)
    double *FLU;
    // workspace: the deviatoric 5x5 block of K_c = C066^-1 + S_local of the
    // voxel's orientation-table row in the crystal frame of its TranMat
    // (Q^T K_c Q, Q the B-basis rotation), packed lower triangle, written by
    // the first Newton launch of each increment (TranMat is fixed until the
    // close-out).  Tiles of 64 voxels, component-major inside a tile:
    // component q of voxel v at ((v / 64) * 15 + q) * 64 + v % 64
    // (VOCE_KCV_COMP * voce_kcv_voxels(V) doubles)
    double *KCV;
    uint32_t *ORI;
    uint8_t *PH;
    double *SLOC;                // per orientation, VOCE_OTAB doubles: S_local = C_gr^-1
                                 // and K_c = C066^-1 + S_local (36 each, row major),
                                 // then the initial TranMat (9)
    PhaseConst *phase;           // [2]
    MacroState *ms;
    void *spec, *usum;           // complex T
    double *part;                // per-block partials
    int nblk_vox;                // blocks of the voxel kernels
    void *tw_x, *tw_y, *tw_z;    // twiddle tables, complex T, length N, exp(-2 pi i j/N)
    int fft_f32;                 // 1: FFT in FP32
    int uniform_phase;           // 1: one crystalline phase (its constants are uniform)
    int two_crss;                // 1: some phase keeps a separate backward CRSS
    // FFT launch choices (tuned per grid size): plane-kernel variant plv
    // (0..4) and x-pass variant xv (0..3, see voce_kernels.hip), x-pass tile tk
    // ((y,kz) columns per block), radix-8 stages allowed in the plane and
    // x passes
    int plv, xv, tk, pl_r8, x_r8;
    int ncb;                     // column blocks of the spectrum layout (see voce_spec_layout)
};

// What one launch sequence covers: a device table of RVE descriptors and,
// for a batch, per-kernel maps block -> (RVE, local block).  A single RVE
// passes a one-entry table and null maps.  Batch-wide template choices:
// fft_f32 must agree across the batch; uniform_phase is the AND and
// two_crss the OR of the members'.
struct LaunchSet {
    const RveDev *tab;
    // map_pfwd: the forward plane blocks plus one finaliser block per RVE
    // (local block 6 nx), which finishes the previous iteration's Newton step
    const int2 *map_pfwd, *map_plane6, *map_plane3, *map_x, *map_vox, *map_one;
    int n_pfwd, n_plane6, n_plane3, n_x, n_vox, n_one;
    int plv, xv;                 // kernel variants (same for every RVE of the set)
    int p2_n;                    // plv 4: the square plane size ny = nz of the set
    int p2_nx, p2_tk;            // xv 3: nx and the x-pass tile of the set
    int sp_ny, sp_nz;            // plv 5: ny and nz of the set (chunks of the split passes)
    size_t pl_lds, x_lds;        // dynamic LDS: the largest over the set's RVEs
    int fft_f32, uniform_phase, two_crss;
};

// voxels the KCV array is allocated for (whole tiles of 64), and doubles per voxel
static inline long voce_kcv_voxels(long V) { return (V + 63) & ~63L; }
#ifndef VOCE_KCV_COMP
#define VOCE_KCV_COMP 15
#endif

#ifdef __cplusplus
extern "C" {
#endif
// host-side launchers (voce_kernels.hip); each returns a hipError_t as int
int voce_launch_begin_increment(const LaunchSet *L, void *stream);
int voce_launch_iteration(const LaunchSet *L, void *stream, void *events /* hipEvent_t[6] or NULL */);
int voce_launch_closeout(const LaunchSet *L, void *stream, int need_skew);
int voce_launch_fft_probe(const LaunchSet *L, void *stream, int which);
int voce_launch_count_active(const RveDev *tab, int n, int *out_dev, void *stream);
int voce_launch_finalize(const LaunchSet *L, void *stream);   // pending iteration finalisers (one block per RVE)
int voce_launch_poll(const RveDev *tab, const int *idx_dev, int n, int *flags_dev, void *stream);
int voce_launch_init_fields(const RveDev *tab1, long V, const double *sig6_by_orient, void *stream);
int voce_grid_supported(int n);
#define VOCE_N_PLV 7                // plane-kernel variants (4: line-parallel square planes,
                                    // 5: split z and y passes, 6: line-parallel split passes, 256 x 256)
// B-basis component a of the crystal-frame Schmid vector of system s of the
// twelve {111}<110> systems in the order and signs of the reference's
// Cu_p.sx.  System s = 3p + j is direction j of plane p; over the four planes
// each (a, j) component is kappa_aj times a Walsh row, (-1)^popcount(w_aj & p),
// with kappa one of +-1/(2 sqrt 3), -1/sqrt 3, -1/2 or 0.
static inline double voce_fcc_schmid(int s, int a)
{
    static const int W[5][3] = {{3, 0, 0}, {3, 0, 0}, {0, 1, 1}, {1, 0, 2}, {0, 3, 0}};
    static const double K[5][3] = {{1, -1, -2}, {-1, -1, 0}, {0, 1, 1}, {1, 0, -1}, {1, 1, 0}};
    const int p = s / 3, j = s % 3, w = W[a][j] & p;
    const double unit = (a == 1) ? 0.5 : 0.5 / sqrt(3.0);
    return ((w == 1 || w == 2) ? -1.0 : 1.0) * K[a][j] * unit;
}
#define VOCE_N_XV 5                 // x-pass variants (3: line-parallel, Nx 16, 32, 64; 4: staged through LDS, Nx 16, 32, 48, 64, 128, 256)
int voce_kernels_prepare(void);
int voce_xgreen_tile(int nx, int xv);
int voce_vox_block(void);   // threads per block of the per-voxel kernels
size_t voce_plane_lds(const RveDev *r, int plv);            // 0: does not fit
int voce_default_plv(const RveDev *d);                      // -1: no plane variant fits
size_t voce_xgreen_lds(const RveDev *r, int xv, int tk);    // 0: does not fit
void voce_spec_layout(RveDev *r);   // sets ncb and nspec from the grid and precision
#ifdef __cplusplus
}
#endif
