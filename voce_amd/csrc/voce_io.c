/*
 * voce_io.c -- the reference's input formats, read on the host.
 *
 *   namelist  GetNameList (io.c:54-122): 80-byte fgets records; lines
 *             starting with '#' or '\n' are skipped; a line holding only
 *             separators ends the parse; the read that reaches end of file
 *             ends the parse *before* its (unterminated) line is used; keys
 *             are case-sensitive; values are separated by ", \t\n"; a key
 *             seen twice keeps its first values.
 *   .sx       PlasticInit (init.c:51-187): fixed line positions, one block
 *             per listed mode, 24 lines skipped for an unused mode.
 *   mesh      InitMicroStruct (init.c:393-423): one 80-byte record per
 *             voxel, "phi1 Phi phi2 i j k grain phase", x slowest; an index
 *             that disagrees with the record's position is abort 1139.
 *   latent    PlasticInit's 12x12 table (init.c:283-381).
 *
 * voce_load_case_voronoi takes the mesh from voce_voronoi (voce_gen.cpp)
 * instead of a file.
 */
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/voce.h"
#include "../../include/voce_batch.h"

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
} namelist_t;

enum { K_INT, K_REAL, K_STR };
typedef struct {
    const char *name;
    size_t off;
    int type;
    int len;
} key_t_;

#define KINT(x, n) {#x, offsetof(namelist_t, x), K_INT, n}
#define KREAL(x, n) {#x, offsetof(namelist_t, x), K_REAL, n}
#define KSTR(x) {#x, offsetof(namelist_t, x), K_STR, 1}

/* io.c:25-52, same order */
static const key_t_ KEYS[26] = {
    KINT(CellDim, 3), KINT(N_phases, 1), KINT(Type_phases, 2), KREAL(TimeStep, 1), KINT(N_steps, 1),
    KINT(Update_Flag, 1), KINT(Hard_Flag, 1), KINT(Tex_Flag, 1), KINT(IterMax, 1), KINT(PrintControl, 2),
    KREAL(Err, 1), KINT(VelGrad_BC_Flag, 6), KREAL(VelGrad_BC, 6), KINT(Stress_BC_Flag, 6),
    KREAL(Stress_BC, 6), KREAL(ElastConst_PhaseI, 4), KREAL(ElastConst_PhaseII, 4), KSTR(Slip_PhaseI),
    KSTR(Slip_PhaseII), KSTR(initial_ms), KREAL(self, 2), KREAL(coplane, 2), KREAL(cross, 2),
    KREAL(glissile, 2), KREAL(hirth, 2), KREAL(lomer, 2)};

static int read_namelist(const char *path, namelist_t *nm, int status[26])
{
    char line[80];
    FILE *fp = fopen(path, "r");
    int k;
    memset(nm, 0, sizeof(*nm));
    for (k = 0; k < 26; ++k) status[k] = 0;
    if (!fp) return VOCE_ERR_FILE;
    for (;;) {
        char *tok;
        if (!fgets(line, sizeof(line), fp)) { /* the buffer keeps its bytes; feof ends */ }
        if (feof(fp)) break;
        if (line[0] == '#' || line[0] == '\n') continue;
        tok = strtok(line, " \t\n");
        if (!tok) break;
        for (k = 0; k < 26; ++k) {
            int j;
            if (strcmp(tok, KEYS[k].name) != 0) continue;
            if (status[k] != 0) {
                status[k] = 4;
                break;
            }
            status[k] = 1;
            for (j = 0; j < KEYS[k].len; ++j) {
                char *v = strtok(NULL, ", \t\n");
                char *base = (char *)nm + KEYS[k].off;
                if (!v) {
                    status[k] = 2;
                    continue;
                }
                if (KEYS[k].type == K_INT) ((int *)base)[j] = (int)atol(v);
                else if (KEYS[k].type == K_REAL) ((double *)base)[j] = atof(v);
                else { strncpy(base, v, 99); base[99] = 0; }
            }
            if (strtok(NULL, ", \t\n")) status[k] = 3;
            break;
        }
    }
    fclose(fp);
    return 0;
}

static void join(char *out, size_t n, const char *base, const char *name)
{
    if (base && name[0] != '/') snprintf(out, n, "%s/%s", base, name);
    else snprintf(out, n, "%s", name);
}

/* init.c:283-381 */
static void latent_table(double S, double C, double X, double G, double Hi, double L, double H[12][12])
{
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
    int i, j;
    for (i = 0; i < 12; ++i)
        for (j = 0; j < 12; ++j) H[i][j] = j >= i ? up[i][j] : up[j][i];
}

/* PlasticInit's parse (init.c:65-187).  Slip vectors are returned as read;
 * voce_ctx_create normalises them (init.c:191-215). */
static int read_material(const char *path, voce_phase *ph)
{
    char buf[80] = {0};
    int i, j, mx = 0, nmax = 0, nmodes = 0, imode[64];
    int nsmx = 0, isect = 0;
    double nrsx = 0, gamd0x = 0, twsh = 0, t0f = 0, t0b = 0, t1 = 0, th0 = 0, th1 = 0, hs = 0, hl = 0;
    FILE *fp = fopen(path, "r");
    if (!fp) return VOCE_ERR_FILE;
#define LINE() (void)!fgets(buf, sizeof(buf), fp)
    LINE();
    LINE();
    LINE(); LINE();
    LINE();
    LINE(); LINE();
    LINE(); sscanf(buf, "%d", &nmax);
    LINE(); LINE();
    LINE(); sscanf(buf, "%d", &nmodes);
    if (nmodes > 64) nmodes = 64;
    LINE(); LINE();
    for (i = 0; i < nmodes; ++i) { LINE(); sscanf(buf, "%d", &imode[i]); }
    LINE();
    for (i = 0; i < nmax; ++i) {
        int used = 0;
        LINE();
        sscanf(buf, "%*s %*s %d %*s", &mx);
        for (j = 0; j < nmodes; ++j)
            if (mx == imode[j]) { used = 1; break; }
        if (!used) {
            for (j = 0; j < 24; ++j) LINE();
            continue;
        }
        LINE(); LINE();
        sscanf(buf, "%d %lf %lf %lf %d", &nsmx, &nrsx, &gamd0x, &twsh, &isect);
        if (nsmx > VOCE_NSYS || nsmx < 0) { fclose(fp); return VOCE_ERR_FILE; }
        LINE(); LINE();
        sscanf(buf, "%lf %lf %lf %lf %lf", &t0f, &t0b, &t1, &th0, &th1);
        LINE(); LINE();
        sscanf(buf, "%lf %lf", &hs, &hl);
        LINE();
        ph->nsys = nsmx;
        ph->nsrs = nrsx;
        ph->gam0 = gamd0x;
        ph->tau0f = t0f;
        ph->tau0b = t0b;
        ph->tau1 = t1;
        ph->thet0 = th0;
        ph->thet1 = th1;
        for (j = 0; j < nsmx; ++j) {
            LINE();
            sscanf(buf, "%lf %lf %lf %lf %lf %lf", &ph->bn[j][0], &ph->bn[j][1], &ph->bn[j][2], &ph->bb[j][0],
                   &ph->bb[j][1], &ph->bb[j][2]);
        }
    }
#undef LINE
    fclose(fp);
    (void)twsh; (void)isect; (void)hs; (void)hl;
    return 0;
}

/* InitMicroStruct's read (init.c:405-427) */
static int read_mesh(const char *path, const int dims[3], double *eul, int *grain, int *phase)
{
    char buf[80];
    long v = 0;
    const long V = (long)dims[0] * dims[1] * dims[2];
    FILE *fp = fopen(path, "r");
    if (!fp) return VOCE_ERR_FILE;
    for (v = 0; v < V; ++v) {
        double a, b, c;
        int ii, jj, kk, g, p;
        long idx;
        if (!fgets(buf, sizeof(buf), fp)) { fclose(fp); return VOCE_ERR_MESH_INDEX; }
        if (sscanf(buf, "%lf %lf %lf %d %d %d %d %d", &a, &b, &c, &ii, &jj, &kk, &g, &p) != 8) {
            fclose(fp);
            return VOCE_ERR_MESH_INDEX;
        }
        idx = ((long)(ii - 1) * dims[1] + jj - 1) * dims[2] + kk - 1;
        if (idx != v) { fclose(fp); return VOCE_ERR_MESH_INDEX; }
        eul[3 * v] = a;
        eul[3 * v + 1] = b;
        eul[3 * v + 2] = c;
        grain[v] = g;
        phase[v] = p;
    }
    fclose(fp);
    return 0;
}

void voce_case_free(voce_case *c)
{
    if (!c) return;
    free(c->euler_deg);
    free(c->grain);
    free(c->phase_id);
    c->euler_deg = NULL;
    c->grain = NULL;
    c->phase_id = NULL;
}

static int load_case_impl(const char *input_path, const char *base_dir, const int *dims_override, int n_grains,
                          uint64_t seed, voce_case *out);

int voce_load_case(const char *input_path, const char *base_dir, voce_case *out)
{
    return load_case_impl(input_path, base_dir, NULL, 0, 0, out);
}

int voce_load_case_voronoi(const char *input_path, const char *base_dir, const int dims[3], int n_grains,
                           uint64_t seed, voce_case *out)
{
    if (n_grains <= 0) return VOCE_ERR_ARG;
    return load_case_impl(input_path, base_dir, dims, n_grains, seed, out);
}

static int load_case_impl(const char *input_path, const char *base_dir, const int *dims_override, int n_grains,
                          uint64_t seed, voce_case *out)
{
    namelist_t nm;
    char path[1200];
    long V;
    int p, rc;
    voce_setup *s;
    if (!input_path || !out) return VOCE_ERR_ARG;
    memset(out, 0, sizeof(*out));
    if ((rc = read_namelist(input_path, &nm, out->namelist_status))) return rc;
    s = &out->setup;
    memcpy(s->dims, dims_override ? dims_override : nm.CellDim, sizeof(s->dims));
    if (s->dims[0] <= 0 || s->dims[1] <= 0 || s->dims[2] <= 0) return VOCE_ERR_FILE;
    if (nm.N_phases < 1 || nm.N_phases > VOCE_NPHASE) return VOCE_ERR_FILE;
    s->n_phases = nm.N_phases;
    s->time_step = nm.TimeStep;
    s->err = nm.Err;
    s->iter_max = nm.IterMax;
    s->update_flag = nm.Update_Flag;
    s->hard_flag = nm.Hard_Flag;
    memcpy(s->velgrad_flag, nm.VelGrad_BC_Flag, sizeof(s->velgrad_flag));
    memcpy(s->velgrad, nm.VelGrad_BC, sizeof(s->velgrad));
    memcpy(s->stress_flag, nm.Stress_BC_Flag, sizeof(s->stress_flag));
    memcpy(s->stress, nm.Stress_BC, sizeof(s->stress));
    s->precision = VOCE_FP64;
    out->n_steps = nm.N_steps;
    out->print_control[0] = nm.PrintControl[0];
    out->print_control[1] = nm.PrintControl[1];
    out->tex_flag = nm.Tex_Flag;
    for (p = 0; p < nm.N_phases; ++p) {
        voce_phase *ph = &s->phase[p];
        const double *ec = p == 0 ? nm.ElastConst_PhaseI : nm.ElastConst_PhaseII;
        ph->type = nm.Type_phases[p];
        ph->elast[0] = ec[0];
        ph->elast[1] = ec[1];
        ph->elast[2] = ec[2];
        if (ph->type == 0) {
            join(path, sizeof(path), base_dir, p == 0 ? nm.Slip_PhaseI : nm.Slip_PhaseII);
            if ((rc = read_material(path, ph))) return rc;
            latent_table(nm.self[p], nm.coplane[p], nm.cross[p], nm.glissile[p], nm.hirth[p], nm.lomer[p], ph->hard);
        }
    }
    V = (long)s->dims[0] * s->dims[1] * s->dims[2];
    out->euler_deg = (double *)malloc(sizeof(double) * 3 * V);
    out->grain = (int *)malloc(sizeof(int) * V);
    out->phase_id = (int *)malloc(sizeof(int) * V);
    if (!out->euler_deg || !out->grain || !out->phase_id) {
        voce_case_free(out);
        return VOCE_ERR_NOMEM;
    }
    if (n_grains > 0) {
        double *eg = (double *)malloc(sizeof(double) * 3 * n_grains);
        long v;
        if (!eg) {
            voce_case_free(out);
            return VOCE_ERR_NOMEM;
        }
        rc = voce_voronoi(s->dims, n_grains, seed, out->grain, eg);
        if (!rc)
            for (v = 0; v < V; ++v) {
                const int g = out->grain[v] - 1;
                out->euler_deg[3 * v] = eg[3 * g];
                out->euler_deg[3 * v + 1] = eg[3 * g + 1];
                out->euler_deg[3 * v + 2] = eg[3 * g + 2];
                out->phase_id[v] = 1;
            }
        free(eg);
        if (rc) {
            voce_case_free(out);
            return rc;
        }
    } else {
        join(path, sizeof(path), base_dir, nm.initial_ms);
        if ((rc = read_mesh(path, s->dims, out->euler_deg, out->grain, out->phase_id))) {
            voce_case_free(out);
            return rc;
        }
    }
    s->euler_deg = out->euler_deg;
    s->grain = out->grain;
    s->phase_id = out->phase_id;
    return 0;
}
