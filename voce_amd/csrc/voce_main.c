/*
 * voce -- host driver with the reference's command line and outputs.
 *
 * `voce input [fp64|mixed]` follows main() (evp.c:7-32): read the namelist,
 * the .sx file and the mesh, set up, run N_steps increments, then tear
 * down.  It writes vm.out and err.out in the working directory with the
 * reference's headers and row formats (init.c:539-540, evolution.c:302,
 * 427, 563-564) and the field dumps that PrintControl and Tex_Flag ask for
 * (evolution.c:569-611).  All solver work goes through include/voce.h.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../include/voce.h"

static double now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int main(int argc, char **argv)
{
    voce_case cs;
    voce_ctx *ctx = NULL;
    voce_incr_report rep;
    FILE *fvm, *ferr;
    int rc, s, it, pc0, pc1;
    double t0;
    if (argc < 2) {
        fprintf(stderr, "usage: voce input [fp64|mixed]\n");
        return 2;
    }
    rc = voce_load_case(argv[1], NULL, &cs);
    if (rc) {
        fprintf(stderr, "voce: cannot read input: %s\n", voce_strerror(rc));
        return rc > 0 ? rc : 1;
    }
    if (argc > 2 && strcmp(argv[2], "mixed") == 0) cs.setup.precision = VOCE_MIXED;
    rc = voce_ctx_create(&cs.setup, &ctx);
    if (rc) {
        fprintf(stderr, "voce: setup failed: %s\n", voce_strerror(rc));
        voce_case_free(&cs);
        return rc > 0 ? rc : 1;
    }
    pc0 = cs.print_control[0];
    pc1 = cs.print_control[1] > 0 ? cs.print_control[1] : 1;
    if ((pc0 == 1) || cs.tex_flag == 1) voce_set_diagnostics(ctx, 1);
    fvm = fopen("vm.out", "w+");
    ferr = fopen("err.out", "w+");
    if (!fvm || !ferr) {
        fprintf(stderr, "voce: cannot open vm.out/err.out\n");
        return 1;
    }
    fprintf(fvm, "EVM\tEVMP\tDVM\tDVMP\tSVM\tSVMSVM_1\tIteration_tot\n");
    fprintf(ferr, "IT\tERRE\tERRS\tSVM\n");
    t0 = now();
    for (s = 1; s <= cs.n_steps; ++s) {
        rc = voce_increment(ctx, &rep);
        if (rc) {
            fprintf(stderr, "voce: increment %d failed: %s\n", s, voce_strerror(rc));
            break;
        }
        if (cs.n_steps != 1) fprintf(ferr, "STEP = %d\n", s);
        for (it = 0; it < rep.iter; ++it)
            fprintf(ferr, "%d\t%e\t%e\t%e\n", it + 1, rep.err_e[it], rep.err_s[it], rep.svm_it[it]);
        fprintf(fvm, "%e\t%e\t%e\t%e\t%e\t%e\t%d\n", rep.e_vm, rep.evmp, rep.d_vm, rep.dvmp, rep.s_vm, rep.s_vm1,
                rep.iter_total);
        fflush(fvm);
        fflush(ferr);
        if (pc0 == 1 && (s / pc1 * pc1 == s)) voce_write_fields(ctx, ".", s, cs.tex_flag, 0);
    }
    if (!rc && cs.tex_flag == 1) voce_write_fields(ctx, ".", 0, 1, 1);
    fprintf(stderr, "voce: %d increments in %.3f s\n", s - 1, now() - t0);
    fclose(fvm);
    fclose(ferr);
    voce_ctx_destroy(ctx);
    voce_case_free(&cs);
    return rc > 0 ? rc : (rc < 0 ? 1 : 0);
}
