/*
 * evp_oracle -- command-line driver for the oracle (test infrastructure).
 * Mirrors the reference's main (evp.c:7-32) and its per-increment output
 * (evolution.c:297-304, 427, 562-612): `evp_oracle input [threads]` writes
 * vm.out and err.out in the working directory, plus the field dumps that
 * PrintControl asks for.  Used to time the CPU baseline on the host cores.
 */
#include <stdio.h>
#include <stdlib.h>
#include <time.h>
#include "oracle.h"

static double now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int main(int argc, char **argv)
{
    int err = 0, n, s, nsteps, it;
    orc_job *J;
    orc_report *rep;
    FILE *fvm, *ferr;
    double t0, t1;
    int pc[2] = {0, 1};
    if (argc < 2) { fprintf(stderr, "usage: evp_oracle input [threads] [max_steps]\n"); return 2; }
    if (argc > 2) orc_set_threads(atoi(argv[2]));
    t0 = now();
    J = orc_setup(argv[1], NULL, &err);
    if (!J) { fprintf(stderr, "setup failed: code %d\n", err); return err > 0 ? err : 1; }
    t1 = now();
    fprintf(stderr, "setup %.3f s, V=%d\n", t1 - t0, orc_nvox(J));
    fvm = fopen("vm.out", "w+");
    ferr = fopen("err.out", "w+");
    fprintf(fvm, "EVM\tEVMP\tDVM\tDVMP\tSVM\tSVMSVM_1\tIteration_tot\n");
    fprintf(ferr, "IT\tERRE\tERRS\tSVM\n");
    nsteps = orc_nsteps(J);
    if (argc > 3 && atoi(argv[3]) > 0 && atoi(argv[3]) < nsteps) nsteps = atoi(argv[3]);
    rep = (orc_report *)malloc(sizeof(orc_report));
    t0 = now();
    for (s = 1; s <= nsteps; s++) {
        n = orc_increment(J, rep);
        if (n) { fprintf(stderr, "increment %d failed: code %d\n", s, n); return n; }
        if (orc_nsteps(J) != 1) fprintf(ferr, "STEP = %d\n", s);
        for (it = 0; it < rep->iter; it++)
            fprintf(ferr, "%d\t%e\t%e\t%e\n", it + 1, rep->err_e[it], rep->err_s[it], rep->svm_it[it]);
        fprintf(fvm, "%e\t%e\t%e\t%e\t%e\t%e\t%d\n", rep->e_vm, rep->evmp, rep->d_vm, rep->dvmp, rep->s_vm,
                rep->s_vm1, rep->iter_total);
        fflush(fvm);
        fflush(ferr);
        (void)pc;
    }
    t1 = now();
    fprintf(stderr, "increments %d wall %.3f s iter_total %d voxel-it/s %.4e\n", nsteps, t1 - t0,
            rep->iter_total, (double)orc_nvox(J) * rep->iter_total / (t1 - t0));
    fclose(fvm);
    fclose(ferr);
    free(rep);
    orc_free(J);
    return 0;
}
