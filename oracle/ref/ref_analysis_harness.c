/*
 * ref_analysis_harness.c -- driver for the reference Analysis program, built
 * from the reference sources where they lie (oracle/ref/Makefile) into
 * oracle/_ref/.  TEST INFRASTRUCTURE: it only produces golden vectors.
 *
 * It repeats the reference main's set-up (Analysis/main.c:3-37,45-51) for one
 * job directory given on the command line, instead of "../%.3lf/data_00%d",
 * and then calls the observables of Analysis/main.c:53-62 selected by a
 * letter string: m = mean_square_displacement(1.4,0,0),
 * o = overlap_correlation_function(1.4,0.3,0,0), s = S4Qt(1.4,0.3),
 * f = Fsqt(1.4,0.3,0), g = guu(1.4, tauAlpha) (the call Analysis/main.c:64 has
 * commented out; guuTauAlphaAct.dat).
 * usage: analyse_ref <sub_dir> <letters> [tauAlpha]
 */
#include <string.h>
#include "global.h"
#include "subroutine.h"

int main(int argc, char **argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s <sub_dir> <letters: m o s f g> [tauAlpha]\n", argv[0]);
        return 2;
    }
    double factor = 1.4000000;
    MPI_Init(&argc, &argv);
    MPI_Comm_size(MPI_COMM_WORLD, &Nprocs);
    MPI_Comm_rank(MPI_COMM_WORLD, &node);
    L = pow((double)N / RHO, 1.0 / 3.0);
    invL = 1.0 / L;
    nq = 15;
    NT = 23 * 12 + 100;
    snprintf(sub_dir, sizeof sub_dir, "%s", argv[1]);
    if (node == 0) read_data_file();
    BCastConstants();
    GetTheOriginConfigs();
    SplitTheSystem();
    const char *w = argv[2];
    if (strchr(w, 'm')) mean_square_displacement(factor, 0, 0);
    if (strchr(w, 'o')) overlap_correlation_function(factor, 0.3, 0, 0);
    if (strchr(w, 's')) S4Qt(factor, 0.3);
    if (strchr(w, 'f')) Fsqt(factor, 0.3, 0);
    if (strchr(w, 'g')) guu(factor, argc > 3 ? atof(argv[3]) : 0.0);
    MPI_Barrier(MPI_COMM_WORLD);
    MPI_Finalize();
    return 0;
}
