/*
 * analyse_block.c -- drop-in for the reference Analysis_Block executable
 * (Analysis_Block/main.c:6-79):  analyse_block <serial> <T> <F0> [N]
 * Reads ./para_file (line 4: start_ensemble end_ensemble, main.c:21-36) and,
 * for every job in that range, ../%.3lf(T)/data_00<job>/{time_index_file.dat,
 * data_coor.out} (read_all_data.c:5-99).  All frames stay in HBM (an
 * ac4_frames store) and, for block edges L/scale with scale = 3, 4, ..., 19
 * (main.c:45-63), libac4 writes BlockAnalysis/MSD/MSDLb*.dat,
 * BlockAnalysis/Relax/RelaxLb*.dat, BlockAnalysis/Relax/RelaxTotLb*.dat and
 * BlockAnalysis/vanHove*.dat (van Hove at time 1500.0, main.c:61).
 * As in the reference, <serial> and <F0> are read but the jobs come from
 * para_file.  N is compile-time in the reference (Analysis_Block/global.h:10,
 * 50000); here it is the optional 4th argument.  AC4_DIR=<dir> analyses that
 * one job directory instead of the para_file range; AC4_BIN=1 reads the
 * frames from data_coor.bin.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "ac4.h"
#include "xalloc.h"
#include "frame_reader.h"

static void die(const char *what, int rc) {
    fprintf(stderr, "analyse_block: %s failed (%d): %s\n", what, rc, ac4_last_error());
    exit(1);
}
#define CK(e) do { int rc_ = (e); if (rc_ != AC4_OK) die(#e, rc_); } while (0)

static int one_job(const char *sub_dir, int n, double L, int bin) {
    int nframes = 0;
    int *fidx = fr_read_index(sub_dir, 0, &nframes);
    if (!fidx) return 1;
    const int gap = ac4_origin_gap(fidx, nframes);
    if (gap <= 0) die("ac4_origin_gap", gap);
    ac4_frames *fs = NULL;
    CK(ac4_frames_create(n, L, nframes, &fs));
    double *x = xmalloc(sizeof(double) * n), *y = xmalloc(sizeof(double) * n), *z = xmalloc(sizeof(double) * n);
    int *ty = xmalloc(sizeof(int) * n);
    frame_reader rd;
    if (fr_open(&rd, sub_dir, n, bin)) return 1;
    for (int fr = 0; fr < nframes; fr++) {
        if (rd.tf) { /* text parsed on the GPU: device arrays */
            const double *dx, *dy, *dz;
            const int *dt;
            if (fr_next_dev(&rd, &dx, &dy, &dz, &dt)) return 1;
            CK(ac4_frames_push_device(fs, dx, dy, dz, fidx[fr]));
            continue;
        }
        if (fr_next(&rd, x, y, z, ty)) return 1;
        CK(ac4_frames_push(fs, x, y, z, fidx[fr]));
    }
    fr_close(&rd);
    printf("%d\t%d\n", n * nframes, nframes);
    double scale = 3.0, cut = L / scale;
    while (scale < 20.0) {
        CK(ac4_frames_block(fs, gap, 1.40000000, 0.3, cut, 1500.0, sub_dir));
        scale = scale + 1.0;
        cut = L / scale;
    }
    ac4_frames_destroy(fs);
    free(x); free(y); free(z); free(ty); free(fidx);
    return 0;
}

int main(int argc, char **argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: %s <serial> <T> <F0> [N]\n", argv[0]);
        return 2;
    }
    double T = atof(argv[2]);
    int n = argc > 4 ? atoi(argv[4]) : 50000;
    double L = pow((double)n / 1.2, 1.0 / 3.00); /* Analysis_Block/main.c:15 */
    int bin = getenv("AC4_BIN") && atoi(getenv("AC4_BIN"));
    if (getenv("AC4_DIR")) return one_job(getenv("AC4_DIR"), n, L, bin);
    FILE *pf = fopen("para_file", "r");
    if (!pf) {
        fprintf(stderr, "\nSorry man PARA file doesn't exit\n");
        return 0; /* main.c:26-29 */
    }
    char buff[1024];
    int start = 0, end = -1;
    for (int line = 1; line <= 4; line++) {
        if (!fgets(buff, sizeof buff, pf)) break;
        if (line == 4 && sscanf(buff, "%d\t%d", &start, &end) != 2) end = -1;
    }
    fclose(pf);
    char main_dir[128], sub_dir[256];
    snprintf(main_dir, sizeof main_dir, "../%.3lf", T);
    printf("%s\n", main_dir);
    for (int job = start; job <= end; job++) {
        snprintf(sub_dir, sizeof sub_dir, "%s/data_00%d", main_dir, job);
        if (one_job(sub_dir, n, L, bin)) return 1;
    }
    return 0;
}
