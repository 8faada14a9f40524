/*
 * analyse.c -- drop-in for the reference Analysis executable
 * (Analysis/main.c:3-80):  analyse <serial> <T> <tauAlpha> [N]
 * reads ../%.3lf(T)/data_00<serial>/{time_index_file.dat,data_coor.out}
 * (read_all_data.c:5-102), streams every frame to libac4.so on the GPU and
 * writes msd.dat, relax_a_0.30.dat and S4qt_a_0.30.dat (the observables
 * Analysis/main.c:53-59 calls), plus Fsqt_a_0.30.dat when AC4_FS=1.
 * N is compile-time in the reference (Analysis/global.h:11, 50000); here it
 * is the optional 4th argument.  AC4_DIR overrides the job directory.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "ac4.h"

static void die(const char *what, int rc) {
    fprintf(stderr, "analyse: %s failed (%d): %s\n", what, rc, ac4_last_error());
    exit(1);
}
#define CK(e) do { int rc_ = (e); if (rc_ != AC4_OK) die(#e, rc_); } while (0)

int main(int argc, char **argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: %s <serial> <T> <tauAlpha> [N]\n", argv[0]);
        return 2;
    }
    int serial = atoi(argv[1]);
    double T = atof(argv[2]);
    int n = argc > 4 ? atoi(argv[4]) : 50000;
    char sub_dir[512], path[640], buff[1024];
    if (getenv("AC4_DIR"))
        snprintf(sub_dir, sizeof sub_dir, "%s", getenv("AC4_DIR"));
    else
        snprintf(sub_dir, sizeof sub_dir, "../%.3lf/data_00%d", T, serial); /* Analysis/main.c:39,44 */
    snprintf(path, sizeof path, "%s/time_index_file.dat", sub_dir);
    FILE *f = fopen(path, "rb");
    if (!f) { fprintf(stderr, "Frameindex file does'nt exist %s\n", path); return 1; }
    int nframes = 0;
    if (!fgets(buff, sizeof buff, f) || sscanf(buff, "total timesteps equals %d", &nframes) != 1 || nframes < 2) {
        fprintf(stderr, "analyse: bad header in %s\n", path);
        return 1;
    }
    int *fidx = malloc(sizeof(int) * nframes);
    for (int i = 0; i < nframes; i++) {
        if (!fgets(buff, sizeof buff, f) || sscanf(buff, "%d", &fidx[i]) != 1) {
            fprintf(stderr, "analyse: short index file\n");
            return 1;
        }
    }
    fclose(f);
    double L = pow((double)n / 1.2, 1.0 / 3.0); /* Analysis/main.c:15 */
    int obs = AC4_OBS_MSD | AC4_OBS_OVERLAP | AC4_OBS_S4;
    if (getenv("AC4_FS") && atoi(getenv("AC4_FS"))) obs |= AC4_OBS_FS;
    ac4_ana *an = NULL;
    CK(ac4_ana_create(n, L, nframes, fidx, 0, 1.4, 0.3, obs, 0, &an));
    double *x = malloc(sizeof(double) * n), *y = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n);
    int *ty = malloc(sizeof(int) * n);
    snprintf(path, sizeof path, "%s/data_coor.out", sub_dir);
    f = fopen(path, "rb");
    if (!f) { fprintf(stderr, "DATA FILE DOES NOT EXIST %s\n", path); return 1; }
    for (int fr = 0; fr < nframes; fr++) {
        for (int i = 0; i < n; i++) {
            if (!fgets(buff, sizeof buff, f)) { fprintf(stderr, "analyse: data ends in frame %d\n", fr); return 1; }
            char *e;
            x[i] = strtod(buff, &e);
            y[i] = strtod(e, &e);
            z[i] = strtod(e, &e);
            ty[i] = (int)strtol(e, &e, 10);
        }
        CK(ac4_ana_push_frame(an, fr, x, y, z, ty));
    }
    fclose(f);
    CK(ac4_ana_finalize(an));
    CK(ac4_ana_write(an, sub_dir));
    printf("Overlap Calculated\n");
    ac4_ana_destroy(an);
    return 0;
}
