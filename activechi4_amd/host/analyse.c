/*
 * analyse.c -- drop-in for the reference Analysis executable
 * (Analysis/main.c:3-80):  analyse <serial> <T> <tauAlpha> [N]
 * reads ../%.3lf(T)/data_00<serial>/{time_index_file.dat,data_coor.out}
 * (read_all_data.c:5-102), streams every frame to libac4.so on the GPU and
 * writes msd.dat, relax_a_0.30.dat and S4qt_a_0.30.dat (the observables
 * Analysis/main.c:53-59 calls), plus Fsqt_a_0.30.dat when AC4_FS=1.
 * N is compile-time in the reference (Analysis/global.h:11, 50000); here it
 * is the optional 4th argument.  AC4_DIR overrides the job directory.
 * AC4_BIN=1 reads the frames from data_coor.bin (written by run with
 * AC4_BIN=1: the exact values data_coor.out parses to, so every output file
 * is the same) instead of parsing the text.  The text is parsed on the GPU
 * (ac4_textframes, the same values as the host reader); AC4_HOST_PARSE=1
 * parses it on the host instead.
 * Observables whose calls Analysis/main.c has commented out: AC4_GUU=1 adds
 * guuTauAlphaAct.dat (guu(factor, tauAlpha), main.c:64; active particles
 * from ActiveFileIndex.dat, read_all_data.c:49-58) and AC4_CAGE=1 adds
 * relaxMPICR1_a_0.30.dat (overlap_cage, main.c:57).  Both keep every frame
 * in HBM (an ac4_frames store, 24 B per particle and frame).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include "ac4.h"
#include "xalloc.h"
#include "frame_reader.h"

static int env_on(const char *k) { return getenv(k) && atoi(getenv(k)); }

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
    char sub_dir[512];
    if (getenv("AC4_DIR"))
        snprintf(sub_dir, sizeof sub_dir, "%s", getenv("AC4_DIR"));
    else
        snprintf(sub_dir, sizeof sub_dir, "../%.3lf/data_00%d", T, serial); /* Analysis/main.c:39,44 */
    int nframes = 0;
    int *fidx = fr_read_index(sub_dir, 1, &nframes);
    if (!fidx) return 1;
    double L = pow((double)n / 1.2, 1.0 / 3.0); /* Analysis/main.c:15 */
    int obs = AC4_OBS_MSD | AC4_OBS_OVERLAP | AC4_OBS_S4;
    if (getenv("AC4_FS") && atoi(getenv("AC4_FS"))) obs |= AC4_OBS_FS;
    ac4_ana *an = NULL;
    CK(ac4_ana_create(n, L, nframes, fidx, 0, 1.4, 0.3, obs, 0, &an));
    double *x = xmalloc(sizeof(double) * n), *y = xmalloc(sizeof(double) * n), *z = xmalloc(sizeof(double) * n);
    int *ty = xmalloc(sizeof(int) * n);
    const int guu = env_on("AC4_GUU"), cage = env_on("AC4_CAGE");
    ac4_frames *fs = NULL;
    if (guu || cage) CK(ac4_frames_create(n, L, nframes, &fs));
    frame_reader rd;
    if (fr_open(&rd, sub_dir, n, env_on("AC4_BIN"))) return 1;
    for (int fr = 0; fr < nframes; fr++) {
        if (rd.tf) { /* text parsed on the GPU: device arrays */
            const double *dx, *dy, *dz;
            const int *dt;
            if (fr_next_dev(&rd, &dx, &dy, &dz, &dt)) return 1;
            CK(ac4_ana_push_frame_device(an, fr, dx, dy, dz, dt));
            if (fs) CK(ac4_frames_push_device(fs, dx, dy, dz, fidx[fr]));
            continue;
        }
        if (fr_next(&rd, x, y, z, ty)) return 1;
        CK(ac4_ana_push_frame(an, fr, x, y, z, ty));
        if (fs) CK(ac4_frames_push(fs, x, y, z, fidx[fr]));
    }
    fr_close(&rd);
    CK(ac4_ana_finalize(an));
    CK(ac4_ana_write(an, sub_dir));
    printf("Overlap Calculated\n");
    ac4_ana_destroy(an);
    char path[640];
    if (fs) {
        const int gap = ac4_origin_gap(fidx, nframes);
        if (gap <= 0) die("ac4_origin_gap", gap);
        if (cage) {
            snprintf(path, sizeof path, "%s/relaxMPICR1_a_%.2lf.dat", sub_dir, 0.3); /* overlap_cage_rel.c:158 */
            CK(ac4_frames_overlap_cage(fs, gap, 1.4, 0.3, 2.5 + 0.35, path, NULL, 0, NULL));
        }
        if (guu) {
            int *act = xcalloc(n, sizeof(int)), k; /* read_all_data.c:49-58 */
            snprintf(path, sizeof path, "%s/ActiveFileIndex.dat", sub_dir);
            FILE *fa = fopen(path, "rb");
            if (!fa) { fprintf(stderr, "analyse: no %s\n", path); return 1; }
            while (fscanf(fa, "%d", &k) == 1)
                if (k >= 0 && k < n) act[k] = 1;
            fclose(fa);
            snprintf(path, sizeof path, "%s/guuTauAlphaAct.dat", sub_dir); /* guu.c:262 */
            CK(ac4_frames_guu(fs, act, gap, 1.4, atof(argv[3]), path, NULL, NULL, NULL, NULL));
            free(act);
        }
        ac4_frames_destroy(fs);
    }
    return 0;
}
