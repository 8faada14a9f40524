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
 * is the same) instead of parsing the text.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
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
    if (getenv("AC4_BIN") && atoi(getenv("AC4_BIN"))) {
        snprintf(path, sizeof path, "%s/data_coor.bin", sub_dir);
        f = fopen(path, "rb");
        if (!f) { fprintf(stderr, "DATA FILE DOES NOT EXIST %s\n", path); return 1; }
        char magic[8];
        int32_t hd[2];
        if (fread(magic, 1, 8, f) != 8 || memcmp(magic, "AC4FRM01", 8) || fread(hd, sizeof hd, 1, f) != 1 ||
            hd[0] != n) {
            fprintf(stderr, "analyse: %s is not a frame file of N=%d\n", path, n);
            return 1;
        }
        for (int fr = 0; fr < nframes; fr++) {
            uint64_t st;
            if (fread(&st, sizeof st, 1, f) != 1 || fread(x, sizeof(double), n, f) != (size_t)n ||
                fread(y, sizeof(double), n, f) != (size_t)n || fread(z, sizeof(double), n, f) != (size_t)n ||
                fread(ty, sizeof(int), n, f) != (size_t)n) {
                fprintf(stderr, "analyse: data ends in frame %d\n", fr);
                return 1;
            }
            CK(ac4_ana_push_frame(an, fr, x, y, z, ty));
        }
        fclose(f);
    } else {
        /* read_all_data.c:61-91: n lines per frame, parsed by ac4_parse_frame
         * (values identical to fscanf's) from a large buffer */
        snprintf(path, sizeof path, "%s/data_coor.out", sub_dir);
        f = fopen(path, "rb");
        if (!f) { fprintf(stderr, "DATA FILE DOES NOT EXIST %s\n", path); return 1; }
        size_t cap = (size_t)n * 96 + (1u << 20), have = 0, pos = 0;
        char *buf = malloc(cap);
        for (int fr = 0; fr < nframes; fr++) {
            for (;;) { /* enough text for n lines? */
                size_t k = pos;
                int lines = 0;
                while (lines < n) {
                    const char *nl = memchr(buf + k, '\n', have - k);
                    if (!nl) break;
                    k = (size_t)(nl - buf) + 1;
                    lines++;
                }
                if (lines == n) break;
                memmove(buf, buf + pos, have - pos);
                have -= pos;
                pos = 0;
                if (have == cap) { cap *= 2; buf = realloc(buf, cap); }
                const size_t got = fread(buf + have, 1, cap - have, f);
                if (got == 0) { fprintf(stderr, "analyse: data ends in frame %d\n", fr); return 1; }
                have += got;
            }
            const long long used = ac4_parse_frame(buf + pos, (long long)(have - pos), n, x, y, z, ty);
            if (used < 0) die("ac4_parse_frame", (int)used);
            pos += (size_t)used;
            CK(ac4_ana_push_frame(an, fr, x, y, z, ty));
        }
        free(buf);
        fclose(f);
    }
    CK(ac4_ana_finalize(an));
    CK(ac4_ana_write(an, sub_dir));
    printf("Overlap Calculated\n");
    ac4_ana_destroy(an);
    return 0;
}
