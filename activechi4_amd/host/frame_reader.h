/*
 * frame_reader.h -- the frames of one job directory, for the drop-in
 * analysis executables: time_index_file.dat, then N lines
 * "%.5lf\t%.5lf\t%.5lf\t%d" per frame from data_coor.out (read through
 * ac4_parse_frame, whose values equal fscanf's; Analysis/read_all_data.c:5-102,
 * Analysis_Block/read_all_data.c:5-99), or the binary sidecar data_coor.bin
 * that run writes with AC4_BIN=1 (the same values).
 */
#pragma once
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "ac4.h"

/* Frame step indices of sub_dir/time_index_file.dat.  header_count = 1: the
 * count on the "total timesteps equals %d" line (Analysis/read_all_data.c);
 * 0: the number of integers after that line (Analysis_Block/read_all_data.c:
 * 23-26).  Returns a malloc'd array or NULL (message printed). */
static int *fr_read_index(const char *sub_dir, int header_count, int *nframes) {
    char path[1024], buff[1024];
    snprintf(path, sizeof path, "%s/time_index_file.dat", sub_dir);
    FILE *f = fopen(path, "rb");
    if (!f) { fprintf(stderr, "Frameindex file does'nt exist %s\n", path); return NULL; }
    int nf = 0, v;
    if (!fgets(buff, sizeof buff, f)) { fprintf(stderr, "empty index file %s\n", path); fclose(f); return NULL; }
    if (header_count) {
        if (sscanf(buff, "total timesteps equals %d", &nf) != 1) nf = -1;
    } else {
        while (fscanf(f, "%d", &v) == 1) nf++;
        rewind(f);
        if (!fgets(buff, sizeof buff, f)) nf = -1;
    }
    if (nf < 2) { fprintf(stderr, "bad frame index file %s\n", path); fclose(f); return NULL; }
    int *fidx = malloc(sizeof(int) * nf);
    if (!fidx) { fprintf(stderr, "out of memory for %d frame indices\n", nf); fclose(f); return NULL; }
    for (int i = 0; i < nf; i++)
        if (!fgets(buff, sizeof buff, f) || sscanf(buff, "%d", &fidx[i]) != 1) {
            fprintf(stderr, "short index file %s\n", path);
            fclose(f);
            free(fidx);
            return NULL;
        }
    fclose(f);
    *nframes = nf;
    return fidx;
}

typedef struct {
    FILE *f;
    int n, bin, frame;
    char *buf;
    size_t cap, have, pos;
    /* text parsed on the GPU (ac4_textframes): the pinned chunk, its fill,
     * the bytes the current batch of frames took, and the batch */
    ac4_textframes *tf;
    char *tb;
    long long tcap, thave, tused;
    int tk, tnf, eof;
} frame_reader;

static int fr_open(frame_reader *r, const char *sub_dir, int n, int bin) {
    char path[1024];
    memset(r, 0, sizeof *r);
    r->n = n;
    r->bin = bin;
    snprintf(path, sizeof path, "%s/%s", sub_dir, bin ? "data_coor.bin" : "data_coor.out");
    r->f = fopen(path, "rb");
    if (!r->f) { fprintf(stderr, "DATA FILE DOES NOT EXIST %s\n", path); return -1; }
    if (bin) {
        char magic[8];
        int32_t hd[2];
        if (fread(magic, 1, 8, r->f) != 8 || memcmp(magic, "AC4FRM01", 8) || fread(hd, sizeof hd, 1, r->f) != 1 ||
            hd[0] != n) {
            fprintf(stderr, "%s is not a frame file of N=%d\n", path, n);
            fclose(r->f);
            r->f = NULL;
            return -1;
        }
    } else if (!(getenv("AC4_HOST_PARSE") && atoi(getenv("AC4_HOST_PARSE")))) {
        /* the text goes to the GPU in chunks; AC4_HOST_PARSE=1 keeps the host reader */
        if (ac4_textframes_create(n, 0, &r->tf) != AC4_OK) {
            fprintf(stderr, "text parser: %s\n", ac4_last_error());
            fclose(r->f);
            r->f = NULL;
            return -1;
        }
        r->tb = ac4_textframes_buffer(r->tf, &r->tcap);
    } else {
        r->cap = (size_t)n * 96 + (1u << 20);
        r->buf = malloc(r->cap);
        if (!r->buf) {
            fprintf(stderr, "out of memory for the frame text buffer\n");
            fclose(r->f);
            r->f = NULL;
            return -1;
        }
    }
    return 0;
}

/* The next frame into x, y, z, ty (n each).  0 on success. */
static int fr_next(frame_reader *r, double *x, double *y, double *z, int *ty) {
    const int n = r->n;
    if (r->bin) {
        uint64_t st;
        if (fread(&st, sizeof st, 1, r->f) != 1 || fread(x, sizeof(double), n, r->f) != (size_t)n ||
            fread(y, sizeof(double), n, r->f) != (size_t)n || fread(z, sizeof(double), n, r->f) != (size_t)n ||
            fread(ty, sizeof(int), n, r->f) != (size_t)n) {
            fprintf(stderr, "frame data ends in frame %d\n", r->frame);
            return -1;
        }
    } else {
        for (;;) { /* enough text for n lines? */
            size_t k = r->pos;
            int lines = 0;
            while (lines < n) {
                const char *nl = memchr(r->buf + k, '\n', r->have - k);
                if (!nl) break;
                k = (size_t)(nl - r->buf) + 1;
                lines++;
            }
            if (lines == n) break;
            memmove(r->buf, r->buf + r->pos, r->have - r->pos);
            r->have -= r->pos;
            r->pos = 0;
            if (r->have == r->cap) { /* a frame longer than the buffer */
                char *nb = realloc(r->buf, 2 * r->cap);
                if (!nb) { fprintf(stderr, "out of memory in frame %d\n", r->frame); return -1; }
                r->buf = nb;
                r->cap *= 2;
            }
            const size_t got = fread(r->buf + r->have, 1, r->cap - r->have, r->f);
            if (got == 0) { fprintf(stderr, "frame data ends in frame %d\n", r->frame); return -1; }
            r->have += got;
        }
        const long long used = ac4_parse_frame(r->buf + r->pos, (long long)(r->have - r->pos), n, x, y, z, ty);
        if (used < 0) { fprintf(stderr, "frame %d: %s\n", r->frame, ac4_last_error()); return -1; }
        r->pos += (size_t)used;
    }
    r->frame++;
    return 0;
}

/* With the GPU parser (r->tf): the next frame as device arrays, valid until
 * the call after the last frame of the current batch.  0 on success. */
static int fr_next_dev(frame_reader *r, const double **x, const double **y, const double **z, const int **ty) {
    while (r->tk == r->tnf) { /* the batch is used up: parse the next one */
        if (r->tused) {
            memmove(r->tb, r->tb + r->tused, (size_t)(r->thave - r->tused));
            r->thave -= r->tused;
            r->tused = 0;
        }
        if (!r->eof && r->thave < r->tcap) {
            const size_t got = fread(r->tb + r->thave, 1, (size_t)(r->tcap - r->thave), r->f);
            r->thave += (long long)got;
            if (got == 0) r->eof = 1;
        }
        /* a last line without its newline still counts (as for ac4_parse_frame) */
        if (r->eof && r->thave > 0 && r->tb[r->thave - 1] != '\n' && r->thave < r->tcap) r->tb[r->thave++] = '\n';
        int nf = 0;
        long long used = 0;
        if (ac4_textframes_parse(r->tf, r->thave, 1 << 30, &nf, &used) != AC4_OK) {
            fprintf(stderr, "frame %d: %s\n", r->frame, ac4_last_error());
            return -1;
        }
        if (nf > 0) {
            r->tnf = nf;
            r->tk = 0;
            r->tused = used;
        } else if (r->eof || r->thave == r->tcap) {
            fprintf(stderr, "frame data ends in frame %d\n", r->frame);
            return -1;
        }
    }
    if (ac4_textframes_get(r->tf, r->tk, x, y, z, ty) != AC4_OK) return -1;
    r->tk++;
    r->frame++;
    return 0;
}

static void fr_close(frame_reader *r) {
    if (r->f) fclose(r->f);
    if (r->tf) ac4_textframes_destroy(r->tf);
    free(r->buf);
    memset(r, 0, sizeof *r);
}
