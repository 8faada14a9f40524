/*
 * oracle_block.c -- CPU restatement of the reference block analysis
 * (Analysis_Block/: main.c, read_all_data.c, msd_pos.c, overlap.c, vanHove.c).
 * TEST INFRASTRUCTURE ONLY: the checker, never the product.
 *
 * For each block edge L/scale, scale = 3..19 (Analysis_Block/main.c:53-63),
 * the box is cut into ncx^3 blocks, ncx = (int)(L / (L/scale)), and every
 * particle is tagged with the block of its wrapped ORIGIN position.  Per
 * (origin, lag) pair and block: the overlap count, the sum of squared
 * displacements, and the block centre-of-mass displacement; the van Hove
 * file lists the block COM displacements at the one lag nearest a time.
 *
 * Parity pins (tests/test_oracle_block.py): byte-identical output files
 * against the reference program built from Analysis_Block/ (oracle/ref/
 * Makefile) on the golden trajectories; closed forms (rigid shift of the
 * whole box: COM displacement = shift; a static trajectory: Q = 1, MSD = 0).
 *
 * Two places where the reference reads uninitialised memory are given the
 * meaning the reference binary built here shows (vanHove.c):
 *  - the lag ladder of :31-44 runs from the first frame with offset 1 and
 *    slot 0 (its locals other/offset/current_frame are never set), and only
 *    the slots it fills are searched (:46-51 reads up to num_of_intervals);
 *  - FrameNum is 0 when no lag lies within 1000 time units of Time, so each
 *    origin is paired with itself.
 * Expression order follows the reference; build with -ffp-contract=off.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include "oracle.h"

#define B_TIME_STEP 0.005 /* Analysis_Block/global.h:14 */
#define B_DIVIDE_TO 2.0   /* Analysis_Block/global.h:16 */

/* Analysis_Block/overlap.c:86-91: while-loop wrap into [0,1) */
static double wrap_loop(double v) {
    while (v >= 1.0) v -= 1.0;
    while (v < 0.0) v += 1.0;
    return v;
}

/* overlap.c:127-138: one minimum-image correction */
static double mi1(double d) {
    if (d >= 0.5)
        d -= 1.0;
    else if (d < -0.5)
        d += 1.0;
    return d;
}

static void corr(double *acc, double d) { /* msd_pos.c:139-150 */
    if (d >= 0.5)
        *acc -= 1.0;
    else if (d < -0.5)
        *acc += 1.0;
}

/* origins visited by the do-while of msd_pos.c:70-195 (frame positions) */
static int origin_sequence(const int *fidx, int nframes, int gap, int *out) {
    int current = 0, k = 0, n = 0;
    do {
        out[n++] = current;
        k++;
        while (current < nframes && fidx[current] != k * gap) current++;
    } while (current < nframes);
    return n;
}

typedef struct {
    int N, ncx, ncell;
    double L, inv;
    const double *X, *Y, *Z;
    int *cell, *nin;           /* of the current origin */
    double *x0, *y0, *z0;      /* wrapped origin frame */
    double *com0;              /* [ncell][3] */
} blk_t;

static void load_origin(blk_t *b, int f) { /* msd_pos.c:72-108 */
    const int N = b->N, ncx = b->ncx, ncxy = ncx * ncx;
    memset(b->nin, 0, sizeof(int) * b->ncell);
    memset(b->com0, 0, sizeof(double) * 3 * b->ncell);
    for (int i = 0; i < N; i++) {
        b->x0[i] = wrap_loop(b->X[(size_t)f * N + i]);
        b->y0[i] = wrap_loop(b->Y[(size_t)f * N + i]);
        b->z0[i] = wrap_loop(b->Z[(size_t)f * N + i]);
    }
    for (int i = 0; i < N; i++) {
        int c = floor(b->x0[i] * b->inv) + floor(b->y0[i] * b->inv) * ncx + floor(b->z0[i] * b->inv) * ncxy;
        b->cell[i] = c;
        b->nin[c]++;
    }
    for (int i = 0; i < N; i++) {
        b->com0[3 * b->cell[i] + 0] += b->x0[i];
        b->com0[3 * b->cell[i] + 1] += b->y0[i];
        b->com0[3 * b->cell[i] + 2] += b->z0[i];
    }
}

/* Per block for the lag frame f: overlap count (overlap.c:119-143), sum of
 * r^2 (msd_pos.c:126-170) and COM with corrections (msd_pos.c:131-150). */
static void pair_blocks(const blk_t *b, int f, double param, double *ov, double *s2, double *com) {
    const int N = b->N;
    const double L = b->L;
    memset(ov, 0, sizeof(double) * b->ncell);
    memset(s2, 0, sizeof(double) * b->ncell);
    memset(com, 0, sizeof(double) * 3 * b->ncell);
    for (int i = 0; i < N; i++) {
        double x = wrap_loop(b->X[(size_t)f * N + i]);
        double y = wrap_loop(b->Y[(size_t)f * N + i]);
        double z = wrap_loop(b->Z[(size_t)f * N + i]);
        double dx = x - b->x0[i], dy = y - b->y0[i], dz = z - b->z0[i];
        const int c = b->cell[i];
        com[3 * c + 0] += x;
        com[3 * c + 1] += y;
        com[3 * c + 2] += z;
        corr(&com[3 * c + 2], dz);
        corr(&com[3 * c + 1], dy);
        corr(&com[3 * c + 0], dx);
        dz = mi1(dz);
        dy = mi1(dy);
        dx = mi1(dx);
        double r2 = L * L * (dx * dx + dy * dy + dz * dz);
        s2[c] += r2;
        if (r2 < param) ov[c] += 1.0;
    }
}

static int count_origins1(const int *fidx, int nframes, int gap) { /* overlap.c:16-23 */
    int current = 0, on = 0;
    do {
        while (current < nframes && fidx[current] != on * gap + 1) current++;
        if (current < nframes) on++;
    } while (current < nframes);
    return on;
}

int or_block_scale(int N, double L, int nframes, const int *fidx, int gap, const double *X, const double *Y,
                   const double *Z, double factor, double param, double cell_cut, double vh_time,
                   const char *sub_dir) {
    blk_t b;
    memset(&b, 0, sizeof b);
    b.N = N; b.L = L; b.X = X; b.Y = Y; b.Z = Z;
    b.ncx = (int)(L / cell_cut); /* msd_pos.c:33 */
    b.ncell = b.ncx * b.ncx * b.ncx;
    b.inv = (double)b.ncx;
    const int ncell = b.ncell;
    const int maxint = count_origins1(fidx, nframes, gap) * gap / B_DIVIDE_TO;
    const int noi = 10 + (int)(log((double)maxint) / log(factor));
    int cap = nframes * 64 + 64;
    int *pc = malloc(sizeof(int) * cap), *po = malloc(sizeof(int) * cap), *ps = malloc(sizeof(int) * cap);
    int mi_ = 0;
    int np = or_walk(fidx, nframes, gap, factor, 1, cap, pc, po, ps, &mi_);
    int *orig = malloc(sizeof(int) * (nframes + 1));
    int nor = origin_sequence(fidx, nframes, gap, orig);
    b.cell = malloc(sizeof(int) * N);
    b.nin = malloc(sizeof(int) * ncell);
    b.x0 = malloc(sizeof(double) * N); b.y0 = malloc(sizeof(double) * N); b.z0 = malloc(sizeof(double) * N);
    b.com0 = malloc(sizeof(double) * 3 * ncell);
    double *ov = malloc(sizeof(double) * ncell), *s2 = malloc(sizeof(double) * ncell);
    double *com = malloc(sizeof(double) * 3 * ncell);
    /* per (block, slot) and per slot accumulators */
    size_t bs = (size_t)ncell * noi;
    double *c = calloc(bs, 8), *x4 = calloc(bs, 8), *msd = calloc(bs, 8);
    int *cnt = calloc(bs, 4);
    double *dt = calloc(noi, 8), *cT = calloc(noi, 8), *x4T = calloc(noi, 8), *mc = calloc(noi, 8), *mc2 = calloc(noi, 8);
    int *cT_n = calloc(noi, 4), *binder = calloc(noi, 4);
    if (np < 0 || maxint <= 0) return -1;

    int p = 0;
    for (int o = 0; o < nor; o++) {
        load_origin(&b, orig[o]);
        for (; p < np && pc[p] == orig[o]; p++) {
            const int t = ps[p];
            if (t >= noi) return -1;
            pair_blocks(&b, po[p], param, ov, s2, com);
            const double dts = B_TIME_STEP * (double)(fidx[po[p]] - fidx[pc[p]]);
            for (int i = 0; i < ncell; i++) { /* msd_pos.c:171-181 */
                const size_t k = (size_t)i * noi + t;
                msd[k] += s2[i] / (double)b.nin[i];
                double dx = (com[3 * i + 0] - b.com0[3 * i + 0]) / b.nin[i];
                double dy = (com[3 * i + 1] - b.com0[3 * i + 1]) / b.nin[i];
                double dz = (com[3 * i + 2] - b.com0[3 * i + 2]) / b.nin[i];
                mc[t] += L * L * (dx * dx + dy * dy + dz * dz);
                mc2[t] += L * L * L * L * (dx * dx * dx * dx + dy * dy * dy * dy + dz * dz * dz * dz);
                binder[t]++;
            }
            for (int i = 0; i < ncell; i++) { /* overlap.c:144-152 */
                const size_t k = (size_t)i * noi + t;
                c[k] += ov[i] / (double)b.nin[i];
                x4[k] += (ov[i] / (double)b.nin[i]) * (ov[i] / (double)b.nin[i]);
                cnt[k]++;
                cT[t] += ov[i] / (double)b.nin[i];
                cT_n[t]++;
            }
            dt[t] = dts;
        }
    }
    /* after the walk b.nin holds the last origin's counts (overlap.c:183) */
    char path[1024];
    const double lb = L / b.inv;
    FILE *fm, *fr, *ft;
    snprintf(path, sizeof path, "%s/BlockAnalysis/MSD/MSDLb%.3lf.dat", sub_dir, lb);
    fm = fopen(path, "wb");
    snprintf(path, sizeof path, "%s/BlockAnalysis/Relax/RelaxLb%.3lf.dat", sub_dir, lb);
    fr = fopen(path, "wb");
    snprintf(path, sizeof path, "%s/BlockAnalysis/Relax/RelaxTotLb%.3lf.dat", sub_dir, lb);
    ft = fopen(path, "wb");
    if (!fm || !fr || !ft) return -1;
    fprintf(fm, "0.0\t0.0\t0.0\t0.0\n");
    double avg_nin = 0.0;
    for (int i = 0; i < ncell; i++) avg_nin += (double)b.nin[i];
    avg_nin /= (double)ncell;
    for (int j = 1; j < noi; j++) {
        double bm = 0.0, bc = 0.0, bx = 0.0;
        int bn = 0, bnm = 0;
        for (int i = 0; i < ncell; i++) {
            const size_t k = (size_t)i * noi + j;
            if (cnt[k]) { /* msd_pos.c:206-212 (same counters as overlap) */
                bm += msd[k] / (double)cnt[k];
                bnm++;
            }
            if (cnt[k]) { /* overlap.c:180-186 */
                bc += c[k] / (double)cnt[k];
                bx += (x4[k] / (double)cnt[k] - c[k] * c[k] / ((double)(cnt[k] * cnt[k]))) * b.nin[i];
                bn++;
                x4T[j] += x4[k];
            }
        }
        if (bnm)
            fprintf(fm, "%.14lf\t%.14lf\t%.14lf\t%.14lf\n", dt[j], bm / (double)bnm, mc[j] / (double)binder[j],
                    mc2[j] / (double)binder[j]);
        if (bn) fprintf(fr, "%.14lf\t%.14lf\t%.14lf\n", dt[j], bc / (double)bn, bx / (double)bn);
        if (cT_n[j])
            fprintf(ft, "%.14lf\t%.14lf\t%.14lf\n", dt[j], cT[j] / (double)cT_n[j],
                    avg_nin * (x4T[j] / (double)cT_n[j] - cT[j] * cT[j] / ((double)cT_n[j] * cT_n[j])));
    }
    fclose(fm);
    fclose(fr);
    fclose(ft);

    /* van Hove (vanHove.c:20-51): the lag nearest vh_time on the first
     * origin's ladder, then every origin paired with origin + FrameNum */
    int frame_num = 0;
    {
        int other = 0, offset = 1, ti = 0;
        const int cf = fidx[0];
        double *dl = calloc(noi + 64, 8);
        while (other < nframes && offset < maxint) {
            const int of = cf + offset;
            while (other < nframes && fidx[other] != of) other++;
            if (other < nframes) {
                dl[ti++] = B_TIME_STEP * (double)(of - cf);
                if ((int)(offset * factor) == offset)
                    offset++;
                else
                    offset = (int)(((double)offset) * factor);
            }
        }
        double inter = 1000.0;
        for (int i = 1; i < noi && i < ti; i++) {
            if (fabs(dl[i] - vh_time) < inter) {
                frame_num = (int)(dl[i] * 200.0);
                inter = fabs(dl[i] - vh_time);
            }
        }
        free(dl);
    }
    snprintf(path, sizeof path, "%s/BlockAnalysis/vanHove%.3lf.dat", sub_dir, lb);
    FILE *fv = fopen(path, "wb");
    if (!fv) return -1;
    for (int o = 0; o < nor; o++) {
        const int cur = orig[o];
        const int want = fidx[cur] + frame_num;
        int other = cur;
        while (other < nframes && fidx[other] != want) other++;
        if (other >= nframes) continue;
        load_origin(&b, cur);
        pair_blocks(&b, other, param, ov, s2, com);
        for (int i = 0; i < ncell; i++)
            for (int a = 0; a < 3; a++)
                fprintf(fv, "%.14lf\n", ((com[3 * i + a] - b.com0[3 * i + a]) / b.nin[i]) * L);
    }
    fclose(fv);
    free(pc); free(po); free(ps); free(orig);
    free(b.cell); free(b.nin); free(b.x0); free(b.y0); free(b.z0); free(b.com0);
    free(ov); free(s2); free(com);
    free(c); free(x4); free(msd); free(cnt);
    free(dt); free(cT); free(x4T); free(mc); free(mc2); free(cT_n); free(binder);
    return 0;
}

/* Analysis_Block/main.c:45-70 for one job directory: read the frames
 * (read_all_data.c:16-84, the frame count is the number of integers after
 * the header line), then block edges L/scale for scale = 3, 4, ..., 19. */
int or_block_analyse_dir(const char *sub_dir, int N, double vh_time) {
    char path[1024], buff[1024];
    snprintf(path, sizeof path, "%s/time_index_file.dat", sub_dir);
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    int nframes = 0, v;
    if (!fgets(buff, 1024, f)) { fclose(f); return -1; }
    while (fscanf(f, "%d", &v) != EOF) nframes++;
    rewind(f);
    int *fidx = malloc(sizeof(int) * (nframes + 10));
    if (!fgets(buff, 1024, f)) { fclose(f); free(fidx); return -1; }
    for (int i = 0; i < nframes; i++) {
        if (!fgets(buff, 1024, f)) { fclose(f); free(fidx); return -1; }
        sscanf(buff, "%d", &fidx[i]);
    }
    fclose(f);
    int gap = or_origin_gap(fidx, nframes);
    size_t tot = (size_t)nframes * N;
    double *X = malloc(sizeof(double) * tot), *Y = malloc(sizeof(double) * tot), *Z = malloc(sizeof(double) * tot);
    snprintf(path, sizeof path, "%s/data_coor.out", sub_dir);
    f = fopen(path, "rb");
    if (!f) return -1;
    int ty;
    for (size_t k = 0; k < tot; k++)
        if (fscanf(f, "%lf %lf %lf %d", &X[k], &Y[k], &Z[k], &ty) != 4) { fclose(f); return -1; }
    fclose(f);
    snprintf(path, sizeof path, "%s/BlockAnalysis", sub_dir);
    mkdir(path, 0755);
    snprintf(path, sizeof path, "%s/BlockAnalysis/MSD", sub_dir);
    mkdir(path, 0755);
    snprintf(path, sizeof path, "%s/BlockAnalysis/Relax", sub_dir);
    mkdir(path, 0755);
    const double L = pow((double)N / 1.2, 1.0 / 3.00); /* Analysis_Block/main.c:15 */
    int rc = 0;
    double scale = 3.0, cut = L / scale;
    while (scale < 20.0) {
        if (or_block_scale(N, L, nframes, fidx, gap, X, Y, Z, 1.40000000, 0.3 * 0.3, cut, vh_time, sub_dir) < 0) rc = -1;
        scale = scale + 1.0;
        cut = L / scale;
    }
    free(X); free(Y); free(Z); free(fidx);
    return rc;
}
