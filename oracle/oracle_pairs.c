/*
 * oracle_pairs.c -- CPU restatement of the reference's pair observables
 * (TEST INFRASTRUCTURE ONLY: the checker, never the product):
 *   or_guu          Analysis/guu.c:3-300, the displacement-displacement
 *                   correlation g_uu(r) of the active particles at the lag
 *                   nearest tau (the frameindx = 3 branch, guuTauAlphaAct.dat);
 *   or_overlap_cage Analysis/overlap_cage_rel.c:4-269, the overlap of each
 *                   particle's displacement relative to the mean displacement
 *                   of its 1.3-sigma neighbours at the origin;
 *   or_frame_printf, or_frame_strtod  glibc sprintf / strtod per line of the
 *                   frame text (MDcode/print.c:10, Analysis/read_all_data.c:81),
 *                   the CPU baseline of the library's frame formatter/parser.
 * One MPI rank: NPP = N, GatherVec and MPI_Reduce return the local values.
 *
 * Parity pins (tests/test_oracle_pairs.py):
 *  - or_guu: byte-identical guuTauAlphaAct.dat against the reference
 *    Analysis program (guu.c compiled from its sources, called by
 *    oracle/ref/ref_analysis_harness.c) on the golden trajectories.
 *  - or_overlap_cage: overlap_cage_rel.c is not in Analysis/makefile's object
 *    list and uses Cell_cut, half_neibs, MAP, Neib, NeibNum, In_cell and
 *    max_part_in_cell, which Analysis/global.h does not define, so no
 *    reference run exists.  The cell grid takes MDcode/global.h:23,30
 *    (half_neibs 13, Cell_cut = 2.5 + 0.35); the neighbour order is the one
 *    GetList's traversal produces.  Pinned by an O(N^2) numpy restatement
 *    of the neighbour sets and the cage-relative displacement, and closed
 *    forms (a rigid translation of the box: Q = 1 at every lag).
 * guu.c leaves FrameNum unset when no lag lies within 1000 time units of
 * tau; it is 0 here (each origin paired with itself).
 * Build with -ffp-contract=off: expression order follows the reference.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define P_PI 3.14159265358979 /* Analysis/global.h:14 */
#define P_RHO 1.2             /* Analysis/global.h:22 */
#define P_TIME_STEP 0.005     /* Analysis/global.h:19 */
#define P_DIVIDE_TO 2.0       /* Analysis/global.h:21 */

static double wrap_loop(double v) { /* read_all_data.c LoadOrigin */
    while (v >= 1.0) v -= 1.0;
    while (v < 0.0) v += 1.0;
    return v;
}

static double mi1(double d) { /* guu.c:157-168 */
    if (d >= 0.5)
        d -= 1.0;
    else if (d < -0.5)
        d += 1.0;
    return d;
}

int or_guu(int N, double L, int nframes, const int *fidx, int gap, const double *X, const double *Y,
           const double *Z, const int *act, double factor, double tau, const char *path, double *hist3,
           int *maxbin_out, int *frame_num_out, int *n_origins_out) {
    const double bin = 0.005;
    const double MaxBin = (int)(0.5 * L / bin); /* guu.c:17 */
    const int maxbin = (int)MaxBin;
    const int maxint = 200 * gap / P_DIVIDE_TO; /* guu.c:30-31: origin_number = 200 */
    const int noi = 10 + (int)(log((double)maxint) / log(factor));
    /* guu.c:67-86: the lag ladder of the first frame */
    double *dt = calloc((size_t)noi + nframes + 64, sizeof(double));
    int ti = 1;
    {
        const int cf = fidx[0];
        int other = 0, offset = 1;
        while (other < nframes && offset < maxint) {
            const int of = cf + offset;
            while (other < nframes && fidx[other] != of) other++;
            if (other < nframes) {
                dt[ti++] = P_TIME_STEP * (double)(of - cf);
                if ((int)(offset * factor) == offset)
                    offset++;
                else
                    offset = (int)(((double)offset) * factor);
            }
        }
    }
    int frame_num = 0;
    double inter = 1000.0; /* guu.c:88-115, FrameNum[3] */
    for (int i = 1; i < ti; i++)
        if (fabs(dt[i] - tau) < inter) {
            frame_num = (int)(dt[i] * 200.0);
            inter = fabs(dt[i] - tau);
        }
    free(dt);
    double *hist = calloc(maxbin, 8), *hist1 = calloc(maxbin, 8), *hist2 = calloc(maxbin, 8);
    double *dxt = calloc(N, 8), *dyt = calloc(N, 8), *dzt = calloc(N, 8), *drt = calloc(N, 8);
    double *wx = malloc(sizeof(double) * N), *wy = malloc(sizeof(double) * N), *wz = malloc(sizeof(double) * N);
    double msd = 0.0, md = 0.0;
    int current = 0, on = 0;
    do { /* guu.c:138-231 */
        const int cf = fidx[current];
        int other = current;
        while (other < nframes && fidx[other] != cf + frame_num) other++;
        if (other >= nframes) break;
        const double *x0 = X + (size_t)current * N, *y0 = Y + (size_t)current * N, *z0 = Z + (size_t)current * N;
        const double *xt = X + (size_t)other * N, *yt = Y + (size_t)other * N, *zt = Z + (size_t)other * N;
        for (int i = 0; i < N; i++) {
            if (act[i] != 1) continue;
            double dx = mi1(xt[i] - x0[i]), dy = mi1(yt[i] - y0[i]), dz = mi1(zt[i] - z0[i]);
            double r2 = (dx * dx + dy * dy + dz * dz);
            double r = L * sqrt(r2);
            drt[i] = r;
            msd += L * L * r2;
            md += r;
            dxt[i] = L * dx;
            dyt[i] = L * dy;
            dzt[i] = L * dz;
        }
        for (int i = 0; i < N; i++) { /* LoadOrigin(origin_number): the same frame, wrapped */
            wx[i] = wrap_loop(x0[i]);
            wy[i] = wrap_loop(y0[i]);
            wz[i] = wrap_loop(z0[i]);
        }
        for (int i = 0; i < N - 1; i++) { /* guu.c:193-223 */
            if (act[i] != 1) continue;
            for (int j = i + 1; j < N; j++) {
                if (act[j] != 1) continue;
                double dx = mi1(wx[j] - wx[i]), dy = mi1(wy[j] - wy[i]), dz = mi1(wz[j] - wz[i]);
                double r2 = L * L * (dx * dx + dy * dy + dz * dz);
                int index = (int)(sqrt(r2) / bin);
                if (index < MaxBin) {
                    hist[index] += 2.0;
                    hist1[index] += 2.0 * (drt[i] * drt[j]);
                    hist2[index] += 2.0 * (dxt[i] * dxt[j] + dyt[i] * dyt[j] + dzt[i] * dzt[j]);
                }
            }
        }
        on++;
        while (current < nframes && fidx[current] != on * gap) current++;
    } while (current < nframes);
    /* guu.c:250-276 */
    double MSDt = msd / ((double)N * (double)on);
    double MDt = md / ((double)N * (double)on);
    double norm = (4.0 * P_PI * (P_RHO)*bin * (double)N * (double)on);
    double norm1 = (4.0 * P_PI * P_RHO * bin * (double)N * (double)on * (double)(MDt * MDt));
    double norm2 = (4.0 * P_PI * P_RHO * bin * (double)N * (double)on * (double)(MSDt));
    int rc = 0;
    if (path) {
        FILE *f = fopen(path, "wb");
        if (!f) {
            rc = -1;
        } else {
            for (int i = 1; i < MaxBin; i++) {
                double r = (double)(i + 1) * bin;
                fprintf(f, "%0.10g\t%.10g\t%.10g\t%.10g\t%.10g\n", r, ((double)hist[i] / (norm * r * r)),
                        ((double)hist1[i] / (norm1 * r * r)), hist2[i] / (norm1 * r * r), hist2[i] / (norm2 * r * r));
            }
            fclose(f);
        }
    }
    if (hist3) {
        memcpy(hist3, hist, sizeof(double) * maxbin);
        memcpy(hist3 + maxbin, hist1, sizeof(double) * maxbin);
        memcpy(hist3 + 2 * (size_t)maxbin, hist2, sizeof(double) * maxbin);
    }
    if (maxbin_out) *maxbin_out = maxbin;
    if (frame_num_out) *frame_num_out = frame_num;
    if (n_origins_out) *n_origins_out = on;
    free(hist); free(hist1); free(hist2);
    free(dxt); free(dyt); free(dzt); free(drt);
    free(wx); free(wy); free(wz);
    return rc;
}

/* ---------------------------------------------------- cage-relative overlap */

/* overlap_cage_rel.c:253-269: |minimum image|^2 * L * L < 1.3 * 1.3 */
static int len_check(double ax, double ay, double az, double bx, double by, double bz, double L) {
    double dx = ax - bx, dy = ay - by, dz = az - bz;
    if (dx < -0.5) dx += 1.0;
    if (dx > 0.5) dx -= 1.0;
    if (dy < -0.5) dy += 1.0;
    if (dy > 0.5) dy -= 1.0;
    if (dz < -0.5) dz += 1.0;
    if (dz > 0.5) dz -= 1.0;
    double dr2 = (dx * dx + dy * dy + dz * dz) * L * L;
    return dr2 < 1.3 * 1.3;
}

static int getcell_I(int n, int ix, int iy, int iz) { /* overlap_cage_rel.c:273-277 */
    return 1 + ((ix - 1 + n * 3) % n) + ((iy - 1 + n * 3) % n) * n + ((iz - 1 + n * 3) % n) * n * n;
}

/* the 13 half-stencil offsets of maps() (overlap_cage_rel.c:297-309) */
static const int HS[13][3] = {{1, 0, 0},  {1, 1, 0},  {0, 1, 0},  {-1, 1, 0}, {1, 0, -1}, {1, 1, -1}, {0, 1, -1},
                              {-1, 1, -1}, {1, 0, 1}, {1, 1, 1},  {0, 1, 1},  {-1, 1, 1}, {0, 0, 1}};

/* GetList (overlap_cage_rel.c:185-251) on wrapped origin positions; nb[i*cap + k]
 * in the order the traversal appends them.  Returns -1 on overflow. */
int or_cage_list(int N, double L, double cell_cut, const double *x0, const double *y0, const double *z0, int cap,
                 int *nn, int *nb) {
    const int n = (int)(L / cell_cut);
    const double cell_size = 1.0 / (double)n;
    const int ncell = n * n * n;
    int *map = malloc(sizeof(int) * 13 * (ncell + 1));
    for (int ix = 1; ix <= n; ix++)
        for (int iy = 1; iy <= n; iy++)
            for (int iz = 1; iz <= n; iz++) {
                const int c = getcell_I(n, ix, iy, iz);
                for (int s = 0; s < 13; s++)
                    map[s * (ncell + 1) + c] = getcell_I(n, ix + HS[s][0], iy + HS[s][1], iz + HS[s][2]);
            }
    int *cnt = calloc(ncell + 1, sizeof(int)), *cell = malloc(sizeof(int) * N);
    for (int i = 0; i < N; i++) {
        cell[i] = 1 + floor(x0[i] / cell_size) + floor(y0[i] / cell_size) * n + floor(z0[i] / cell_size) * n * n;
        cnt[cell[i]]++;
    }
    int *st = calloc(ncell + 2, sizeof(int));
    for (int c = 0; c <= ncell; c++) st[c + 1] = st[c] + cnt[c];
    int *mem = malloc(sizeof(int) * N), *fill = calloc(ncell + 1, sizeof(int));
    for (int i = 0; i < N; i++) mem[st[cell[i]] + fill[cell[i]]++] = i; /* index order within a cell */
    memset(nn, 0, sizeof(int) * N);
    int rc = 0;
#define ADD(a, b)                                                  \
    do {                                                           \
        if (nn[a] >= cap || nn[b] >= cap) { rc = -1; goto done; }  \
        nb[(size_t)(a) * cap + nn[a]++] = (b);                     \
        nb[(size_t)(b) * cap + nn[b]++] = (a);                     \
    } while (0)
    for (int i = 0; i < N; i++) {
        const int c = cell[i];
        for (int q = st[c]; q < st[c + 1]; q++) {
            const int k = mem[q];
            if (k > i && len_check(x0[i], y0[i], z0[i], x0[k], y0[k], z0[k], L)) ADD(i, k);
        }
        for (int s = 0; s < 13; s++) {
            const int c2 = map[s * (ncell + 1) + c];
            for (int q = st[c2]; q < st[c2 + 1]; q++) {
                const int k = mem[q];
                if (len_check(x0[i], y0[i], z0[i], x0[k], y0[k], z0[k], L)) ADD(i, k);
            }
        }
    }
#undef ADD
done:
    free(map); free(cnt); free(cell); free(st); free(mem); free(fill);
    return rc;
}

static int count_origins1(const int *fidx, int nframes, int gap) { /* overlap_cage_rel.c:23-30 */
    int current = 0, on = 0;
    do {
        while (current < nframes && fidx[current] != on * gap + 1) current++;
        if (current < nframes) on++;
    } while (current < nframes);
    return on;
}

int or_overlap_cage(int N, double L, int nframes, const int *fidx, int gap, const double *X, const double *Y,
                    const double *Z, double factor, double a, double cell_cut, const char *path, int *pair_counts,
                    int cap_pairs, int *npairs_out) {
    const double parameter = a * a;
    const int maxint = count_origins1(fidx, nframes, gap) * gap / P_DIVIDE_TO;
    const int noi = 10 + (int)(log((double)maxint) / log(factor));
    const int cap = 400; /* Neib[N][400] */
    double *c = calloc(noi, 8), *x4 = calloc(noi, 8), *delta = calloc(noi, 8);
    int *counters = calloc(noi, sizeof(int));
    int *nn = malloc(sizeof(int) * N), *nb = malloc(sizeof(int) * (size_t)N * cap);
    double *wx = malloc(sizeof(double) * N), *wy = malloc(sizeof(double) * N), *wz = malloc(sizeof(double) * N);
    double *DX = malloc(sizeof(double) * N), *DY = malloc(sizeof(double) * N), *DZ = malloc(sizeof(double) * N);
    int current = 0, on = 0, np = 0, rc = 0;
    do { /* overlap_cage_rel.c:59-154 */
        const int cf = fidx[current];
        const double *x0 = X + (size_t)current * N, *y0 = Y + (size_t)current * N, *z0 = Z + (size_t)current * N;
        for (int i = 0; i < N; i++) {
            wx[i] = wrap_loop(x0[i]);
            wy[i] = wrap_loop(y0[i]);
            wz[i] = wrap_loop(z0[i]);
        }
        if (or_cage_list(N, L, cell_cut, wx, wy, wz, cap, nn, nb) < 0) { rc = -1; break; }
        int other = current, offset = 1, t = 1;
        while (other < nframes && offset < maxint) {
            const int of = cf + offset;
            while (other < nframes && fidx[other] != of) other++;
            if (other < nframes) {
                const double *xt = X + (size_t)other * N, *yt = Y + (size_t)other * N, *zt = Z + (size_t)other * N;
                for (int i = 0; i < N; i++) { /* unwrapped difference, one correction (:81-100) */
                    DX[i] = L * mi1(xt[i] - x0[i]);
                    DY[i] = L * mi1(yt[i] - y0[i]);
                    DZ[i] = L * mi1(zt[i] - z0[i]);
                }
                double sum = 0.0;
                for (int i = 0; i < N; i++) { /* :109-131 */
                    double dx = 0.0, dy = 0.0, dz = 0.0;
                    for (int j = 0; j < nn[i]; j++) {
                        const int part = nb[(size_t)i * cap + j];
                        dx += DX[part];
                        dy += DY[part];
                        dz += DZ[part];
                    }
                    dx = dx / (double)nn[i];
                    dy = dy / (double)nn[i];
                    dz = dz / (double)nn[i];
                    double r2 = (DX[i] * DX[i] + DY[i] * DY[i] + DZ[i] * DZ[i]);
                    r2 += (dx * dx + dy * dy + dz * dz);
                    r2 -= 2.00 * (DX[i] * dx + DY[i] * dy + DZ[i] * dz);
                    if (r2 < parameter) sum += 1.0;
                }
                if (pair_counts && np < cap_pairs) pair_counts[np] = (int)sum;
                np++;
                double SUM = sum / (double)N;
                if (t >= noi) { rc = -1; break; }
                c[t] += SUM;
                x4[t] += SUM * SUM;
                counters[t]++;
                delta[t] = P_TIME_STEP * (double)(of - cf);
                if ((int)(offset * factor) == offset)
                    offset++;
                else
                    offset = (int)(((double)offset) * factor);
                t++;
            }
        }
        if (rc) break;
        on++;
        while (current < nframes && fidx[current] != on * gap) current++;
    } while (current < nframes);
    if (!rc && path) { /* :156-167 */
        FILE *f = fopen(path, "wb");
        if (!f) {
            rc = -1;
        } else {
            fprintf(f, "0.0\t1.0\t0.0\n");
            for (int k = 1; k < noi; k++)
                if (counters[k])
                    fprintf(f, "%.14lf\t%.14lf\t%.14lf\n", delta[k], c[k] / (double)counters[k],
                            x4[k] / (double)counters[k] - c[k] * c[k] / ((double)(counters[k] * counters[k])));
            fclose(f);
        }
    }
    if (npairs_out) *npairs_out = np;
    free(c); free(x4); free(delta); free(counters);
    free(nn); free(nb); free(wx); free(wy); free(wz); free(DX); free(DY); free(DZ);
    return rc;
}

/* ------------------------------------------------------- trajectory text I/O */

/* MDcode/print.c:10: one "%.5lf\t%.5lf\t%.5lf\t%d\n" line per particle with
 * glibc's printf; returns the bytes written (buf must hold 80*n). */
long long or_frame_printf(int n, const double *x, const double *y, const double *z, const int *t, char *buf) {
    char *o = buf;
    for (int i = 0; i < n; i++) o += sprintf(o, "%.5lf\t%.5lf\t%.5lf\t%d\n", x[i], y[i], z[i], t[i]);
    return (long long)(o - buf);
}

/* Analysis/read_all_data.c:81: the values fscanf("%lf %lf %lf %d") returns,
 * via strtod/strtol; returns the bytes consumed or -1. */
long long or_frame_strtod(const char *buf, int n, double *x, double *y, double *z, int *t) {
    char *e = (char *)buf;
    for (int i = 0; i < n; i++) {
        const char *s0 = e;
        x[i] = strtod(e, &e);
        y[i] = strtod(e, &e);
        z[i] = strtod(e, &e);
        t[i] = (int)strtol(e, &e, 10);
        if (e == s0) return -1;
        if (*e == '\n') e++;
    }
    return (long long)(e - buf);
}
