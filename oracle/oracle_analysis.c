/*
 * oracle_analysis.c -- CPU restatement of the reference Analysis observables
 * (TEST INFRASTRUCTURE ONLY: the checker, never the product).
 *
 * One MPI rank: NPP = N, every MPI_Reduce returns the local value.
 * Expression order follows the reference; build with -ffp-contract=off.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define OR_PI 3.14159265358979       /* Analysis/global.h:14 */
#define OR_TIME_STEP 0.005           /* Analysis/global.h:19 */
#define OR_DIVIDE_TO 2.0             /* Analysis/global.h:21 */
#define OR_NQ_MAX 162

int or_origin_gap(const int *fidx, int nframes) {
    /* Analysis/read_all_data.c:36-48 */
    int index = 0;
    while (index + 1 < nframes && fidx[index] + 1 == fidx[index + 1]) index++;
    int dummy = index;
    int consecutive = 0;
    while (consecutive < dummy) {
        if (index + 1 >= nframes) return -1;
        if (fidx[index] + 1 == fidx[index + 1])
            consecutive++;
        else
            consecutive = 0;
        index++;
    }
    return fidx[index - dummy];
}

int or_getq(int *vx, int *vy, int *vz, int *q) {
    /* Analysis/S4Qt.c:297-356 (nq = 15, Analysis/main.c:41) */
    const int nq = 15;
    int count = 1;
    q[0] = 1;
    vx[0] = 1;
    vy[0] = 0;
    vz[0] = 0;
    for (int k = 1; k < 3; k++)
        for (int i = k; i < 22; i++) {
            q[count] = k * k + i * i;
            vx[count] = 0;
            vy[count] = k;
            vz[count] = i;
            count++;
        }
    for (int i = 1; i <= nq; i++)
        for (int j = 1; j <= i; j++) {
            q[count] = i * i + j * j + j * j;
            vx[count] = i;
            vy[count] = j;
            vz[count] = j;
            count++;
        }
    count--;
    int i, j, inc, v, vec[3];
    inc = 1;
    do {
        inc *= 3;
        inc++;
    } while (inc <= count);
    do {
        inc /= 3;
        for (i = inc + 1; i <= count; i++) {
            v = q[i];
            vec[0] = vx[i];
            vec[1] = vy[i];
            vec[2] = vz[i];
            j = i;
            while (q[j - inc] > v) {
                q[j] = q[j - inc];
                vx[j] = vx[j - inc];
                vy[j] = vy[j - inc];
                vz[j] = vz[j - inc];
                j -= inc;
                if (j <= inc) break;
            }
            q[j] = v;
            vx[j] = vec[0];
            vy[j] = vec[1];
            vz[j] = vec[2];
        }
    } while (inc > 1);
    return count + 1;
}

static int degen_of(int x, int y, int z) { /* S4Qt.c:72-76 */
    if (y != z && y != x && z != x) return 6;
    if (y == z && z == x) return 1;
    return 3;
}

static int count_origins(const int *fidx, int nframes, int gap, int plus) {
    /* msd_pos.c:17-23 (plus=0), overlap.c:16-22 (plus=1) */
    int current = 0, on = 0;
    do {
        while (current < nframes && fidx[current] != on * gap + plus) current++;
        if (current < nframes) on++;
    } while (current < nframes);
    return on;
}

int or_walk(const int *fidx, int nframes, int gap, double factor, int rule,
            int cap, int *pc, int *po, int *ps, int *maxint) {
    int on = count_origins(fidx, nframes, gap, rule);
    int maximal_interval = on * gap / OR_DIVIDE_TO; /* msd_pos.c:28 */
    if (maxint) *maxint = maximal_interval;
    int current = 0, origin_number = 0, np = 0;
    do { /* msd_pos.c:44-107 */
        int current_frame = fidx[current];
        int other = current, offset = 1, time_index = 1;
        while (other < nframes && offset < maximal_interval) {
            int other_frame = current_frame + offset;
            while (other < nframes && fidx[other] != other_frame) other++;
            if (other < nframes) {
                if (np >= cap) return -1;
                pc[np] = current;
                po[np] = other;
                ps[np] = time_index;
                np++;
                if ((int)(offset * factor) == offset)
                    offset++;
                else
                    offset = (int)(((double)offset) * factor);
                time_index++;
            }
        }
        origin_number++;
        while (current < nframes && fidx[current] != origin_number * gap) current++;
    } while (current < nframes);
    return np;
}

#define WALK_ALLOC(rule)                                                        \
    int cap = nframes * 64 + 64;                                                \
    int *pc = malloc(sizeof(int) * cap), *po = malloc(sizeof(int) * cap),       \
        *ps = malloc(sizeof(int) * cap);                                        \
    int maxint = 0;                                                             \
    int np = or_walk(fidx, nframes, gap, factor, rule, cap, pc, po, ps, &maxint);

#define WALK_FREE() free(pc); free(po); free(ps);

int or_msd(int N, double L, int nframes, const int *fidx, int gap,
           const double *X, const double *Y, const double *Z, const int *TY,
           double factor, int flag, int nslots, double *m0, double *m1,
           double *m2, int *counters, double *delta_t, double *pair_vals,
           int *num_intervals) {
    WALK_ALLOC(0)
    double TotPart = (double)N;
    if (flag == 1) TotPart = 0.8 * N;
    int noi = 10 + (int)(log((double)maxint) / log(factor)) + 10; /* msd_pos.c:29 */
    if (num_intervals) *num_intervals = noi;
    for (int k = 0; k < nslots; k++) { m0[k] = 0.0; m1[k] = 0.0; m2[k] = 0.0; counters[k] = 0; delta_t[k] = 0.0; }
    for (int p = 0; p < np; p++) {
        const double *x0 = X + (size_t)pc[p] * N, *y0 = Y + (size_t)pc[p] * N, *z0 = Z + (size_t)pc[p] * N;
        const double *x = X + (size_t)po[p] * N, *y = Y + (size_t)po[p] * N, *z = Z + (size_t)po[p] * N;
        const int *ty = TY + (size_t)po[p] * N;
        double sum = 0.0, sumSq = 0.0, sumCube = 0.0;
        for (int i = 0; i < N; i++) { /* msd_pos.c:64-76 */
            if (ty[i] != 1 || flag == 0) {
                double dx = x[i] - x0[i];
                double dy = y[i] - y0[i];
                double dz = z[i] - z0[i];
                double r2 = L * L * (dx * dx + dy * dy + dz * dz);
                double r3 = r2 * r2;
                double r4 = L * L * L * L * (dx * dx * dx * dx + dy * dy * dy * dy + dz * dz * dz * dz);
                sum += r2;
                sumCube += r3;
                sumSq += r4;
            }
        }
        if (pair_vals) { pair_vals[3 * p] = sum; pair_vals[3 * p + 1] = sumSq; pair_vals[3 * p + 2] = sumCube; }
        int t = ps[p];
        if (t >= nslots) { WALK_FREE(); return -1; }
        m0[t] += sum / TotPart; /* msd_pos.c:87-93 */
        m1[t] += sumSq / TotPart;
        m2[t] += sumCube / TotPart;
        counters[t]++;
        delta_t[t] = OR_TIME_STEP * (double)(fidx[po[p]] - fidx[pc[p]]);
    }
    WALK_FREE();
    return np;
}

static inline double minimg_ana(double d) { /* overlap.c:67-78 */
    if (d >= 0.5)
        d -= 1.0;
    else if (d < -0.5)
        d += 1.0;
    return d;
}

int or_overlap(int N, double L, int nframes, const int *fidx, int gap,
               const double *X, const double *Y, const double *Z,
               const int *TY, double factor, double a, int flag, int nslots,
               double *c, double *x4, int *counters, double *delta_t,
               int *pair_count, int *num_intervals) {
    double parameter = a * a; /* overlap.c:5 */
    WALK_ALLOC(1)
    double TotPart = (double)N;
    if (flag == 1) TotPart = 0.8 * N;
    int noi = 10 + (int)(log((double)maxint) / log(factor)); /* overlap.c:29 */
    if (num_intervals) *num_intervals = noi;
    for (int k = 0; k < nslots; k++) { c[k] = 0.0; x4[k] = 0.0; counters[k] = 0; delta_t[k] = 0.0; }
    for (int p = 0; p < np; p++) {
        const double *x0 = X + (size_t)pc[p] * N, *y0 = Y + (size_t)pc[p] * N, *z0 = Z + (size_t)pc[p] * N;
        const double *x = X + (size_t)po[p] * N, *y = Y + (size_t)po[p] * N, *z = Z + (size_t)po[p] * N;
        const int *ty = TY + (size_t)po[p] * N;
        double sum = 0.0;
        int cnt = 0;
        for (int i = 0; i < N; i++) { /* overlap.c:61-84 */
            if (ty[i] != 1 || flag == 0) {
                double dx = x[i] - x0[i];
                double dy = y[i] - y0[i];
                double dz = z[i] - z0[i];
                dz = minimg_ana(dz);
                dy = minimg_ana(dy);
                dx = minimg_ana(dx);
                double r2 = L * L * (dx * dx + dy * dy + dz * dz);
                if (r2 < parameter) { sum += 1.0; cnt++; }
            }
        }
        sum = sum / TotPart;
        if (pair_count) pair_count[p] = cnt;
        int t = ps[p];
        if (t >= nslots) { WALK_FREE(); return -1; }
        c[t] += sum; /* overlap.c:89-94 */
        x4[t] += sum * sum;
        counters[t]++;
        delta_t[t] = OR_TIME_STEP * (double)(fidx[po[p]] - fidx[pc[p]]);
    }
    WALK_FREE();
    return np;
}

/* single-if wrap of S4Qt.c:102-113 / Fsqt.c:105-116 */
static inline double wrap1(double v) {
    if (v >= 1.0) v -= 1.0;
    if (v < 0.0) v += 1.0;
    return v;
}

/* q components for permutation s of a vector with degeneracy d, in the order
 * S4Qt.c:133-214 / Fsqt.c:139-221 evaluate them. */
static void perm_q(int d, int s, int a, int b, int cc, int *o) {
    static const int p3[3][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}};
    static const int p6[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    int v[3] = {a, b, cc};
    const int *p = d == 6 ? p6[s] : p3[s];
    o[0] = v[p[0]];
    o[1] = v[p[1]];
    o[2] = v[p[2]];
}

/* Phase sums of one particle at (px,py,pz) for every q, added into cq/sq
 * exactly as S4Qt.c:133-215 (S4) and Fsqt.c:139-221 (Fs) do, and the Q[n]
 * assignment side effect. */
static void add_phases(int nq, const int *vx, const int *vy, const int *vz,
                       const int *dg, double invL, double L, double px,
                       double py, double pz, double *cq, double *sq,
                       double *qval) {
    for (int n = 0; n < nq; n++) {
        int d = dg[n], o[3];
        if (d == 1) {
            double qx = 2.0 * OR_PI * invL * (double)(vx[n]);
            double qy = 2.0 * OR_PI * invL * (double)(vy[n]);
            double qz = 2.0 * OR_PI * invL * (double)(vz[n]);
            qval[n] = qx * qx + qy * qy + qz * qz;
            cq[n] = cq[n] + ((double)(cos((qx * L * px) + (qy * L * py) + (qz * L * pz))));
            sq[n] = sq[n] + ((double)(sin((qx * L * px) + (qy * L * py) + (qz * L * pz))));
        } else {
            double cosine = 0.0, sine = 0.0;
            for (int s = 0; s < d; s++) {
                perm_q(d, s, vx[n], vy[n], vz[n], o);
                double qx = 2.0 * OR_PI * invL * (double)(o[0]);
                double qy = 2.0 * OR_PI * invL * (double)(o[1]);
                double qz = 2.0 * OR_PI * invL * (double)(o[2]);
                if (s == 0 || d == 6) qval[n] = qx * qx + qy * qy + qz * qz;
                cosine += ((double)(cos((qx * L * px) + (qy * L * py) + (qz * L * pz))));
                sine += ((double)(sin((qx * L * px) + (qy * L * py) + (qz * L * pz))));
            }
            cq[n] += (cosine) / (d == 3 ? 3.0 : 6.0);
            sq[n] += (sine) / (d == 3 ? 3.0 : 6.0);
        }
    }
}

static int phase_observable(int is_s4, int N, double L, int nframes,
                            const int *fidx, int gap, const double *X,
                            const double *Y, const double *Z, const int *TY,
                            double factor, double a, int flag, int nslots,
                            double *cs, double *ss, double *cs2, double *ss2,
                            int *counters, double *delta_t, double *qval,
                            double *pair_cs, int *pair_mask) {
    double parameter = a * a;
    double invL = 1.0 / L;
    int vx[OR_NQ_MAX], vy[OR_NQ_MAX], vz[OR_NQ_MAX], qq[OR_NQ_MAX + 8], dg[OR_NQ_MAX];
    int nq = or_getq(vx, vy, vz, qq);
    for (int n = 0; n < nq; n++) dg[n] = degen_of(vx[n], vy[n], vz[n]);
    WALK_ALLOC(1)
    double *x0 = malloc(sizeof(double) * N), *y0 = malloc(sizeof(double) * N), *z0 = malloc(sizeof(double) * N);
    double cq[OR_NQ_MAX], sq[OR_NQ_MAX];
    for (int n = 0; n < nq; n++) qval[n] = 0.0;
    for (int k = 0; k < nslots; k++) {
        counters[k] = 0;
        delta_t[k] = 0.0;
        for (int n = 0; n < nq; n++) {
            cs[k * nq + n] = 0.0; ss[k * nq + n] = 0.0; cs2[k * nq + n] = 0.0; ss2[k * nq + n] = 0.0;
        }
    }
    int cur_loaded = -1;
    for (int p = 0; p < np; p++) {
        if (pc[p] != cur_loaded) { /* LoadFramePP(current,0): x0 persists across lags */
            cur_loaded = pc[p];
            memcpy(x0, X + (size_t)pc[p] * N, sizeof(double) * N);
            memcpy(y0, Y + (size_t)pc[p] * N, sizeof(double) * N);
            memcpy(z0, Z + (size_t)pc[p] * N, sizeof(double) * N);
        }
        const double *xr = X + (size_t)po[p] * N, *yr = Y + (size_t)po[p] * N, *zr = Z + (size_t)po[p] * N;
        const int *ty = TY ? TY + (size_t)po[p] * N : NULL;
        for (int n = 0; n < nq; n++) { cq[n] = 0.0; sq[n] = 0.0; }
        int nmask = 0;
        for (int i = 0; i < N; i++) {
            int use = is_s4 || ty[i] != 1 || flag == 0;
            if (!use) continue;
            x0[i] = wrap1(x0[i]); /* S4Qt.c:102-107 / Fsqt.c:105-110 */
            y0[i] = wrap1(y0[i]);
            z0[i] = wrap1(z0[i]);
            double x = wrap1(xr[i]), y = wrap1(yr[i]), z = wrap1(zr[i]);
            double dx = (x - x0[i]), dy = (y - y0[i]), dz = (z - z0[i]);
            dz = minimg_ana(dz);
            dy = minimg_ana(dy);
            dx = minimg_ana(dx);
            if (is_s4) {
                double r2 = L * L * (dx * dx + dy * dy + dz * dz); /* S4Qt.c:131-132 */
                if (r2 < parameter) {
                    nmask++;
                    add_phases(nq, vx, vy, vz, dg, invL, L, x0[i], y0[i], z0[i], cq, sq, qval);
                }
            } else {
                add_phases(nq, vx, vy, vz, dg, invL, L, dx, dy, dz, cq, sq, qval);
            }
        }
        if (pair_cs) {
            memcpy(pair_cs + (size_t)p * 2 * nq, cq, sizeof(double) * nq);
            memcpy(pair_cs + (size_t)p * 2 * nq + nq, sq, sizeof(double) * nq);
        }
        if (pair_mask) pair_mask[p] = nmask;
        int t = ps[p];
        if (t >= nslots) { np = -1; break; }
        counters[t]++; /* S4Qt.c:228-236 */
        delta_t[t] = OR_TIME_STEP * (double)(fidx[po[p]] - fidx[pc[p]]);
        for (int n = 0; n < nq; n++) {
            ss[t * nq + n] += sq[n];
            cs[t * nq + n] += cq[n];
            ss2[t * nq + n] += sq[n] * sq[n];
            cs2[t * nq + n] += cq[n] * cq[n];
        }
    }
    free(x0); free(y0); free(z0);
    WALK_FREE();
    return np;
}

int or_s4(int N, double L, int nframes, const int *fidx, int gap,
          const double *X, const double *Y, const double *Z, double factor,
          double a, int nslots, double *cs, double *ss, double *cs2,
          double *ss2, int *counters, double *delta_t, double *qval,
          double *pair_cs, int *pair_mask) {
    return phase_observable(1, N, L, nframes, fidx, gap, X, Y, Z, NULL, factor, a, 0,
                            nslots, cs, ss, cs2, ss2, counters, delta_t, qval, pair_cs, pair_mask);
}

int or_fsqt(int N, double L, int nframes, const int *fidx, int gap,
            const double *X, const double *Y, const double *Z, const int *TY,
            double factor, double a, int flag, int nslots, double *cs,
            double *ss, double *cs2, double *ss2, int *counters,
            double *delta_t, double *qval, double *pair_cs) {
    return phase_observable(0, N, L, nframes, fidx, gap, X, Y, Z, TY, factor, a, flag,
                            nslots, cs, ss, cs2, ss2, counters, delta_t, qval, pair_cs, NULL);
}

int or_write_msd(const char *path, int noi, const double *m0, const double *m1,
                 const double *m2, const int *counters, const double *delta_t) {
    FILE *f = fopen(path, "wb"); /* msd_pos.c:112-120 */
    if (!f) return -1;
    fprintf(f, "0.0\t0.0\t0.0\t0.0\n");
    for (int k = 1; k < noi; k++)
        if (counters[k])
            fprintf(f, "%.14lf\t%.14lf\t%.14lf\t%.14lf\n", delta_t[k], m0[k] / (double)counters[k],
                    m1[k] / (double)counters[k], m2[k] / (double)counters[k]);
    fclose(f);
    return 0;
}

int or_write_overlap(const char *path, int noi, const double *c, const double *x4,
                     const int *counters, const double *delta_t) {
    FILE *f = fopen(path, "wb"); /* overlap.c:108-121 */
    if (!f) return -1;
    fprintf(f, "0.0\t1.0\t0.0\n");
    for (int k = 1; k < noi; k++)
        if (counters[k])
            fprintf(f, "%.14lf\t%.14lf\t%.14lf\n", delta_t[k], c[k] / (double)counters[k],
                    x4[k] / (double)counters[k] - c[k] * c[k] / ((double)(counters[k] * counters[k])));
    fclose(f);
    return 0;
}

int or_write_s4(const char *path, int N, int nslots, const double *cs_in,
                const double *ss_in, const double *cs2_in, const double *ss2_in,
                const int *counters, const double *delta_t, const double *qval) {
    const int nq = OR_NQ_MAX;
    int vx[OR_NQ_MAX], vy[OR_NQ_MAX], vz[OR_NQ_MAX], qq[OR_NQ_MAX + 8];
    or_getq(vx, vy, vz, qq);
    FILE *f = fopen(path, "wb"); /* S4Qt.c:251-271 */
    if (!f) return -1;
    for (int k = 0; k < nslots; k++) {
        if (counters[k] != 0) {
            for (int n = 0; n < nq; n++) {
                double dg = (double)degen_of(vx[n], vy[n], vz[n]);
                double C = cs_in[k * nq + n], C2 = cs2_in[k * nq + n];
                double S = ss_in[k * nq + n], S2 = ss2_in[k * nq + n];
                C = C / (double)(counters[k]);
                C2 = C2 / (double)(counters[k]);
                C2 = dg * (C2 - C * C) / (double)(N);
                C = dg * C / (double)(N);
                S = S / (double)(counters[k]);
                S2 = S2 / (double)(counters[k]);
                S2 = dg * (S2 - S * S) / (double)(N);
                S = dg * S / (double)(N);
                fprintf(f, "%0.10g\t%0.10g\t%.10g\t%.10g\t%.10g\n", delta_t[k], qval[n], C, S, (S2 + C2));
            }
            fprintf(f, "\n\n");
        }
    }
    fclose(f);
    return 0;
}

int or_write_fsqt(const char *path, double totpart, int nslots, const double *cs_in,
                  const double *ss_in, const double *cs2_in, const double *ss2_in,
                  const int *counters, const double *delta_t, const double *qval) {
    const int nq = OR_NQ_MAX;
    FILE *f = fopen(path, "wb"); /* Fsqt.c:257-278 */
    if (!f) return -1;
    for (int n = 0; n < nq; n++) {
        for (int k = 0; k < nslots; k++) {
            if (counters[k] != 0) {
                double C = cs_in[k * nq + n], C2 = cs2_in[k * nq + n];
                double S = ss_in[k * nq + n], S2 = ss2_in[k * nq + n];
                C = C / (double)(counters[k]);
                C2 = C2 / (double)(counters[k]);
                C2 = (double)(C2 - C * C) / totpart;
                C = (double)C / totpart;
                S = S / (double)(counters[k]);
                S2 = S2 / (double)(counters[k]);
                S2 = (double)(S2 - S * S) / totpart;
                S = (double)S / totpart;
                fprintf(f, "%0.10g\t%0.10g\t%.10g\t%.10g\t%.10g\n", delta_t[k], qval[n], C, S, (S2 + C2));
            }
        }
        fprintf(f, "\n\n");
    }
    fclose(f);
    return 0;
}

int or_analyse_dir(const char *sub_dir, int N, int do_msd, int do_overlap,
                   int do_s4, int do_fs) {
    char path[1024], buff[1024];
    /* read_all_data.c:14-31 */
    snprintf(path, sizeof path, "%s/time_index_file.dat", sub_dir);
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    int nframes = 0;
    if (!fgets(buff, 1024, f) || sscanf(buff, "total timesteps equals %d", &nframes) != 1) { fclose(f); return -1; }
    int *fidx = malloc(sizeof(int) * (nframes + 10));
    for (int i = 0; i < nframes; i++) {
        if (!fgets(buff, 1024, f)) { fclose(f); free(fidx); return -1; }
        sscanf(buff, "%d", &fidx[i]);
    }
    fclose(f);
    int gap = or_origin_gap(fidx, nframes);
    /* read_all_data.c:61-91 */
    size_t tot = (size_t)(nframes + 1) * N;
    double *X = malloc(sizeof(double) * tot), *Y = malloc(sizeof(double) * tot), *Z = malloc(sizeof(double) * tot);
    int *TY = malloc(sizeof(int) * tot);
    snprintf(path, sizeof path, "%s/data_coor.out", sub_dir);
    f = fopen(path, "rb");
    if (!f) return -1;
    size_t i = 0;
    while (i < tot && fscanf(f, "%lf\t%lf\t%lf\t%d", &X[i], &Y[i], &Z[i], &TY[i]) != EOF) {
        i++;
        if ((i / N) > (size_t)nframes) break;
    }
    fclose(f);
    double L = or_ana_box_length(N);
    double factor = 1.4000000;
    int nslots = nframes + 64, noi = 0, rc = 0;
    if (do_msd) {
        double *m0 = calloc(nslots, 8), *m1 = calloc(nslots, 8), *m2 = calloc(nslots, 8), *dt = calloc(nslots, 8);
        int *cn = calloc(nslots, 4);
        if (or_msd(N, L, nframes, fidx, gap, X, Y, Z, TY, factor, 0, nslots, m0, m1, m2, cn, dt, NULL, &noi) < 0) rc = -1;
        snprintf(path, sizeof path, "%s/msd.dat", sub_dir);
        or_write_msd(path, noi, m0, m1, m2, cn, dt);
        free(m0); free(m1); free(m2); free(dt); free(cn);
    }
    if (do_overlap) {
        double *c = calloc(nslots, 8), *x4 = calloc(nslots, 8), *dt = calloc(nslots, 8);
        int *cn = calloc(nslots, 4);
        if (or_overlap(N, L, nframes, fidx, gap, X, Y, Z, TY, factor, 0.3, 0, nslots, c, x4, cn, dt, NULL, &noi) < 0) rc = -1;
        snprintf(path, sizeof path, "%s/relax_a_%.2lf.dat", sub_dir, 0.3);
        or_write_overlap(path, noi, c, x4, cn, dt);
        free(c); free(x4); free(dt); free(cn);
    }
    if (do_s4 || do_fs) {
        size_t sz = (size_t)nslots * OR_NQ_MAX;
        double *cs = calloc(sz, 8), *ss = calloc(sz, 8), *cs2 = calloc(sz, 8), *ss2 = calloc(sz, 8), *dt = calloc(nslots, 8);
        double qv[OR_NQ_MAX];
        int *cn = calloc(nslots, 4);
        if (do_s4) {
            if (or_s4(N, L, nframes, fidx, gap, X, Y, Z, factor, 0.3, nframes, cs, ss, cs2, ss2, cn, dt, qv, NULL, NULL) < 0) rc = -1;
            snprintf(path, sizeof path, "%s/S4qt_a_%.2lf.dat", sub_dir, 0.3);
            or_write_s4(path, N, nframes, cs, ss, cs2, ss2, cn, dt, qv);
        }
        if (do_fs) {
            if (or_fsqt(N, L, nframes, fidx, gap, X, Y, Z, TY, factor, 0.3, 0, nframes, cs, ss, cs2, ss2, cn, dt, qv, NULL) < 0) rc = -1;
            snprintf(path, sizeof path, "%s/Fsqt_a_%.2lf.dat", sub_dir, 0.3);
            or_write_fsqt(path, (double)N, nframes, cs, ss, cs2, ss2, cn, dt, qv);
        }
        free(cs); free(ss); free(cs2); free(ss2); free(dt); free(cn);
    }
    free(X); free(Y); free(Z); free(TY); free(fidx);
    return rc;
}

void or_chi4_ensemble(int nrep, int nlag, const double *sq, const double *sq2, const int *cnt, const int *n_r,
                      double *out3) {
    double nbar = 0.0;
    for (int r = 0; r < nrep; r++) nbar += (double)n_r[r];
    nbar /= (double)nrep;
    for (int l = 0; l < nlag; l++) {
        double s = 0.0, s2 = 0.0;
        long long c = 0;
        for (int r = 0; r < nrep; r++) {
            s += sq[(size_t)r * nlag + l];
            s2 += sq2[(size_t)r * nlag + l];
            c += cnt[(size_t)r * nlag + l];
        }
        double q = 0.0, x4 = 0.0;
        if (c > 0) { /* overlap.c:116-118 */
            double dc = (double)c;
            q = s / dc;
            x4 = s2 / dc - s * s / (dc * dc);
        }
        out3[3 * l] = nbar * x4;
        out3[3 * l + 1] = q;
        out3[3 * l + 2] = (double)c;
    }
}
