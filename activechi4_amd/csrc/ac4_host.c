/*
 * ac4_host.c -- host-side scalar bookkeeping of libac4 (C99): error buffer,
 * frame schedule, origin detection, (origin, lag) walk, ensemble chi4 and the
 * reference's output formats.  No per-particle work happens here.
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "ac4.h"
#include "ac4_host.h"

static __thread char g_err[1024];

void ac4_set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

const char *ac4_last_error(void) { return g_err; }

/* MDcode/main.c:17-41: shell sort of a[1..n] (Numerical Recipes indexing) */
static void shell_sort_1based(int n, uint64_t *a) {
    int i, j, inc = 1;
    uint64_t v;
    do {
        inc *= 3;
        inc++;
    } while (inc <= n);
    do {
        inc /= 3;
        for (i = inc + 1; i <= n; i++) {
            v = a[i];
            j = i;
            while (a[j - inc] > v) {
                a[j] = a[j - inc];
                j -= inc;
                if (j <= inc) break;
            }
            a[j] = v;
        }
    } while (inc > 1);
}

int ac4_schedule(uint64_t run_length, int n_origins, double factor, uint64_t *out, int cap) {
    if (!out || n_origins <= 0 || cap < 4 || !(factor > 1.0)) {
        ac4_set_error("ac4_schedule: bad arguments");
        return AC4_EARG;
    }
    /* MDcode/main.c:71-95: origins k*linearInterval, offsets x factor */
    uint64_t cur = 0, lin = run_length / (uint64_t)n_origins, maxint = run_length;
    for (uint64_t k = 0; k < (uint64_t)n_origins; k++) {
        if (cur + 2 >= (uint64_t)cap) { ac4_set_error("ac4_schedule: cap %d too small", cap); return AC4_ECAP; }
        out[cur++] = k * lin;
        uint64_t off = 1;
        while (off < maxint) {
            uint64_t idx = k * lin + off;
            if (idx < run_length) {
                if (cur + 2 >= (uint64_t)cap) { ac4_set_error("ac4_schedule: cap %d too small", cap); return AC4_ECAP; }
                out[cur++] = idx;
            }
            if ((uint64_t)(off * factor) == off)
                off++;
            else
                off = (uint64_t)((double)off * factor);
        }
    }
    /* shellSort(current, a) reads a[current], one word past the filled range;
     * a freshly allocated heap block holds zero there (MDcode/main.c:483). */
    for (uint64_t t = cur; t < (uint64_t)cap; t++) out[t] = 0;
    shell_sort_1based((int)cur, out);
    /* duplicate removal, MDcode/main.c:97-104 */
    uint64_t i = 0, j = 0;
    while (j < cur) {
        while (j + 1 < (uint64_t)cap && out[j] == out[i]) j++;
        i++;
        out[i] = out[j];
    }
    return (int)i;
}

int ac4_origin_gap(const int *fidx, int nframes) {
    /* Analysis/read_all_data.c:36-48 */
    if (!fidx || nframes < 2) { ac4_set_error("ac4_origin_gap: need >= 2 frames"); return AC4_EARG; }
    int index = 0;
    while (index + 1 < nframes && fidx[index] + 1 == fidx[index + 1]) index++;
    int run = index, consecutive = 0;
    while (consecutive < run) {
        if (index + 1 >= nframes) { ac4_set_error("ac4_origin_gap: no second origin run"); return AC4_EARG; }
        if (fidx[index] + 1 == fidx[index + 1])
            consecutive++;
        else
            consecutive = 0;
        index++;
    }
    return fidx[index - run];
}

int ac4_walk(const int *fidx, int nframes, int gap, double factor, int rule, ac4_walk_t *w) {
    /* origins counted on frames k*gap (+1 for rule 1): msd_pos.c:17-23, overlap.c:16-22 */
    int cur = 0, on = 0;
    do {
        while (cur < nframes && fidx[cur] != on * gap + rule) cur++;
        if (cur < nframes) on++;
    } while (cur < nframes);
    w->maximal_interval = (int)(on * gap / 2.0); /* DIVIDE_TO, msd_pos.c:28 */
    w->n_origins_counted = on;
    int cap = 64, np = 0;
    w->cur = malloc(sizeof(int) * cap);
    w->oth = malloc(sizeof(int) * cap);
    w->slot = malloc(sizeof(int) * cap);
    int current = 0, origin_number = 0, max_slot = 0;
    do { /* msd_pos.c:44-107 */
        int current_frame = fidx[current];
        int other = current, offset = 1, time_index = 1;
        while (other < nframes && offset < w->maximal_interval) {
            int other_frame = current_frame + offset;
            while (other < nframes && fidx[other] != other_frame) other++;
            if (other < nframes) {
                if (np == cap) {
                    cap *= 2;
                    w->cur = realloc(w->cur, sizeof(int) * cap);
                    w->oth = realloc(w->oth, sizeof(int) * cap);
                    w->slot = realloc(w->slot, sizeof(int) * cap);
                }
                w->cur[np] = current;
                w->oth[np] = other;
                w->slot[np] = time_index;
                if (time_index > max_slot) max_slot = time_index;
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
    w->np = np;
    w->max_slot = max_slot;
    return AC4_OK;
}

void ac4_walk_free(ac4_walk_t *w) {
    free(w->cur);
    free(w->oth);
    free(w->slot);
    memset(w, 0, sizeof *w);
}

/* Degeneracy of a q vector, Analysis/S4Qt.c:72-76 */
int ac4_degen(int x, int y, int z) {
    if (y != z && y != x && z != x) return 6;
    if (y == z && z == x) return 1;
    return 3;
}

int ac4_write_msd(const char *path, int noi, const double *m0, const double *m1, const double *m2,
                  const int *counters, const double *delta_t) {
    FILE *f = fopen(path, "wb"); /* msd_pos.c:112-120 */
    if (!f) { ac4_set_error("cannot write %s", path); return AC4_EIO; }
    fprintf(f, "0.0\t0.0\t0.0\t0.0\n");
    for (int k = 1; k < noi; k++)
        if (counters[k])
            fprintf(f, "%.14lf\t%.14lf\t%.14lf\t%.14lf\n", delta_t[k], m0[k] / (double)counters[k],
                    m1[k] / (double)counters[k], m2[k] / (double)counters[k]);
    fclose(f);
    return AC4_OK;
}

int ac4_write_overlap(const char *path, int noi, const double *c, const double *x4, const int *counters,
                      const double *delta_t) {
    FILE *f = fopen(path, "wb"); /* overlap.c:108-121 */
    if (!f) { ac4_set_error("cannot write %s", path); return AC4_EIO; }
    fprintf(f, "0.0\t1.0\t0.0\n");
    for (int k = 1; k < noi; k++)
        if (counters[k])
            fprintf(f, "%.14lf\t%.14lf\t%.14lf\n", delta_t[k], c[k] / (double)counters[k],
                    x4[k] / (double)counters[k] - c[k] * c[k] / ((double)(counters[k] * counters[k])));
    fclose(f);
    return AC4_OK;
}

int ac4_write_phase(const char *path, int is_s4, int N, double totpart, int nslots, const int *vx, const int *vy,
                    const int *vz, const double *cs, const double *ss, const double *cs2, const double *ss2,
                    const int *counters, const double *delta_t, const double *qval) {
    FILE *f = fopen(path, "wb");
    if (!f) { ac4_set_error("cannot write %s", path); return AC4_EIO; }
    const int nq = AC4_NQ;
    if (is_s4) { /* S4Qt.c:251-271: slot-major, Degen-weighted, / N */
        for (int k = 0; k < nslots; k++) {
            if (counters[k] == 0) continue;
            for (int n = 0; n < nq; n++) {
                double dg = (double)ac4_degen(vx[n], vy[n], vz[n]);
                double C = cs[k * nq + n] / (double)(counters[k]);
                double C2 = cs2[k * nq + n] / (double)(counters[k]);
                C2 = dg * (C2 - C * C) / (double)(N);
                C = dg * C / (double)(N);
                double S = ss[k * nq + n] / (double)(counters[k]);
                double S2 = ss2[k * nq + n] / (double)(counters[k]);
                S2 = dg * (S2 - S * S) / (double)(N);
                S = dg * S / (double)(N);
                fprintf(f, "%0.10g\t%0.10g\t%.10g\t%.10g\t%.10g\n", delta_t[k], qval[n], C, S, (S2 + C2));
            }
            fprintf(f, "\n\n");
        }
    } else { /* Fsqt.c:257-278: q-major, / TotPart */
        for (int n = 0; n < nq; n++) {
            for (int k = 0; k < nslots; k++) {
                if (counters[k] == 0) continue;
                double C = cs[k * nq + n] / (double)(counters[k]);
                double C2 = cs2[k * nq + n] / (double)(counters[k]);
                C2 = (double)(C2 - C * C) / totpart;
                C = (double)C / totpart;
                double S = ss[k * nq + n] / (double)(counters[k]);
                double S2 = ss2[k * nq + n] / (double)(counters[k]);
                S2 = (double)(S2 - S * S) / totpart;
                S = (double)S / totpart;
                fprintf(f, "%0.10g\t%0.10g\t%.10g\t%.10g\t%.10g\n", delta_t[k], qval[n], C, S, (S2 + C2));
            }
            fprintf(f, "\n\n");
        }
    }
    fclose(f);
    return AC4_OK;
}

/* Analysis/S4Qt.c:297-356 (nq = 15): the 162 lattice vectors sorted on |q|^2
 * by the same shell sort (indices 1..161, so its tie order is reproduced). */
int ac4_getq(int *vx, int *vy, int *vz, int *q) {
    int count = 1;
    q[0] = 1; vx[0] = 1; vy[0] = 0; vz[0] = 0;
    for (int k = 1; k < 3; k++)
        for (int i = k; i < 22; i++) { q[count] = k * k + i * i; vx[count] = 0; vy[count] = k; vz[count] = i; count++; }
    for (int i = 1; i <= 15; i++)
        for (int j = 1; j <= i; j++) { q[count] = i * i + j * j + j * j; vx[count] = i; vy[count] = j; vz[count] = j; count++; }
    count--;
    int inc = 1;
    do { inc *= 3; inc++; } while (inc <= count);
    do {
        inc /= 3;
        for (int i = inc + 1; i <= count; i++) {
            int v = q[i], a = vx[i], b = vy[i], c = vz[i], j = i;
            while (q[j - inc] > v) {
                q[j] = q[j - inc]; vx[j] = vx[j - inc]; vy[j] = vy[j - inc]; vz[j] = vz[j - inc];
                j -= inc;
                if (j <= inc) break;
            }
            q[j] = v; vx[j] = a; vy[j] = b; vz[j] = c;
        }
    } while (inc > 1);
    return count + 1;
}

/* ---- seeded initial state (replaces MDcode/initialise.c:34-101 and
 * active.c:4-43, whose streams come from time() and glibc rand()) ---- */
typedef struct { uint32_t k0, k1; uint64_t ctr; uint32_t buf[4]; int left; } ac4_rng;

static void ac4_philox_host(const uint32_t c_in[4], const uint32_t k_in[2], uint32_t out[4]);

static uint32_t rng_u32(ac4_rng *g) {
    if (g->left == 0) {
        uint32_t c[4] = {(uint32_t)g->ctr, (uint32_t)(g->ctr >> 32), 0x696e6974u, 0u}, k[2] = {g->k0, g->k1};
        ac4_philox_host(c, k, g->buf);
        g->ctr++;
        g->left = 4;
    }
    return g->buf[4 - g->left--];
}
static double rng_u01(ac4_rng *g) {
    uint32_t a = rng_u32(g), b = rng_u32(g);
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}
static void rng_perm(ac4_rng *g, int *p, int n) { /* uniform Fisher-Yates */
    for (int i = 0; i < n; i++) p[i] = i;
    for (int i = n - 1; i > 0; i--) {
        int j = (int)(rng_u01(g) * (double)(i + 1));
        int t = p[i]; p[i] = p[j]; p[j] = t;
    }
}

static void ac4_philox_host(const uint32_t c_in[4], const uint32_t k_in[2], uint32_t out[4]) {
    /* Philox-4x32-10, Salmon et al. SC 2011 (same as the device version) */
    uint32_t c0 = c_in[0], c1 = c_in[1], c2 = c_in[2], c3 = c_in[3], k0 = k_in[0], k1 = k_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0, hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0; c1 = lo1; c2 = hi0 ^ c3 ^ k1; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

int ac4_init_lattice(int n, double T, int n_act, uint64_t seed, double *x, double *y, double *z, double *xu,
                     double *yu, double *zu, double *px, double *py, double *pz, int *type, int *act_tag, int *kx,
                     int *ky, int *kz) {
    if (n < 2 || !(T > 0) || !x || !y || !z || !px || !py || !pz || !type) {
        ac4_set_error("ac4_init_lattice: bad arguments");
        return AC4_EARG;
    }
    ac4_rng g = {(uint32_t)seed, (uint32_t)(seed >> 32), 0, {0, 0, 0, 0}, 0};
    int t = 1;
    while (t * t * t < n) t++; /* initialise.c:35-37 */
    int nsite = t * t * t;
    int *perm = malloc(sizeof(int) * (size_t)(nsite > n ? nsite : n));
    if (!perm) { ac4_set_error("ac4_init_lattice: out of memory"); return AC4_EARG; }
    rng_perm(&g, perm, nsite);
    double dr = 1.0 / (double)t;
    for (int i = 0; i < n; i++) { /* simple-cubic sites, chosen at random */
        int s = perm[i];
        x[i] = dr * (double)(s % t);
        y[i] = dr * (double)((s / t) % t);
        z[i] = dr * (double)(s / (t * t));
        if (xu) { xu[i] = x[i]; yu[i] = y[i]; zu[i] = z[i]; }
    }
    int na = (int)(.80 * (double)n); /* global.h:16 */
    rng_perm(&g, perm, n);
    for (int i = 0; i < n; i++) type[perm[i]] = i < na ? 0 : 1;
    /* Maxwell-Boltzmann momenta (Box-Muller), zero drift, T = sum p^2 / DOF */
    double sx = 0, sy = 0, sz = 0;
    for (int i = 0; i < n; i++) {
        double v[3];
        for (int k = 0; k < 3; k++) {
            double u1 = rng_u01(&g), u2 = rng_u01(&g);
            v[k] = sqrt(-2.0 * log(1.0 - u1)) * cos(2.0 * 3.14159265358979323846 * u2);
        }
        px[i] = v[0]; py[i] = v[1]; pz[i] = v[2];
        sx += v[0]; sy += v[1]; sz += v[2];
    }
    double p2 = 0.0;
    for (int i = 0; i < n; i++) {
        px[i] -= sx / n; py[i] -= sy / n; pz[i] -= sz / n;
        p2 += px[i] * px[i] + py[i] * py[i] + pz[i] * pz[i];
    }
    double scale = sqrt(T * (3.0 * (double)n - 3.0) / p2);
    for (int i = 0; i < n; i++) { px[i] *= scale; py[i] *= scale; pz[i] *= scale; }
    /* active particles and initial directions, MDcode/active.c:24-41 */
    if (n_act < 0) n_act = (int)((double)n * 0.1000);
    if (act_tag) {
        rng_perm(&g, perm, n);
        for (int i = 0; i < n; i++) act_tag[i] = -1;
        for (int a = 0; a < n_act; a++) act_tag[perm[a]] = a;
    }
    if (kx && ky && kz)
        for (int a = 0; a < n_act; a++) {
            int v = a < n_act / 2 ? 1 : -1;
            kx[a] = v; ky[a] = v; kz[a] = v;
        }
    free(perm);
    return AC4_OK;
}

/* InitialiseNoseHoover (MDcode/time_evolve.c:476-493): eta = 0, chain masses
 * Q0 = DOF*T*tau^2, Q1 = Q2 = T*tau^2 (tau = 0.2), p_eta_k ~ N(0, T*Q_k).  The
 * reference draws p_eta from a time-seeded GSL stream; here the draw is a
 * xorshift64 stream (Marsaglia 2003) from `seed` through Box-Muller, so a
 * job's chain is reproducible from its seed. */
void ac4_chain_init(int n, double T, uint64_t seed, double *c9) {
    const double tq = .20;
    for (int k = 0; k < 3; k++) { c9[k] = 0.0; c9[6 + k] = T * tq * tq; }
    c9[6] = (3.0 * n - 3.0) * T * tq * tq;
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 1;
    for (int k = 0; k < 3; k++) {
        double u[2];
        for (int m = 0; m < 2; m++) {
            st ^= st >> 12; st ^= st << 25; st ^= st >> 27;
            u[m] = (double)((st * 2685821657736338717ull) >> 11) * (1.0 / 9007199254740992.0);
        }
        c9[3 + k] = sqrt(T * c9[6 + k]) * sqrt(-2.0 * log(1.0 - u[0])) * cos(2.0 * 3.14159265358979 * u[1]);
    }
}

/* ---------------------------------------------------- trajectory text I/O
 * data_coor.out lines are "%.5lf\t%.5lf\t%.5lf\t%d\n" (MDcode/print.c:4-16).
 * fmt5 prints a double exactly as glibc's %.5lf: the value m*2^e is
 * multiplied by 10^5 in 128-bit integers and rounded to nearest with ties to
 * even (the correctly rounded decimal that printf produces), the sign kept
 * for values that round to zero.  Non-finite or huge values (never in a
 * trajectory) go through snprintf. */
static char *put_u64(char *o, uint64_t v) {
    char t[24];
    int k = 0;
    do { t[k++] = (char)('0' + v % 10); v /= 10; } while (v);
    while (k) *o++ = t[--k];
    return o;
}

static char *fmt5(char *o, double v) {
    uint64_t bits;
    memcpy(&bits, &v, sizeof bits);
    const int neg = (int)(bits >> 63), ex = (int)((bits >> 52) & 0x7ff);
    const uint64_t frac = bits & ((1ull << 52) - 1);
    if (ex == 0x7ff || fabs(v) >= 1e12) return o + sprintf(o, "%.5lf", v);
    const uint64_t m = ex ? (frac | (1ull << 52)) : frac;
    const int e = ex ? ex - 1075 : -1074; /* v = m * 2^e */
    unsigned __int128 q = 0;
    if (m != 0) {
        const unsigned __int128 t = (unsigned __int128)m * 100000u; /* < 2^70 */
        if (e >= 0) {
            q = t << e; /* |v| < 1e12: e <= 0 in practice */
        } else if (-e <= 71) {
            const int sh = -e;
            const unsigned __int128 one = 1;
            const unsigned __int128 rem = t & ((one << sh) - 1), half = one << (sh - 1);
            q = t >> sh;
            if (rem > half || (rem == half && (q & 1))) q++;
        } /* else t / 2^sh < 1/2: rounds to 0 */
    }
    if (neg) *o++ = '-';
    o = put_u64(o, (uint64_t)(q / 100000u));
    *o++ = '.';
    uint64_t fp = (uint64_t)(q % 100000u);
    for (int d = 4; d >= 0; d--) { o[d] = (char)('0' + fp % 10); fp /= 10; }
    return o + 5;
}

long long ac4_format_frame(int n, const double *x, const double *y, const double *z, const int *type, char *out,
                           long long cap) {
    if (n < 0 || !x || !y || !z || !type || !out || cap < 80LL * n) {
        ac4_set_error("format_frame: bad arguments (buffer needs 80 bytes a line)");
        return AC4_EARG;
    }
    /* A line of values below 1e12 in magnitude takes at most 72 bytes.  A
     * frame holding larger or non-finite values (a run that blew up) is
     * measured first: %.5lf of such a value can take over 300 bytes. */
    int wide = 0;
    for (int i = 0; i < n && !wide; i++)
        wide = !(fabs(x[i]) < 1e12 && fabs(y[i]) < 1e12 && fabs(z[i]) < 1e12);
    if (wide) {
        long long need = 0;
        for (int i = 0; i < n; i++)
            need += snprintf(NULL, 0, "%.5lf\t%.5lf\t%.5lf\t%d\n", x[i], y[i], z[i], type[i]);
        if (need > cap) {
            ac4_set_error("format_frame: the frame holds values of 1e12 or more and needs %lld bytes (buffer %lld)",
                          need, cap);
            return AC4_ECAP;
        }
    }
    char *o = out;
    for (int i = 0; i < n; i++) {
        o = fmt5(o, x[i]);
        *o++ = '\t';
        o = fmt5(o, y[i]);
        *o++ = '\t';
        o = fmt5(o, z[i]);
        *o++ = '\t';
        if (type[i] < 0) { *o++ = '-'; o = put_u64(o, (uint64_t)(-(long long)type[i])); }
        else o = put_u64(o, (uint64_t)type[i]);
        *o++ = '\n';
    }
    return (long long)(o - out);
}

/* One number of a data_coor.out line: "[-]digits.digits" with at most 15
 * significant digits is D / 10^f with D < 2^53, and (double)D / 10^f is one
 * correctly rounded division, i.e. the value strtod returns; anything else
 * (exponents, long mantissas) goes through strtod. */
static const char *parse_num(const char *p, double *out) {
    static const double p10[16] = {1e0, 1e1, 1e2, 1e3, 1e4, 1e5, 1e6, 1e7, 1e8, 1e9, 1e10, 1e11, 1e12, 1e13, 1e14, 1e15};
    const char *s = p;
    while (*s == ' ' || *s == '\t') s++;
    const char *st = s;
    int neg = 0, nd = 0, nf = 0;
    if (*s == '-' || *s == '+') { neg = *s == '-'; s++; }
    uint64_t d = 0;
    while (*s >= '0' && *s <= '9' && nd < 16) { d = d * 10 + (uint64_t)(*s - '0'); s++; nd++; }
    if (*s == '.') {
        s++;
        while (*s >= '0' && *s <= '9' && nd < 16) { d = d * 10 + (uint64_t)(*s - '0'); s++; nd++; nf++; }
    }
    if (nd == 0 || nd > 15 || (*s >= '0' && *s <= '9') || *s == 'e' || *s == 'E') {
        char *e;
        *out = strtod(st, &e);
        return e == st ? NULL : e;  /* NULL: no number here */
    }
    const double v = (double)d / p10[nf];
    *out = neg ? -v : v;
    return s;
}

long long ac4_parse_frame(const char *text, long long len, int n, double *x, double *y, double *z, int *type) {
    if (!text || len < 0 || n < 0 || !x || !y || !z || !type) {
        ac4_set_error("parse_frame: bad arguments");
        return AC4_EARG;
    }
    /* each line is parsed from a NUL-terminated copy, so no conversion reads
     * past the line or past text + len */
    char line[1024];
    const char *p = text, *end = text + len;
    for (int i = 0; i < n; i++) {
        if (p >= end) { ac4_set_error("parse_frame: text ends at line %d", i); return AC4_EIO; }
        const char *nl = memchr(p, '\n', (size_t)(end - p));
        const size_t m = (size_t)((nl ? nl : end) - p);
        if (m >= sizeof line) { ac4_set_error("parse_frame: line %d is longer than 1023 bytes", i); return AC4_EIO; }
        memcpy(line, p, m);
        line[m] = '\0';
        const char *q = line;
        if (q) q = parse_num(q, &x[i]);
        if (q) q = parse_num(q, &y[i]);
        if (q) q = parse_num(q, &z[i]);
        char *e = NULL;
        if (q) {
            type[i] = (int)strtol(q, &e, 10);
            if (e == q) q = NULL;
        }
        if (!q) { ac4_set_error("parse_frame: line %d is not \"x y z type\"", i); return AC4_EIO; }
        p = nl ? nl + 1 : end;
    }
    return (long long)(p - text);
}
