/*
 * ac4_tune.c -- autotuner for the batched MD step.  No reference counterpart
 * (the reference has one fixed CPU code path).  The replica set is brought to
 * temperature once; every candidate kernel shape then starts from that same
 * state, replays `steps` graph steps, and the fastest shape is kept for the
 * handles created afterwards (ac4_set_tuning).  Candidates: force-kernel lanes
 * per particle tpp in {1, 2, 4, 8, 16} (so 256/tpp particles per workgroup)
 * crossed with the thread-per-particle and 16/32/64-lane-team list builders,
 * then the Verlet skin {0.30, 0.35, 0.45, 0.55} for the best of those (the
 * skin sets the list radii, the rebuild interval and the list-cell size),
 * then the replicas per launch: the set split into G in {1, 2, 3, 4} launch
 * groups of balanced particle count (ac4_group_plan), one handle and stream
 * per group, stepped together (ac4_md_step_graph_multi) so that one group's
 * kernels fill the gaps of another's.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "ac4.h"
#include "ac4_host.h"

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

typedef struct {
    int n, nact;
    double *d;   /* x y z xu yu zu px py pz (9n) */
    int *i;      /* type act_tag (2n), kx ky kz (3(nact+1)) */
    double c9[9];
} rep_state;

static int load_all(ac4_md *md, int n_rep, rep_state *st) {
    for (int r = 0; r < n_rep; r++) {
        const int n = st[r].n, na = st[r].nact + 1;
        double *d = st[r].d;
        int *q = st[r].i;
        int rc = ac4_md_set_state(md, r, d, d + n, d + 2 * n, d + 3 * n, d + 4 * n, d + 5 * n, d + 6 * n,
                                  d + 7 * n, d + 8 * n, q, q + n, q + 2 * n, q + 2 * n + na, q + 2 * n + 2 * na);
        if (rc == AC4_OK) rc = ac4_md_set_chain(md, r, st[r].c9);
        if (rc != AC4_OK) return rc;
    }
    return AC4_OK;
}

static void free_states(rep_state *st, int n_rep) {
    if (!st) return;
    for (int r = 0; r < n_rep; r++) { free(st[r].d); free(st[r].i); }
    free(st);
}

int ac4_group_plan(int n_rep, const ac4_replica_params *p, int groups, int *group_of) {
    if (n_rep < 1 || !p || groups < 1 || groups > n_rep || !group_of) {
        ac4_set_error("ac4_group_plan: bad arguments");
        return AC4_EARG;
    }
    int *ord = malloc(sizeof(int) * (size_t)n_rep);
    long long *load = calloc((size_t)groups, sizeof(long long));
    if (!ord || !load) { free(ord); free(load); ac4_set_error("ac4_group_plan: out of host memory"); return AC4_ECAP; }
    for (int r = 0; r < n_rep; r++) ord[r] = r;
    for (int a = 1; a < n_rep; a++) { /* largest N first, lower index first among equals */
        int v = ord[a], k = a - 1;
        while (k >= 0 && (p[ord[k]].n < p[v].n || (p[ord[k]].n == p[v].n && ord[k] > v))) { ord[k + 1] = ord[k]; k--; }
        ord[k + 1] = v;
    }
    for (int a = 0; a < n_rep; a++) {
        int best = 0;
        for (int g = 1; g < groups; g++)
            if (load[g] < load[best]) best = g;
        group_of[ord[a]] = best;
        load[best] += p[ord[a]].n;
    }
    free(ord);
    free(load);
    return AC4_OK;
}

/* particle-steps/s of `steps` graph steps of the replica set split into
 * `groups` launch groups (one handle each, stepped together) */
static double time_shape(int n_rep, const ac4_replica_params *p, int precision, int steps, rep_state *st,
                         long long nsum, int tpp, int mode, double skin, int groups) {
    ac4_md *md[4] = {NULL, NULL, NULL, NULL};
    double rate = 0.0;
    int *gof = malloc(sizeof(int) * (size_t)n_rep), *idx = malloc(sizeof(int) * (size_t)n_rep);
    ac4_replica_params *gp = malloc(sizeof(ac4_replica_params) * (size_t)n_rep);
    rep_state *gs = malloc(sizeof(rep_state) * (size_t)n_rep);
    int ok = gof && idx && gp && gs && groups >= 1 && groups <= 4 && groups <= n_rep &&
             ac4_group_plan(n_rep, p, groups, gof) == AC4_OK && ac4_set_tuning(tpp, mode) == AC4_OK &&
             ac4_set_skin(skin) == AC4_OK;
    for (int g = 0; g < groups && ok; g++) {
        int m = 0;
        for (int r = 0; r < n_rep; r++)
            if (gof[r] == g) { gp[m] = p[r]; gs[m] = st[r]; m++; }
        ok = ac4_md_create(m, gp, precision, &md[g]) == AC4_OK && load_all(md[g], m, gs) == AC4_OK;
    }
    if (ok) ok = ac4_md_step_graph_multi(md, groups, 20) == AC4_OK;
    for (int g = 0; g < groups && ok; g++) ok = ac4_md_sync(md[g]) == AC4_OK;
    if (ok) {
        const double t0 = now_s();
        ok = ac4_md_step_graph_multi(md, groups, steps) == AC4_OK;
        for (int g = 0; g < groups && ok; g++) ok = ac4_md_sync(md[g]) == AC4_OK;  // sync reports list overflow
        const double t1 = now_s();
        if (ok && t1 > t0) rate = (double)nsum * steps / (t1 - t0);
    }
    for (int g = 0; g < 4; g++)
        if (md[g]) ac4_md_destroy(md[g]);
    free(gof); free(idx); free(gp); free(gs);
    return rate;
}

int ac4_autotune(int n_rep, const ac4_replica_params *p, int precision, int relax_steps, int steps, double *best4,
                 double *table, int table_cap, int *n_cand) {
    if (n_rep < 1 || !p || relax_steps < 1 || steps < 1) {
        ac4_set_error("ac4_autotune: bad arguments");
        return AC4_EARG;
    }
    static const int tpps[5] = {1, 2, 4, 8, 16}, modes[4] = {1, 2, 3, 4};
    static const double skins[4] = {0.30, 0.35, 0.45, 0.55};
    rep_state *st = calloc((size_t)n_rep, sizeof *st);
    ac4_md *md = NULL;
    int rc = ac4_set_tuning(0, 3);
    if (rc == AC4_OK) rc = ac4_set_skin(0.35);
    if (rc == AC4_OK) rc = ac4_md_create(n_rep, p, precision, &md);
    long long nsum = 0;
    for (int r = 0; r < n_rep && rc == AC4_OK; r++) {
        const int n = p[r].n, nact = p[r].n_act >= 0 ? p[r].n_act : (int)((double)n * 0.1000);
        st[r].n = n;
        st[r].nact = nact;
        nsum += n;
        st[r].d = malloc(sizeof(double) * 9 * (size_t)n);
        st[r].i = malloc(sizeof(int) * (2 * (size_t)n + 3 * (size_t)(nact + 1)));
        double *d = st[r].d;
        int *q = st[r].i, na = nact + 1;
        rc = ac4_init_lattice(n, p[r].T, nact, p[r].seed, d, d + n, d + 2 * n, d + 3 * n, d + 4 * n, d + 5 * n,
                              d + 6 * n, d + 7 * n, d + 8 * n, q, q + n, q + 2 * n, q + 2 * n + na,
                              q + 2 * n + 2 * na);
        ac4_chain_init(n, p[r].T, p[r].seed, st[r].c9);
    }
    if (rc == AC4_OK) rc = load_all(md, n_rep, st);
    if (rc == AC4_OK) rc = ac4_md_relax(md, relax_steps, 1.5, 1.0);
    for (int r = 0; r < n_rep && rc == AC4_OK; r++) {
        const int n = st[r].n, na = st[r].nact + 1;
        double *d = st[r].d;
        int *q = st[r].i;
        rc = ac4_md_get_state(md, r, d, d + n, d + 2 * n, d + 3 * n, d + 4 * n, d + 5 * n, d + 6 * n, d + 7 * n,
                              d + 8 * n, q + 2 * n, q + 2 * n + na, q + 2 * n + 2 * na);
    }
    ac4_md_destroy(md);
    if (rc != AC4_OK) {
        free_states(st, n_rep);
        ac4_set_tuning(0, 3);
        ac4_set_skin(0.35);
        return rc;
    }
    double best = -1.0, bs = 0.35;
    int bt = 0, bm = 3, bg = 1, nc = 0;
#define RECORD(T_, M_, S_, G_, V_)                                                    \
    do {                                                                               \
        if (table && nc < table_cap) {                                                 \
            table[5 * nc] = (T_); table[5 * nc + 1] = (M_);                            \
            table[5 * nc + 2] = (S_); table[5 * nc + 3] = (G_); table[5 * nc + 4] = (V_); \
        }                                                                              \
        nc++;                                                                          \
    } while (0)
    /* stage 1: workgroup shape and list builder at the reference skin */
    for (int a = 0; a < 5; a++)
        for (int b = 0; b < 4; b++) {
            const double rate = time_shape(n_rep, p, precision, steps, st, nsum, tpps[a], modes[b], 0.35, 1);
            RECORD(tpps[a], modes[b], 0.35, 1, rate);
            if (rate > best) { best = rate; bt = tpps[a]; bm = modes[b]; }
        }
    /* stage 2: the skin (list radius and cell size) for that shape */
    for (int k = 0; k < 4; k++) {
        if (skins[k] == 0.35) continue;
        const double rate = time_shape(n_rep, p, precision, steps, st, nsum, bt, bm, skins[k], 1);
        RECORD(bt, bm, skins[k], 1, rate);
        if (rate > best) { best = rate; bs = skins[k]; }
    }
    /* stage 3: replicas per launch (launch groups stepped together) */
    for (int g = 2; g <= 4 && g <= n_rep; g++) {
        const double rate = time_shape(n_rep, p, precision, steps, st, nsum, bt, bm, bs, g);
        RECORD(bt, bm, bs, g, rate);
        if (rate > best) { best = rate; bg = g; }
    }
#undef RECORD
    free_states(st, n_rep);
    if (n_cand) *n_cand = nc;
    if (best <= 0.0) {
        ac4_set_tuning(0, 3);
        ac4_set_skin(0.35);
        ac4_set_error("ac4_autotune: no candidate ran");
        return AC4_EHIP;
    }
    if (best4) { best4[0] = bt; best4[1] = bm; best4[2] = bs; best4[3] = bg; }
    rc = ac4_set_tuning(bt, bm);
    return rc == AC4_OK ? ac4_set_skin(bs) : rc;
}
