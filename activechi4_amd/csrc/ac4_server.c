/*
 * ac4_server.c -- request loop for MD jobs on one GPU (continuous batching).
 * No reference counterpart: the reference runs one job per mpirun
 * (MDcode/readme:1-11).  Jobs (N, T, f0, tau_p, steps, seed) arrive over
 * time; the server keeps a fixed set of device-resident slots, admits queued
 * jobs into a free slot of their size class (FIFO order; a later job may take
 * a free slot of its own class while an earlier one waits), advances every
 * slot together in chunks of `chunk` steps with replayed hipGraphs, and
 * retires finished jobs.
 *
 * - The slots are split into launch groups of balanced capacity, one MD
 *   handle and stream each, stepped together (ac4_md_step_graph_multi: one
 *   group's force kernel runs beside another's list rebuild).  Every group
 *   uses the kernel shape of the server's total capacity, so a job's
 *   arithmetic does not depend on the group its slot is in.
 *
 * - Slots come from an HBM byte budget (ac4_server_plan): size classes a
 *   factor 2 apart over the request range [n_min, n_max], slot counts per
 *   class in proportion to the share of a log-uniform N distribution each
 *   class takes, as many as the budget holds at ac4_md_slot_bytes per slot.
 * - Frames (one every chunk) stay on the device: each job's overlap Q(t) is
 *   accumulated in place (ac4_md_overlap_*, keyed on the job's own step
 *   counter) and copied out with the final x_u when the job finishes.  No
 *   per-chunk copy of the slots' positions.
 * - A replica whose error flag is set (list or cell over capacity, runaway
 *   coordinate) fails alone: its job is marked failed and its slot freed and
 *   cleared, the other jobs keep stepping.
 * Admission never changes a running replica's arithmetic (slot-local loads),
 * and a job's results do not depend on the other jobs in the batch.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "ac4.h"
#include "ac4_host.h"

int ac4_md_load_slot(ac4_md *md, int r, const double *x, const double *y, const double *z, const double *xu,
                     const double *yu, const double *zu, const double *px, const double *py, const double *pz,
                     const int *type, const int *act_tag, const int *kx, const int *ky, const int *kz);

typedef struct {
    ac4_job job;
    int state; /* 0 queued, 1 running, 2 done, 3 failed */
    int slot, frames, preemptions, err;
    long steps_done, target;
    double t_start, t_done;
    double *result;  /* final x_u,y_u,z_u (3n) */
    double *ov;      /* per-lag {sum Q, sum Q^2, origins} of the job (3*(n_lags+1)) */
    char *ckpt;      /* exact checkpoint of a preempted job, resumed on admission */
    long long ckpt_len;
} job_rec;

struct ac4_server {
    ac4_md **mds; /* launch groups: one handle and stream each */
    int ng;
    int nslots, chunk, relax; /* relax: ramp steps rounded up to whole chunks */
    int *cap, *slot_job, *err;
    int *slot_h, *slot_l;     /* slot -> its group, and its index in that group's handle */
    int *gerr;                /* error bits of one group's slots */
    job_rec *jobs;
    int njobs, jcap;
    int ov_on, ov_nl;         /* in-situ overlap: on, lags + 1 */
    long long particle_steps;
};

/* Slot classes a factor 2 apart from the first power of two >= n_min up to
 * n_max (the last class n_max rounded up to 256).  Class k holds the N in
 * (c_{k-1}, c_k] and gets a slot count in proportion to
 * ln(min(c_k, n_max) / max(c_{k-1}, n_min)), the share of a log-uniform N
 * over [n_min, n_max] it takes (at least one slot); the total is the largest
 * that fits hbm_bytes at ac4_md_slot_bytes per slot. */
int ac4_server_plan(double hbm_bytes, int n_min, int n_max, int n_origin, int n_lags, int *caps, int caps_cap) {
    if (!(hbm_bytes > 0) || n_min < 2 || n_max < n_min || n_origin < 0 || n_lags < 0 || !caps || caps_cap < 1) {
        ac4_set_error("ac4_server_plan: bad arguments");
        return AC4_EARG;
    }
    int cls[40], nc = 0;
    long long c = 256;
    while (c < n_min) c *= 2;
    while (c < n_max && nc < 39) { cls[nc++] = (int)c; c *= 2; }
    cls[nc++] = ((n_max + 255) / 256) * 256;
    double share[40], bytes[40], tot_share = 0.0;
    for (int k = 0; k < nc; k++) {
        const double lo = k ? fmax((double)cls[k - 1], (double)n_min) : (double)n_min;
        const double hi = fmin((double)cls[k], (double)n_max);
        share[k] = hi > lo ? log(hi / lo) : 0.0;
        tot_share += share[k];
        long long b = ac4_md_slot_bytes(cls[k], n_origin, n_lags);
        if (b < 0) return (int)b;
        bytes[k] = (double)b;
    }
    /* largest total M whose per-class counts max(1, round(M*share)) fit */
    int best[40], have = 0;
    for (int m = 1; m <= 100000; m++) {
        int cnt[40];
        double need = 0.0;
        int slots = 0;
        for (int k = 0; k < nc; k++) {
            cnt[k] = (int)floor((double)m * share[k] / tot_share + 0.5);
            if (cnt[k] < 1) cnt[k] = 1;
            need += cnt[k] * bytes[k];
            slots += cnt[k];
        }
        if (need > hbm_bytes || slots > caps_cap) break;
        memcpy(best, cnt, sizeof(int) * nc);
        have = 1;
    }
    if (!have) {
        ac4_set_error("ac4_server_plan: %.3g bytes hold no slot of every class (a %d-particle slot takes %.3g)",
                      hbm_bytes, cls[nc - 1], bytes[nc - 1]);
        return AC4_ECAP;
    }
    int ns = 0;
    for (int k = 0; k < nc; k++)
        for (int q = 0; q < best[k]; q++) caps[ns++] = cls[k];
    return ns;
}

/* Launch groups when the caller leaves it to the server: SRV_GROUPS groups
 * once the slots hold at least SRV_GROUP_MIN particles per group (on the
 * 64-replica ensemble three groups of about 51k particles ran 1.57e9
 * particle-steps/s against 1.42e9 for one handle; a fourth stream exceeds
 * the process's four hardware queues), else one. */
#define SRV_GROUPS 3
#define SRV_GROUP_MIN 16384
#define SRV_GROUPS_MAX 64

int ac4_server_create_groups(int nslots, const int *slot_cap, int precision, int chunk, int relax_steps,
                             int groups, ac4_server **out) {
    if (nslots < 1 || !slot_cap || chunk < 1 || relax_steps < 1 || groups < 0 || groups > SRV_GROUPS_MAX || !out) {
        ac4_set_error("ac4_server_create: bad arguments (groups in 0..%d)", SRV_GROUPS_MAX);
        return AC4_EARG;
    }
    ac4_server *s = calloc(1, sizeof *s);
    if (!s) { ac4_set_error("ac4_server_create: out of host memory"); return AC4_ECAP; }
    double cap_tot = 0.0;
    for (int k = 0; k < nslots; k++) cap_tot += slot_cap[k] > 0 ? slot_cap[k] : 0;
    int ng = groups ? groups
                    : ((nslots >= SRV_GROUPS && cap_tot >= (double)SRV_GROUP_MIN * SRV_GROUPS) ? SRV_GROUPS : 1);
    if (ng > nslots) ng = nslots;
    s->ng = ng;
    s->nslots = nslots;
    s->chunk = chunk;
    /* the ramp does not advance a replica's step counter: whole chunks keep
     * every production step count a multiple of the chunk at a tick */
    s->relax = ((relax_steps + chunk - 1) / chunk) * chunk;
    s->cap = malloc(sizeof(int) * nslots);
    s->slot_job = malloc(sizeof(int) * nslots);
    s->err = calloc((size_t)nslots, sizeof(int));
    s->slot_h = malloc(sizeof(int) * nslots);
    s->slot_l = malloc(sizeof(int) * nslots);
    s->gerr = malloc(sizeof(int) * nslots);
    s->mds = calloc((size_t)ng, sizeof(ac4_md *));
    int *ord = malloc(sizeof(int) * nslots), *caps_g = malloc(sizeof(int) * nslots);
    double *load = calloc((size_t)ng, sizeof(double));
    if (!s->cap || !s->slot_job || !s->err || !s->slot_h || !s->slot_l || !s->gerr || !s->mds || !ord || !caps_g ||
        !load) {
        free(ord); free(caps_g); free(load);
        ac4_server_destroy(s);
        ac4_set_error("ac4_server_create: out of host memory");
        return AC4_ECAP;
    }
    for (int k = 0; k < nslots; k++) { s->cap[k] = slot_cap[k]; s->slot_job[k] = -1; ord[k] = k; }
    /* slots to groups: largest capacity first (equal ones in slot order), each
     * to the group with the least capacity so far (the lowest on a tie) */
    for (int a = 1; a < nslots; a++) {
        const int v = ord[a];
        int b = a - 1;
        while (b >= 0 && slot_cap[ord[b]] < slot_cap[v]) { ord[b + 1] = ord[b]; b--; }
        ord[b + 1] = v;
    }
    for (int a = 0; a < nslots; a++) {
        int g = 0;
        for (int q = 1; q < ng; q++) if (load[q] < load[g]) g = q;
        s->slot_h[ord[a]] = g;
        load[g] += slot_cap[ord[a]];
    }
    int rc = AC4_OK;
    for (int g = 0; g < ng && rc == AC4_OK; g++) {
        int n = 0;
        for (int k = 0; k < nslots; k++)
            if (s->slot_h[k] == g) { s->slot_l[k] = n; caps_g[n++] = slot_cap[k]; }
        rc = ac4_md_create_slots_rule(n, caps_g, precision, cap_tot, &s->mds[g]);
    }
    free(ord); free(caps_g); free(load);
    if (rc != AC4_OK) { ac4_server_destroy(s); return rc; }
    *out = s;
    return AC4_OK;
}

int ac4_server_create(int nslots, const int *slot_cap, int precision, int chunk, int relax_steps,
                      ac4_server **out) {
    return ac4_server_create_groups(nslots, slot_cap, precision, chunk, relax_steps, 0, out);
}

int ac4_server_create_budget(double hbm_bytes, int n_min, int n_max, int precision, int chunk, int relax_steps,
                             int origin_every, int n_origin, int n_lags, double a, ac4_server **out) {
    const int ov = n_origin > 0;
    int caps_cap = 65536;
    int *caps = malloc(sizeof(int) * caps_cap);
    if (!caps) { ac4_set_error("ac4_server_create_budget: out of host memory"); return AC4_ECAP; }
    int ns = ac4_server_plan(hbm_bytes, n_min, n_max, ov ? n_origin : 0, ov ? n_lags : 0, caps, caps_cap);
    if (ns < 0) { free(caps); return ns; }
    int rc = ac4_server_create(ns, caps, precision, chunk, relax_steps, out);
    free(caps);
    if (rc != AC4_OK) return rc;
    if (ov && (rc = ac4_server_enable_overlap(*out, origin_every, n_origin, n_lags, a)) != AC4_OK) {
        ac4_server_destroy(*out);
        *out = NULL;
    }
    return rc;
}

int ac4_server_enable_overlap(ac4_server *s, int origin_every, int n_origin, int n_lags, double a) {
    if (!s || s->njobs > 0 || s->ov_on) {
        ac4_set_error("enable_overlap: call once, before the first job is submitted");
        return AC4_ESTATE;
    }
    for (int g = 0; g < s->ng; g++) {
        int rc = ac4_md_overlap_enable(s->mds[g], s->chunk, origin_every, n_origin, n_lags, a);
        if (rc != AC4_OK) return rc;
    }
    s->ov_on = 1;
    s->ov_nl = n_lags + 1;
    return AC4_OK;
}

int ac4_server_slots(const ac4_server *s, int *caps) {
    if (!s) return AC4_EARG;
    if (caps) memcpy(caps, s->cap, sizeof(int) * s->nslots);
    return s->nslots;
}

int ac4_server_destroy(ac4_server *s) {
    if (!s) return AC4_OK;
    if (s->mds)
        for (int g = 0; g < s->ng; g++)
            if (s->mds[g]) ac4_md_destroy(s->mds[g]);
    for (int j = 0; j < s->njobs; j++) { free(s->jobs[j].result); free(s->jobs[j].ov); free(s->jobs[j].ckpt); }
    free(s->jobs); free(s->cap); free(s->slot_job); free(s->err);
    free(s->slot_h); free(s->slot_l); free(s->gerr); free(s->mds);
    free(s);
    return AC4_OK;
}

int ac4_server_submit(ac4_server *s, const ac4_job *job) {
    if (!s || !job || job->n < 2 || job->steps < 1) { ac4_set_error("submit: bad job"); return AC4_EARG; }
    int fits = 0;
    for (int k = 0; k < s->nslots; k++) fits |= s->cap[k] >= job->n;
    if (!fits) { ac4_set_error("submit: no slot can hold N=%d", job->n); return AC4_ECAP; }
    if (s->njobs == s->jcap) {
        int nc = s->jcap ? 2 * s->jcap : 64;
        job_rec *nj = realloc(s->jobs, sizeof(job_rec) * nc);
        if (!nj) { ac4_set_error("submit: out of host memory"); return AC4_ECAP; }
        s->jobs = nj;
        s->jcap = nc;
    }
    job_rec *r = &s->jobs[s->njobs];
    memset(r, 0, sizeof *r);
    r->job = *job;
    r->slot = -1;
    /* production steps rounded up to whole chunks, plus the relaxation ramp */
    long prod = ((job->steps + s->chunk - 1) / s->chunk) * s->chunk;
    r->target = prod + s->relax;
    return s->njobs++;
}

static int admit(ac4_server *s, int j, int slot, double now) {
    job_rec *r = &s->jobs[j];
    const ac4_job *q = &r->job;
    ac4_md *md = s->mds[s->slot_h[slot]];
    const int ls = s->slot_l[slot];
    int tau = (int)floor(q->tau_p / 0.005 + 0.5); /* tau_p in time units, dt = 0.005 */
    if (tau < 1) tau = 1;
    ac4_replica_params p = {q->n, 1.2, q->T, q->f0, -1, tau, q->seed};
    int rc = ac4_md_assign(md, ls, &p);
    if (rc != AC4_OK) return rc;
    if (r->ckpt) { /* a preempted job continues exactly where it stopped (its Q(t) origins and sums included) */
        rc = ac4_md_restore(md, ls, r->ckpt, r->ckpt_len);
        if (rc != AC4_OK) return rc;
        free(r->ckpt);
        r->ckpt = NULL;
        r->state = 1;
        r->slot = slot;
        s->slot_job[slot] = j;
        return AC4_OK;
    }
    if (s->ov_on && (rc = ac4_md_overlap_reset(md, ls)) != AC4_OK) return rc;
    int n = q->n, nact = (int)((double)n * 0.1000);
    double *buf = malloc(sizeof(double) * 9 * (size_t)n);
    int *ib = malloc(sizeof(int) * (2 * (size_t)n + 3 * (size_t)(nact + 1)));
    if (!buf || !ib) {
        free(buf);
        free(ib);
        ac4_set_error("admit: out of host memory");
        return AC4_ECAP;
    }
    double *x = buf, *y = x + n, *z = y + n, *xu = z + n, *yu = xu + n, *zu = yu + n, *px = zu + n, *py = px + n,
           *pz = py + n;
    int *type = ib, *act = type + n, *kx = act + n, *ky = kx + nact + 1, *kz = ky + nact + 1;
    rc = ac4_init_lattice(n, q->T, nact, q->seed, x, y, z, xu, yu, zu, px, py, pz, type, act, kx, ky, kz);
    if (rc == AC4_OK)
        rc = ac4_md_load_slot(md, ls, x, y, z, xu, yu, zu, px, py, pz, type, act, kx, ky, kz);
    free(buf);
    free(ib);
    if (rc != AC4_OK) return rc;
    /* chain for the Nose-Hoover phase, drawn from the job seed; it starts
     * after the relaxation ramp */
    double c[9];
    ac4_chain_init(n, q->T, q->seed, c);
    if ((rc = ac4_md_set_chain(md, ls, c)) != AC4_OK) return rc;
    if ((rc = ac4_md_relax_slot(md, ls, s->relax, 1.5, 1.0)) != AC4_OK) return rc;
    r->state = 1;
    r->slot = slot;
    r->t_start = now;
    s->slot_job[slot] = j;
    return AC4_OK;
}

/* A finished job: final x_u in reference order and its Q(t) sums. */
static int retire(ac4_server *s, int k, job_rec *r) {
    const int n = r->job.n;
    ac4_md *md = s->mds[s->slot_h[k]];
    const int ls = s->slot_l[k];
    r->result = malloc(sizeof(double) * 3 * (size_t)n);
    if (!r->result) { ac4_set_error("retire: out of host memory"); return AC4_ECAP; }
    int rc = ac4_md_get_state(md, ls, NULL, NULL, NULL, r->result, r->result + n, r->result + 2 * n, NULL, NULL,
                              NULL, NULL, NULL, NULL);
    if (rc != AC4_OK) return rc;
    if (s->ov_on) {
        const int nl = s->ov_nl;
        r->ov = malloc(sizeof(double) * 3 * (size_t)nl);
        int *cnt = malloc(sizeof(int) * (size_t)nl);
        if (!r->ov || !cnt) { free(cnt); ac4_set_error("retire: out of host memory"); return AC4_ECAP; }
        rc = ac4_md_overlap_get(md, ls, r->ov, r->ov + nl, cnt);
        for (int l = 0; l < nl; l++) r->ov[2 * nl + l] = (double)cnt[l];
        free(cnt);
        if (rc != AC4_OK) return rc;
    }
    return AC4_OK;
}

int ac4_server_tick(ac4_server *s, double now) {
    if (!s) return AC4_EARG;
    /* admission: queued jobs in arrival order, each into a free slot of its
     * class (the smallest slot capacity that holds it).  The kernel shape is
     * fixed when the server is created and every kernel sums in a fixed
     * order, so a job's arithmetic depends on its own parameters and the
     * server's configuration only, never on which other jobs are running. */
    /* two passes: jobs that have not run yet, then preempted jobs (a
     * preemption yields the slot to the work that was waiting) */
    for (int pass = 0; pass < 2; pass++)
    for (int j = 0; j < s->njobs; j++) {
        if (s->jobs[j].state != 0 || s->jobs[j].job.t_arrive > now) continue;
        if ((s->jobs[j].ckpt != NULL) != (pass == 1)) continue;
        int cls = 0;
        for (int k = 0; k < s->nslots; k++)
            if (s->cap[k] >= s->jobs[j].job.n && (cls == 0 || s->cap[k] < cls)) cls = s->cap[k];
        int best = -1;
        for (int k = 0; k < s->nslots && best < 0; k++)
            if (s->slot_job[k] < 0 && s->cap[k] == cls) best = k;
        if (best >= 0) {
            int rc = admit(s, j, best, now);
            if (rc != AC4_OK) return rc;
        }
    }
    int running = 0;
    for (int k = 0; k < s->nslots; k++) running += s->slot_job[k] >= 0;
    if (!running) return 0;
    /* the groups that hold a running job, stepped together */
    ac4_md *act[SRV_GROUPS_MAX];
    int gact[SRV_GROUPS_MAX], na = 0;
    for (int g = 0; g < s->ng; g++) {
        int busy = 0;
        for (int k = 0; k < s->nslots; k++) busy |= s->slot_h[k] == g && s->slot_job[k] >= 0;
        if (busy) { gact[na] = g; act[na++] = s->mds[g]; }
    }
    int rc = ac4_md_step_graph_multi(act, na, s->chunk);
    if (rc != AC4_OK) return rc;
    /* every running job's frame of this chunk, counted on the device */
    for (int a = 0; a < na; a++)
        if (s->ov_on && (rc = ac4_md_overlap_sample(act[a])) != AC4_OK) return rc;
    for (int k = 0; k < s->nslots; k++) s->err[k] = 0;
    for (int a = 0; a < na; a++) {
        if ((rc = ac4_md_errors(act[a], s->gerr)) != AC4_OK) return rc;
        for (int k = 0; k < s->nslots; k++)
            if (s->slot_h[k] == gact[a]) s->err[k] = s->gerr[s->slot_l[k]];
    }
    for (int k = 0; k < s->nslots; k++) {
        int j = s->slot_job[k];
        if (j < 0) continue;
        job_rec *r = &s->jobs[j];
        if (s->err[k]) { /* this job alone fails; its slot is cleared for the next one */
            r->state = 3;
            r->err = s->err[k];
            r->t_done = now;
            s->slot_job[k] = -1;
            if ((rc = ac4_md_release(s->mds[s->slot_h[k]], s->slot_l[k])) != AC4_OK) return rc;
            continue;
        }
        r->steps_done += s->chunk;
        r->frames++;
        s->particle_steps += (long long)r->job.n * s->chunk;
        if (r->steps_done >= r->target) {
            if ((rc = retire(s, k, r)) != AC4_OK) return rc;
            r->state = 2;
            r->t_done = now;
            s->slot_job[k] = -1;
            if ((rc = ac4_md_release(s->mds[s->slot_h[k]], s->slot_l[k])) != AC4_OK) return rc;
        }
    }
    return running;
}

int ac4_server_preempt(ac4_server *s, int id) {
    if (!s || id < 0 || id >= s->njobs) { ac4_set_error("preempt: bad id"); return AC4_EARG; }
    job_rec *r = &s->jobs[id];
    if (r->state != 1) { ac4_set_error("preempt: job %d is not running", id); return AC4_ESTATE; }
    ac4_md *md = s->mds[s->slot_h[r->slot]];
    const int ls = s->slot_l[r->slot];
    long long len = ac4_md_checkpoint_size(md, ls);
    if (len < 0) return (int)len;
    r->ckpt = malloc((size_t)len);
    if (!r->ckpt) { ac4_set_error("preempt: cannot allocate %lld bytes", len); return AC4_ECAP; }
    int rc = ac4_md_checkpoint(md, ls, r->ckpt, len);
    if (rc != AC4_OK) { free(r->ckpt); r->ckpt = NULL; return rc; }
    r->ckpt_len = len;
    s->slot_job[r->slot] = -1;
    rc = ac4_md_release(md, ls);
    r->slot = -1;
    r->state = 0; /* queued again; its arrival time keeps its place in the FIFO */
    r->preemptions++;
    return rc;
}

int ac4_server_pending(const ac4_server *s) {
    int p = 0;
    for (int j = 0; j < s->njobs; j++) p += s->jobs[j].state < 2;
    return p;
}

int ac4_server_job(const ac4_server *s, int id, double *t_start, double *t_done, long *steps_done, int *frames,
                   int *state) {
    if (!s || id < 0 || id >= s->njobs) { ac4_set_error("job: bad id"); return AC4_EARG; }
    const job_rec *r = &s->jobs[id];
    if (t_start) *t_start = r->t_start;
    if (t_done) *t_done = r->t_done;
    if (steps_done) *steps_done = r->steps_done;
    if (frames) *frames = r->frames;
    if (state) *state = r->state;
    return AC4_OK;
}

int ac4_server_job_error(const ac4_server *s, int id) {
    if (!s || id < 0 || id >= s->njobs) { ac4_set_error("job_error: bad id"); return AC4_EARG; }
    return s->jobs[id].err;
}

int ac4_server_result(const ac4_server *s, int id, double *xu, double *yu, double *zu) {
    if (!s || id < 0 || id >= s->njobs || s->jobs[id].state != 2) {
        ac4_set_error("result: job %d not finished", id);
        return AC4_ESTATE;
    }
    const job_rec *r = &s->jobs[id];
    int n = r->job.n;
    memcpy(xu, r->result, sizeof(double) * n);
    memcpy(yu, r->result + n, sizeof(double) * n);
    memcpy(zu, r->result + 2 * n, sizeof(double) * n);
    return AC4_OK;
}

int ac4_server_job_overlap(const ac4_server *s, int id, double *sq, double *sq2, int *cnt) {
    if (!s || id < 0 || id >= s->njobs || s->jobs[id].state != 2 || !s->ov_on) {
        ac4_set_error("job_overlap: job %d not finished or the server keeps no overlap", id);
        return AC4_ESTATE;
    }
    const job_rec *r = &s->jobs[id];
    const int nl = s->ov_nl;
    for (int l = 0; l < nl; l++) {
        if (sq) sq[l] = r->ov[l];
        if (sq2) sq2[l] = r->ov[nl + l];
        if (cnt) cnt[l] = (int)r->ov[2 * nl + l];
    }
    return AC4_OK;
}

long long ac4_server_particle_steps(const ac4_server *s) { return s ? s->particle_steps : 0; }

ac4_md *ac4_server_handle(ac4_server *s) { return s ? s->mds[0] : NULL; }

ac4_md *ac4_server_slot_handle(ac4_server *s, int slot, int *local) {
    if (!s || slot < 0 || slot >= s->nslots) { ac4_set_error("slot_handle: bad slot"); return NULL; }
    if (local) *local = s->slot_l[slot];
    return s->mds[s->slot_h[slot]];
}

int ac4_server_groups(const ac4_server *s) { return s ? s->ng : AC4_EARG; }

int ac4_server_job_slot(const ac4_server *s, int id) {
    if (!s || id < 0 || id >= s->njobs) { ac4_set_error("job_slot: bad id"); return AC4_EARG; }
    return s->jobs[id].slot;
}
