/*
 * ac4_server.c -- request loop for MD jobs on one GPU (continuous batching).
 * No reference counterpart: the reference runs one job per mpirun
 * (MDcode/readme:1-11).  Jobs (N, T, f0, tau_p, steps, seed) arrive over
 * time; the server keeps a fixed set of device-resident slots (capacities
 * chosen by the caller from an HBM budget), admits queued jobs into a free
 * slot of their size class (FIFO order; a later job may take a free slot of
 * its own class while an earlier one waits), advances every slot together in chunks
 * of `chunk` steps with one replayed hipGraph, copies every running job's
 * frame (x_u) to the host after each chunk, and retires finished jobs.
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
    int state; /* 0 queued, 1 running, 2 done */
    int slot, frames, preemptions;
    long steps_done, target;
    double t_start, t_done;
    double *result;  /* final x_u,y_u,z_u (3n) */
    char *ckpt;      /* exact checkpoint of a preempted job, resumed on admission */
    long long ckpt_len;
} job_rec;

struct ac4_server {
    ac4_md *md;
    int nslots, chunk, relax;
    int *cap, *slot_job;
    job_rec *jobs;
    int njobs, jcap;
    double *hx, *hy, *hz; /* export buffers (ntot each) */
    int ntot;
    long long particle_steps;
};

int ac4_server_create(int nslots, const int *slot_cap, int precision, int chunk, int relax_steps,
                      ac4_server **out) {
    if (nslots < 1 || !slot_cap || chunk < 1 || relax_steps < 1 || !out) {
        ac4_set_error("ac4_server_create: bad arguments");
        return AC4_EARG;
    }
    ac4_server *s = calloc(1, sizeof *s);
    int rc = ac4_md_create_slots(nslots, slot_cap, precision, &s->md);
    if (rc != AC4_OK) { free(s); return rc; }
    s->nslots = nslots;
    s->chunk = chunk;
    s->relax = relax_steps;
    s->cap = malloc(sizeof(int) * nslots);
    s->slot_job = malloc(sizeof(int) * nslots);
    for (int k = 0; k < nslots; k++) { s->cap[k] = slot_cap[k]; s->slot_job[k] = -1; }
    s->ntot = ac4_md_total_particles(s->md);
    s->hx = malloc(sizeof(double) * s->ntot);
    s->hy = malloc(sizeof(double) * s->ntot);
    s->hz = malloc(sizeof(double) * s->ntot);
    *out = s;
    return AC4_OK;
}

int ac4_server_destroy(ac4_server *s) {
    if (!s) return AC4_OK;
    ac4_md_destroy(s->md);
    for (int j = 0; j < s->njobs; j++) { free(s->jobs[j].result); free(s->jobs[j].ckpt); }
    free(s->jobs); free(s->cap); free(s->slot_job); free(s->hx); free(s->hy); free(s->hz);
    free(s);
    return AC4_OK;
}

int ac4_server_submit(ac4_server *s, const ac4_job *job) {
    if (!s || !job || job->n < 2 || job->steps < 1) { ac4_set_error("submit: bad job"); return AC4_EARG; }
    int fits = 0;
    for (int k = 0; k < s->nslots; k++) fits |= s->cap[k] >= job->n;
    if (!fits) { ac4_set_error("submit: no slot can hold N=%d", job->n); return AC4_ECAP; }
    if (s->njobs == s->jcap) {
        s->jcap = s->jcap ? 2 * s->jcap : 64;
        s->jobs = realloc(s->jobs, sizeof(job_rec) * s->jcap);
    }
    job_rec *r = &s->jobs[s->njobs];
    memset(r, 0, sizeof *r);
    r->job = *job;
    r->slot = -1;
    /* production steps rounded up to whole chunks, plus the relaxation ramp */
    long prod = ((job->steps + s->chunk - 1) / s->chunk) * s->chunk;
    r->target = prod + ((s->relax + s->chunk - 1) / s->chunk) * s->chunk;
    return s->njobs++;
}

static int admit(ac4_server *s, int j, int slot, double now) {
    job_rec *r = &s->jobs[j];
    const ac4_job *q = &r->job;
    int tau = (int)floor(q->tau_p / 0.005 + 0.5); /* tau_p in time units, dt = 0.005 */
    if (tau < 1) tau = 1;
    ac4_replica_params p = {q->n, 1.2, q->T, q->f0, -1, tau, q->seed};
    int rc = ac4_md_assign(s->md, slot, &p);
    if (rc != AC4_OK) return rc;
    if (r->ckpt) { /* a preempted job continues exactly where it stopped */
        rc = ac4_md_restore(s->md, slot, r->ckpt, r->ckpt_len);
        if (rc != AC4_OK) return rc;
        free(r->ckpt);
        r->ckpt = NULL;
        r->state = 1;
        r->slot = slot;
        s->slot_job[slot] = j;
        return AC4_OK;
    }
    int n = q->n, nact = (int)((double)n * 0.1000);
    double *buf = malloc(sizeof(double) * 9 * (size_t)n);
    int *ib = malloc(sizeof(int) * (2 * (size_t)n + 3 * (size_t)(nact + 1)));
    double *x = buf, *y = x + n, *z = y + n, *xu = z + n, *yu = xu + n, *zu = yu + n, *px = zu + n, *py = px + n,
           *pz = py + n;
    int *type = ib, *act = type + n, *kx = act + n, *ky = kx + nact + 1, *kz = ky + nact + 1;
    rc = ac4_init_lattice(n, q->T, nact, q->seed, x, y, z, xu, yu, zu, px, py, pz, type, act, kx, ky, kz);
    if (rc == AC4_OK)
        rc = ac4_md_load_slot(s->md, slot, x, y, z, xu, yu, zu, px, py, pz, type, act, kx, ky, kz);
    free(buf);
    free(ib);
    if (rc != AC4_OK) return rc;
    /* chain for the Nose-Hoover phase, drawn from the job seed; it starts
     * after the relaxation ramp */
    double c[9];
    ac4_chain_init(n, q->T, q->seed, c);
    if ((rc = ac4_md_set_chain(s->md, slot, c)) != AC4_OK) return rc;
    if ((rc = ac4_md_relax_slot(s->md, slot, s->relax, 1.5, 1.0)) != AC4_OK) return rc;
    r->state = 1;
    r->slot = slot;
    r->t_start = now;
    s->slot_job[slot] = j;
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
    int rc = ac4_md_step_graph(s->md, s->chunk);
    if (rc != AC4_OK) return rc;
    /* frame of every running job after the chunk (chunk = frame interval) */
    if ((rc = ac4_md_export_all(s->md, s->hx, s->hy, s->hz)) != AC4_OK) return rc;
    for (int k = 0; k < s->nslots; k++) {
        int j = s->slot_job[k];
        if (j < 0) continue;
        job_rec *r = &s->jobs[j];
        r->steps_done += s->chunk;
        r->frames++;
        s->particle_steps += (long long)r->job.n * s->chunk;
        if (r->steps_done >= r->target) {
            int off = ac4_md_slot_offset(s->md, k), n = r->job.n;
            r->result = malloc(sizeof(double) * 3 * (size_t)n);
            memcpy(r->result, s->hx + off, sizeof(double) * n);
            memcpy(r->result + n, s->hy + off, sizeof(double) * n);
            memcpy(r->result + 2 * n, s->hz + off, sizeof(double) * n);
            r->state = 2;
            r->t_done = now;
            s->slot_job[k] = -1;
            if ((rc = ac4_md_release(s->md, k)) != AC4_OK) return rc;
        }
    }
    return running;
}

int ac4_server_preempt(ac4_server *s, int id) {
    if (!s || id < 0 || id >= s->njobs) { ac4_set_error("preempt: bad id"); return AC4_EARG; }
    job_rec *r = &s->jobs[id];
    if (r->state != 1) { ac4_set_error("preempt: job %d is not running", id); return AC4_ESTATE; }
    long long len = ac4_md_checkpoint_size(s->md, r->slot);
    if (len < 0) return (int)len;
    r->ckpt = malloc((size_t)len);
    if (!r->ckpt) { ac4_set_error("preempt: cannot allocate %lld bytes", len); return AC4_ECAP; }
    int rc = ac4_md_checkpoint(s->md, r->slot, r->ckpt, len);
    if (rc != AC4_OK) { free(r->ckpt); r->ckpt = NULL; return rc; }
    r->ckpt_len = len;
    s->slot_job[r->slot] = -1;
    rc = ac4_md_release(s->md, r->slot);
    r->slot = -1;
    r->state = 0; /* queued again; its arrival time keeps its place in the FIFO */
    r->preemptions++;
    return rc;
}

int ac4_server_pending(const ac4_server *s) {
    int p = 0;
    for (int j = 0; j < s->njobs; j++) p += s->jobs[j].state != 2;
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

long long ac4_server_particle_steps(const ac4_server *s) { return s ? s->particle_steps : 0; }
