/*
 * run.c -- drop-in for the reference MD executable (MDcode/main.c:391-492):
 *   run <jobno> <restart_eq 0|1> <restart_run 0|1> [N]
 * reads ./para_file (MDcode/main.c:410-426), writes "%.3lf"(T)/data_00<jobno>/
 * with data_coor.out, time_index_file.dat, stuff.dat, restart_file.dat,
 * ActiveFileIndex.dat, ActiveDirection*.dat and ini.dat in the reference's
 * formats.  All particle work runs in libac4.so on the GPU.
 *
 * N is compile-time in the reference (MDcode/global.h:15, 50000); here it is
 * the optional 4th argument (default 50000).  Environment:
 *   AC4_SEED           RNG key (default derived from jobno)
 *   AC4_FP32=1         FP32 pair forces
 *   AC4_BIN=1          also write data_coor.bin: per frame the step, the
 *                      values the text parses to (x_u, y_u, z_u) and types
 *   AC4_RESUME=1       with restart_run=1: continue the interrupted run
 *                      exactly from restart_file.ac4 (the reference's
 *                      restart_run starts the production again from step 0
 *                      with a new chain, MDcode/main.c:219-236)
 *   AC4_RESTART_EVERY  production steps between restart files (100000)
 *   AC4_STOP_AT        end the run after the restart written at this step
 *                      (for time-limited batch jobs; resume with AC4_RESUME)
 */
#include <errno.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>
#include "ac4.h"
#include "xalloc.h"

#define TIME_STEP 0.005 /* MDcode/global.h:33 */
#define TAU_PER 200     /* MDcode/global.h:58 */

static void die(const char *what, int rc) {
    fprintf(stderr, "run: %s failed (%d): %s\n", what, rc, ac4_last_error());
    exit(1);
}
#define CK(e) do { int rc_ = (e); if (rc_ != AC4_OK) die(#e, rc_); } while (0)

typedef struct {
    int n, nact;
    double *x, *y, *z, *xu, *yu, *zu, *px, *py, *pz;
    int *type, *act, *kx, *ky, *kz;
} state_t;

static void mkdir_p(const char *p) {
    char buf[512];
    snprintf(buf, sizeof buf, "%s", p);
    for (char *s = buf + 1; *s; s++)
        if (*s == '/') { *s = 0; mkdir(buf, 0755); *s = '/'; }
    mkdir(buf, 0755);
}

static void rm_rf(const char *p) {
    char cmd[600];
    snprintf(cmd, sizeof cmd, "rm -rf '%s'", p); /* MDcode/main.c:458-460 */
    if (system(cmd) != 0) fprintf(stderr, "run: could not remove %s\n", p);
}

static double gauss(uint64_t *s) { /* xorshift64* + Box-Muller for p_eta draws */
    double u[2];
    for (int k = 0; k < 2; k++) {
        *s ^= *s >> 12; *s ^= *s << 25; *s ^= *s >> 27;
        u[k] = ((*s * 2685821657736338717ull) >> 11) * (1.0 / 9007199254740992.0);
    }
    return sqrt(-2.0 * log(1.0 - u[0])) * cos(2.0 * 3.14159265358979 * u[1]);
}

/* InitialiseNoseHoover, MDcode/time_evolve.c:476-493: eta = 0, p_eta drawn
 * with variance T*Q_k.  The masses Q_k are the ones the library gives every
 * new replica (Q0 = DOF*T*tau^2, Q1 = Q2 = T*tau^2, tau = 0.2); they do not
 * change during a run. */
static void init_chain(ac4_md *md, double T, uint64_t *s) {
    double c[9];
    CK(ac4_md_get_chain(md, 0, c));
    for (int i = 0; i < 3; i++) c[i] = 0.0;
    for (int i = 0; i < 3; i++) c[3 + i] = sqrt(T * c[6 + i]) * gauss(s);
    CK(ac4_md_set_chain(md, 0, c));
}

/* fixDrift, MDcode/initialise.c:104-138 */
static void fix_drift(state_t *S) {
    double a = 0, b = 0, c = 0;
    for (int i = 0; i < S->n; i++) { a += S->px[i]; b += S->py[i]; c += S->pz[i]; }
    a /= S->n; b /= S->n; c /= S->n;
    for (int i = 0; i < S->n; i++) { S->px[i] -= a; S->py[i] -= b; S->pz[i] -= c; }
}

static void pull(ac4_md *md, state_t *S) {
    CK(ac4_md_get_state(md, 0, S->x, S->y, S->z, S->xu, S->yu, S->zu, S->px, S->py, S->pz, S->kx, S->ky, S->kz));
}
static void push(ac4_md *md, state_t *S) {
    CK(ac4_md_set_state(md, 0, S->x, S->y, S->z, S->xu, S->yu, S->zu, S->px, S->py, S->pz, S->type, S->act, S->kx,
                        S->ky, S->kz));
}

static void print_stuff(ac4_md *md, const char *file, int n) { /* MDcode/print.c:60-98 */
    double o[8];
    CK(ac4_md_stats(md, 0, o));
    FILE *f = fopen(file, "ab");
    if (!f) return;
    fprintf(f, "%g\t%g\t%g\t%g\t%.15lf\t%g\n", o[0] / (double)n, o[2] / (double)n, o[3], o[4], o[5], o[6]);
    fclose(f);
}

/* frame text buffer and the sidecar's parsed copy */
static char *g_txt;
static double *g_bx, *g_by, *g_bz;
static int *g_bt;

static void print_data(ac4_md *md, state_t *S, const char *file, const char *bin_file, long step) {
    /* MDcode/print.c:4-16, one formatted block and one write per frame; the
     * lines are formatted on the device, the host only joins them */
    const long long cap = 80LL * S->n;
    if (!g_txt) g_txt = xmalloc((size_t)cap);
    const long long k = ac4_md_format_frame(md, 0, g_txt, cap);
    if (k < 0) die("ac4_md_format_frame", (int)k);
    FILE *f = fopen(file, "ab");
    if (!f) return;
    fwrite(g_txt, 1, (size_t)k, f);
    fclose(f);
    if (!bin_file) return;
    if (!g_bx) {
        g_bx = xmalloc(sizeof(double) * S->n); g_by = xmalloc(sizeof(double) * S->n);
        g_bz = xmalloc(sizeof(double) * S->n); g_bt = xmalloc(sizeof(int) * S->n);
    }
    if (ac4_parse_frame(g_txt, k, S->n, g_bx, g_by, g_bz, g_bt) != k) die("ac4_parse_frame", AC4_EIO);
    f = fopen(bin_file, "ab");
    if (!f) return;
    if (ftell(f) == 0) { /* header: magic, N */
        const char magic[8] = {'A', 'C', '4', 'F', 'R', 'M', '0', '1'};
        const int32_t hd[2] = {S->n, 0};
        fwrite(magic, 1, 8, f);
        fwrite(hd, sizeof hd, 1, f);
    }
    const uint64_t st = (uint64_t)step;
    fwrite(&st, sizeof st, 1, f);
    fwrite(g_bx, sizeof(double), (size_t)S->n, f);
    fwrite(g_by, sizeof(double), (size_t)S->n, f);
    fwrite(g_bz, sizeof(double), (size_t)S->n, f);
    fwrite(g_bt, sizeof(int), (size_t)S->n, f);
    fclose(f);
}

/* restart_file.ac4: the loop position plus an exact device checkpoint */
static void write_exact(ac4_md *md, const char *file, long next, int index) {
    long long len = ac4_md_checkpoint_size(md, 0);
    if (len < 0) die("ac4_md_checkpoint_size", (int)len);
    char *buf = xmalloc((size_t)len);
    CK(ac4_md_checkpoint(md, 0, buf, len));
    FILE *f = fopen(file, "wb");
    if (!f) { free(buf); return; }
    const char magic[8] = {'A', 'C', '4', 'R', 'U', 'N', '0', '1'};
    const int64_t hd[3] = {next, index, len};
    fwrite(magic, 1, 8, f);
    fwrite(hd, sizeof hd, 1, f);
    fwrite(buf, 1, (size_t)len, f);
    fclose(f);
    free(buf);
}

static int read_exact(ac4_md *md, const char *file, long *next, int *index) {
    FILE *f = fopen(file, "rb");
    if (!f) return -1;
    char magic[8];
    int64_t hd[3];
    if (fread(magic, 1, 8, f) != 8 || memcmp(magic, "AC4RUN01", 8) || fread(hd, sizeof hd, 1, f) != 1) {
        fclose(f);
        return -1;
    }
    if (hd[2] != ac4_md_checkpoint_size(md, 0)) {  /* written for another N or build */
        fprintf(stderr, "%s: checkpoint of %lld bytes, this run's is %lld\n", file, (long long)hd[2],
                ac4_md_checkpoint_size(md, 0));
        fclose(f);
        return -1;
    }
    char *buf = xmalloc((size_t)hd[2]);
    const int ok = fread(buf, 1, (size_t)hd[2], f) == (size_t)hd[2];
    fclose(f);
    if (!ok) { free(buf); return -1; }
    CK(ac4_md_restore(md, 0, buf, hd[2]));
    free(buf);
    *next = (long)hd[0];
    *index = (int)hd[1];
    return 0;
}

static void write_restart(ac4_md *md, state_t *S, const char *file, double L, int iter, int last) {
    pull(md, S); /* MDcode/main.c:272-288 */
    FILE *f = fopen(file, "wb");
    if (!f) return;
    fprintf(f, "%.14lf\t%d\t%d\n", L, iter, last);
    for (int i = 0; i < S->n; i++)
        fprintf(f, "%.14lf\t%.14lf\t%.14lf\t%.14lf\t%.14lf\t%.14lf\t%d\n", S->xu[i], S->yu[i], S->zu[i], S->px[i],
                S->py[i], S->pz[i], S->type[i]);
    fclose(f);
}

static void save_dirs(state_t *S, const char *file) { /* MDcode/active.c:87-99 */
    FILE *f = fopen(file, "wb");
    if (!f) return;
    for (int i = 0; i < S->nact; i++) fprintf(f, "%d\t%d\t%d\n", S->kx[i], S->ky[i], S->kz[i]);
    fclose(f);
}

static int read_restart(state_t *S, const char *file, double *L) { /* MDcode/main.c:290-305 */
    FILE *f = fopen(file, "rb");
    if (!f) return -1;
    int it, idx;
    if (fscanf(f, "%lf\t%d\t%d", L, &it, &idx) != 3) { fclose(f); return -1; }
    for (int i = 0; i < S->n; i++)
        if (fscanf(f, "%lf\t%lf\t%lf\t%lf\t%lf\t%lf\t%d\n", &S->xu[i], &S->yu[i], &S->zu[i], &S->px[i], &S->py[i],
                   &S->pz[i], &S->type[i]) != 7) { fclose(f); return -1; }
    fclose(f);
    for (int i = 0; i < S->n; i++) { /* initialise_restart_state: x = wrap(x_u), x_u = x (:307-326) */
        double *c[3] = {&S->x[i], &S->y[i], &S->z[i]}, u[3] = {S->xu[i], S->yu[i], S->zu[i]};
        for (int k = 0; k < 3; k++) {
            double v = u[k];
            while (v >= 1.0) v -= 1.0;
            while (v < 0.0) v += 1.0;
            *c[k] = v;
        }
        S->xu[i] = S->x[i]; S->yu[i] = S->y[i]; S->zu[i] = S->z[i];
    }
    return 0;
}

static int read_ints(const char *file, int *a, int n, int per_line) {
    FILE *f = fopen(file, "rb");
    if (!f) return -1;
    for (int i = 0; i < n * per_line; i++)
        if (fscanf(f, "%d", &a[i]) != 1) { fclose(f); return -1; }
    fclose(f);
    return 0;
}

int main(int argc, char **argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: %s <jobno> <restart_eq 0|1> <restart_run 0|1> [N]\n", argv[0]);
        return 2;
    }
    int jobno = atoi(argv[1]), restart_eq = atoi(argv[2]), restart_run = atoi(argv[3]);
    int n = argc > 4 ? atoi(argv[4]) : 50000;
    double eqDuration, prodDuration, density, T, P, f0;
    int frame_step;
    char line[1028];
    FILE *pf = fopen("para_file", "r"); /* MDcode/main.c:410-426 */
    if (!pf) { fprintf(stderr, "run: para_file not found\n"); return 1; }
    if (!fgets(line, sizeof line, pf) || !fgets(line, sizeof line, pf) ||
        sscanf(line, "%lf %lf %d", &eqDuration, &prodDuration, &frame_step) != 3 || !fgets(line, sizeof line, pf) ||
        !fgets(line, sizeof line, pf) || sscanf(line, "%lf %lf %lf", &density, &T, &P) != 3 ||
        !fgets(line, sizeof line, pf) || !fgets(line, sizeof line, pf) || sscanf(line, "%lf", &f0) != 1) {
        fprintf(stderr, "run: malformed para_file\n");
        return 1;
    }
    fclose(pf);
    if (frame_step <= 0) frame_step = 1;
    char main_dir[128], sub_dir[256], data_file[320], stuff_file[320], restart_file[320], path[320];
    snprintf(main_dir, sizeof main_dir, "%.3lf", T);
    snprintf(sub_dir, sizeof sub_dir, "%s/data_00%d", main_dir, jobno);
    snprintf(data_file, sizeof data_file, "%s/data_coor.out", sub_dir);
    snprintf(stuff_file, sizeof stuff_file, "%s/stuff.dat", sub_dir);
    snprintf(restart_file, sizeof restart_file, "%s/restart_file.dat", sub_dir);
    char exact_file[320], bin_path[320];
    snprintf(exact_file, sizeof exact_file, "%s/restart_file.ac4", sub_dir);
    snprintf(bin_path, sizeof bin_path, "%s/data_coor.bin", sub_dir);
    const char *bin_file = getenv("AC4_BIN") && atoi(getenv("AC4_BIN")) ? bin_path : NULL;
    const int resume = restart_run && getenv("AC4_RESUME") && atoi(getenv("AC4_RESUME"));
    const long every = getenv("AC4_RESTART_EVERY") ? atol(getenv("AC4_RESTART_EVERY")) : 100000L;
    const long stop_at = getenv("AC4_STOP_AT") ? atol(getenv("AC4_STOP_AT")) : -1L;
    if (every < 1) { fprintf(stderr, "run: AC4_RESTART_EVERY must be positive\n"); return 2; }
    const char *es = getenv("AC4_SEED");
    uint64_t seed = es ? strtoull(es, NULL, 0) : (uint64_t)(0x5eed0000ull + (uint64_t)jobno);
    int prec = getenv("AC4_FP32") && atoi(getenv("AC4_FP32")) ? AC4_FP32 : AC4_FP64;

    state_t S;
    memset(&S, 0, sizeof S);
    S.n = n;
    S.nact = (int)((double)n * 0.1000);
    double **d[] = {&S.x, &S.y, &S.z, &S.xu, &S.yu, &S.zu, &S.px, &S.py, &S.pz};
    for (int k = 0; k < 9; k++) *d[k] = xmalloc(sizeof(double) * n);
    S.type = xmalloc(sizeof(int) * n);
    S.act = xmalloc(sizeof(int) * n);
    S.kx = xmalloc(sizeof(int) * (S.nact + 1));
    S.ky = xmalloc(sizeof(int) * (S.nact + 1));
    S.kz = xmalloc(sizeof(int) * (S.nact + 1));

    double L = pow((double)n / density, 0.33333333); /* MDcode/main.c:465 */
    uint64_t gs = seed ^ 0x9e3779b97f4a7c15ull;
    if (!restart_eq && !restart_run) {
        mkdir_p(main_dir);
        rm_rf(sub_dir);
        mkdir_p(sub_dir);
        CK(ac4_init_lattice(n, T, S.nact, seed, S.x, S.y, S.z, S.xu, S.yu, S.zu, S.px, S.py, S.pz, S.type, S.act, S.kx,
                            S.ky, S.kz));
        snprintf(path, sizeof path, "%s/ActiveFileIndex.dat", sub_dir); /* active.c:9-28 */
        FILE *f = fopen(path, "wb");
        for (int a = 0; a < S.nact; a++)
            for (int i = 0; i < n; i++)
                if (S.act[i] == a) { fprintf(f, "%d\n", i); break; }
        fclose(f);
        snprintf(path, sizeof path, "%s/ini.dat", sub_dir); /* initialise.c:94-97, print.c:18-26 */
        f = fopen(path, "wb");
        for (int i = 0; i < n; i++) fprintf(f, "%.14lf\t%.14lf\t%.14lf\t%d\n", S.x[i], S.y[i], S.z[i], S.type[i]);
        fclose(f);
    } else {
        if (read_restart(&S, restart_file, &L)) { fprintf(stderr, "run: cannot read %s\n", restart_file); return 1; }
        int *idx = xmalloc(sizeof(int) * S.nact);
        snprintf(path, sizeof path, "%s/ActiveFileIndex.dat", sub_dir);
        if (read_ints(path, idx, S.nact, 1)) { fprintf(stderr, "run: cannot read %s\n", path); return 1; }
        for (int i = 0; i < n; i++) S.act[i] = -1;
        for (int a = 0; a < S.nact; a++) S.act[idx[a]] = a;
        free(idx);
        int *k3 = xmalloc(sizeof(int) * 3 * S.nact);
        snprintf(path, sizeof path, "%s/%s", sub_dir, restart_run ? "ActiveDirectionRun.dat" : "ActiveDirection.dat");
        if (read_ints(path, k3, S.nact, 3)) { fprintf(stderr, "run: cannot read %s\n", path); return 1; }
        for (int a = 0; a < S.nact; a++) { S.kx[a] = k3[3 * a]; S.ky[a] = k3[3 * a + 1]; S.kz[a] = k3[3 * a + 2]; }
        free(k3);
    }
    ac4_replica_params p = {n, density, T, f0, S.nact, TAU_PER, seed};
    ac4_md *md = NULL;
    CK(ac4_md_create(1, &p, prec, &md));
    printf("\n\n -----------------------------------------------------------\n\n");
    printf("STARTING MOLECULAR DYNAMICS WITH %d SPHERES \n\n", n);

    if (!restart_eq && !restart_run) {
        /* coldstart(0.1,T) + breakxal(0.1,T) (MDcode/initialise.c:166-201):
         * momenta rescaled every step along 0.5T -> T, then 1.5T -> T, at
         * 0.1 temperature units per time unit */
        int ramp = (int)(0.5 * T / (0.1 * TIME_STEP));
        if (ramp < 1) ramp = 1;
        push(md, &S);
        CK(ac4_md_relax(md, ramp, 0.5, 1.0));
        CK(ac4_md_relax(md, ramp, 1.5, 1.0));
        pull(md, &S);
    }
    if (!restart_run) { /* equilibrate, MDcode/main.c:123-206 */
        fix_drift(&S);
        push(md, &S);
        init_chain(md, T, &gs);
        long steps = (long)(eqDuration / TIME_STEP);
        snprintf(path, sizeof path, "%s/ActiveDirection.dat", sub_dir);
        { /* the reference's 1000 timing steps before equilibration, MDcode/main.c:171-177 */
            time_t t0 = time(0);
            CK(ac4_md_step(md, 1000));
            CK(ac4_md_sync(md));
            printf("Time for 10000 Steps= %ld\n", (long)(time(0) - t0));
        }
        print_stuff(md, stuff_file, n);
        for (long i = 0; i < steps;) {
            long next = steps;
            long a = (i / frame_step + 1) * (long)frame_step, b = (i / 20000 + 1) * 20000L;
            if (a < next) next = a;
            if (b < next) next = b;
            CK(ac4_md_step(md, (int)(next - i)));
            i = next;
            if (i % frame_step == 0) print_stuff(md, stuff_file, n);
            if (i % 20000 == 0) { write_restart(md, &S, restart_file, L, (int)i, 0); save_dirs(&S, path); }
            if (i % 100000 == 0) CK(ac4_md_fix_drift(md, 0, NULL)); /* MDcode/main.c:188 */
        }
        pull(md, &S);
        for (int i = 0; i < n; i++) { S.xu[i] = S.x[i]; S.yu[i] = S.y[i]; S.zu[i] = S.z[i]; } /* :197-203 */
        printf("\nsystem is now equilibriated\n");
    }
    /* frame schedule, MDcode/main.c:479-485 */
    unsigned long long runLength = (unsigned long long)(prodDuration / TIME_STEP);
    int cap = 200 * (12 + (int)(log((double)runLength / 2.0 + 2.0) / log(1.4))) + 64;
    uint64_t *next_idx = xmalloc(sizeof(uint64_t) * cap);
    int nfr = ac4_schedule(runLength, 200, 1.4, next_idx, cap);
    if (nfr < 0) die("ac4_schedule", nfr);
    snprintf(path, sizeof path, "%s/time_index_file.dat", sub_dir);
    FILE *tf = fopen(path, "wb");
    fprintf(tf, "total timesteps equals %d\n", nfr);
    for (int i = 0; i < nfr; i++) fprintf(tf, "%llu\n", (unsigned long long)next_idx[i]);
    fclose(tf);
    /* run, MDcode/main.c:208-270 */
    long i0 = 0;
    int index = 0;
    if (resume) {
        if (read_exact(md, exact_file, &i0, &index)) {
            fprintf(stderr, "run: cannot resume from %s\n", exact_file);
            return 1;
        }
        printf("Run resumed exactly at step %ld\n", i0);
    } else {
        fix_drift(&S);
        push(md, &S);
        init_chain(md, T, &gs);
    }
    snprintf(path, sizeof path, "%s/ActiveDirectionRun.dat", sub_dir);
    long steps = (long)(prodDuration / TIME_STEP);
    printf("Starting run and generating hell lot of data\n");
    for (long i = i0; i < steps;) {
        /* the step index i is the one about to finish; events fire after it */
        long ev = steps - 1;
        if (index < cap && (long)next_idx[index] >= i && (long)next_idx[index] < ev) ev = (long)next_idx[index];
        long a = ((i + frame_step - 1) / frame_step) * (long)frame_step;
        long b = ((i + every - 1) / every) * every;
        long c = ((i + 100000 - 1) / 100000) * 100000L;
        if (a < ev) ev = a;
        if (b < ev) ev = b;
        if (c < ev) ev = c;
        CK(ac4_md_step(md, (int)(ev - i + 1)));
        if (index < cap && (unsigned long long)ev == next_idx[index]) {
            print_data(md, &S, data_file, bin_file, ev);
            index++;
        }
        if (ev % frame_step == 0) print_stuff(md, stuff_file, n);
        i = ev + 1;
        if (ev % 100000 == 0) CK(ac4_md_fix_drift(md, 0, NULL)); /* MDcode/main.c:261, before the restart */
        if (ev % every == 0) {
            write_restart(md, &S, restart_file, L, (int)ev, index - 1);
            save_dirs(&S, path);
            write_exact(md, exact_file, i, index);
            if (stop_at >= 0 && ev >= stop_at) {
                CK(ac4_md_sync(md));
                printf("Stopped after step %ld (resume with AC4_RESUME=1)\n", ev);
                ac4_md_destroy(md);
                return 0;
            }
        }
    }
    CK(ac4_md_sync(md));
    printf("Here it ends\n");
    ac4_md_destroy(md);
    return 0;
}
