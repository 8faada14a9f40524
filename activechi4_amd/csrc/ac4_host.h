/* ac4_host.h -- internal host helpers of libac4 (not part of the public ABI). */
#ifndef AC4_HOST_H
#define AC4_HOST_H
#ifdef __cplusplus
extern "C" {
#endif

void ac4_set_error(const char *fmt, ...);

/* ac4_md_create_slots with the automatic force team width taken for a total
 * capacity of cap_rule (> 0) instead of the handle's own: the launch groups
 * of a server share one kernel shape.  (ac4.h, included first, declares
 * ac4_md.) */
int ac4_md_create_slots_rule(int nslots, const int *cap_n, int precision, double cap_rule, ac4_md **out);

/* The (origin, lag) walk of Analysis/msd_pos.c:44-107 (rule 0) and
 * overlap.c:42-106 (rule 1): pairs in walk order. */
typedef struct {
    int np, max_slot, maximal_interval, n_origins_counted;
    int *cur, *oth, *slot;
} ac4_walk_t;
int ac4_walk(const int *fidx, int nframes, int gap, double factor, int rule, ac4_walk_t *w);
void ac4_walk_free(ac4_walk_t *w);
int ac4_degen(int x, int y, int z);
int ac4_getq(int *vx, int *vy, int *vz, int *q);
int ac4_write_msd(const char *path, int noi, const double *m0, const double *m1, const double *m2,
                  const int *counters, const double *delta_t);
int ac4_write_overlap(const char *path, int noi, const double *c, const double *x4, const int *counters,
                      const double *delta_t);
int ac4_write_phase(const char *path, int is_s4, int N, double totpart, int nslots, const int *vx, const int *vy,
                    const int *vz, const double *cs, const double *ss, const double *cs2, const double *ss2,
                    const int *counters, const double *delta_t, const double *qval);

#ifdef __cplusplus
}
#endif
#endif
