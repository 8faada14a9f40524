/*
 * oracle.h -- CPU restatement of the activeChi4 hot path (TEST INFRASTRUCTURE).
 *
 * This library is the checker, never the product.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  Every
 * function restates one reference routine in plain single-threaded C99 FP64
 * with the reference's own expression order (build with -ffp-contract=off);
 * the citation on each declaration is the reference file:line it follows
 * (paths relative to the reference checkout).
 *
 * Pinning (what checks each function independently of this restatement):
 *   or_md_box_length, or_cells_setup ........ closed forms (tests/test_oracle_md.py)
 *   or_cell_of, or_maps ..................... brute force on tiny grids
 *   or_half_list ............................ pair set == O(N^2) brute force;
 *                                             full counts == brute-force counts
 *   or_forces_list .......................... O(N^2) brute force, Newton-3,
 *                                             V(rc)=F(rc)=0, finite-difference
 *                                             of the energy
 *   or_nhc_steps ............................ Heff conservation (f0=0), chain
 *                                             algebra, NVE limit
 *   or_shuffle .............................. invariants (multiset kept); the
 *                                             swap rule checked on hand cases
 *   or_schedule ............................. brute-force sorted-unique
 *                                             construction + hand cases
 *   or_origin_gap, or_walk .................. reference Analysis run on
 *                                             golden trajectories
 *   or_msd, or_overlap, or_s4, or_fsqt,
 *   or_getq, or_write_* ..................... byte-identical output files vs
 *                                             the reference Analysis program
 *                                             (golden fixtures) + closed forms
 *   or_chi4_ensemble ........................ closed forms from per-origin
 *                                             samples (numpy population
 *                                             variance x Nbar, the
 *                                             reference's one-run X4 for one
 *                                             replica, identical replicas,
 *                                             zero for constant samples)
 */
#ifndef AC4_ORACLE_H
#define AC4_ORACLE_H
#include <stdint.h>

/* ------------------------------------------------------------------ MD */
typedef struct {
    int ncx, ncy, ncz, ncell;
    double inv_cx, inv_cy, inv_cz;   /* InvCellSize = L/CellSize (box units) */
} or_cells;

/* MDcode/main.c:465 */
double or_md_box_length(int n, double density);
/* Analysis/main.c:15 */
double or_ana_box_length(int n);
/* MDcode/Parallel.c:72-80 (one process: npx=npy=npz=1) */
void or_cells_setup(double L, or_cells *c);
/* MDcode/cell_list+forces.c:73-74 */
int or_cell_of(const or_cells *c, double x, double y, double z);
/* MDcode/cell_list+forces.c:24-54; map is [ncell][13] */
void or_maps(const or_cells *c, int *map);
/* MDcode/cell_list+forces.c:58-146.  Builds the half Verlet list in the
 * reference's traversal order.  list is [n][maxneib].  Returns 0, or -1 when
 * a particle overflows maxneib or a cell overflows 600. cell_out may be NULL. */
int or_half_list(const or_cells *c, int n, const double *x, const double *y,
                 const double *z, double L, int maxneib, int *cnt, int *list,
                 int *cell_out);
/* For every i: #{j != i : ((dx*dx+dy*dy)+dz*dz)*L*L < 2.85^2}, minimum image
 * as MDcode/cell_list+forces.c:89-98 (brute force O(N^2)). */
void or_full_counts_brute(int n, const double *x, const double *y,
                          const double *z, double L, int *cnt);
/* MDcode/cell_list+forces.c:153-227 (single process: factor = 1).
 * out3 = {pe, virial/3, vir_act}. */
void or_forces_list(int n, const double *x, const double *y, const double *z,
                    const double *xu, const double *yu, const double *zu,
                    const int *type, double L, const int *cnt, const int *list,
                    int maxneib, const int *act_tag, const int *kx,
                    const int *ky, const int *kz, double f0, double *fx,
                    double *fy, double *fz, double *out3);
/* Same physics over all pairs i<j in index order (brute-force check). */
void or_forces_brute(int n, const double *x, const double *y, const double *z,
                     const int *type, double L, const int *act_tag,
                     const int *kx, const int *ky, const int *kz, double f0,
                     double *fx, double *fy, double *fz, double *out2);
/* Forces on m chosen particles by an O(N) pass each (f3: [m][3]). */
void or_forces_sample(int n, const double *x, const double *y, const double *z,
                      const int *type, double L, const int *act_tag,
                      const int *kx, const int *ky, const int *kz, double f0,
                      int m, const int *idx, double *f3);

/* MDcode/time_evolve.c:476-482: q[0]=DOF*T*tau^2, q[1..2]=T*tau^2 */
void or_nhc_masses(int n, double T, double *q3);

/* One process, nsteps x AdvanceTimeNVTNoseHoover (MDcode/time_evolve.c:495-609)
 * with the half list rebuilt under the reference criterion
 * 3*maxmaxD > skin^2 (:508), or the safe one 4*maxmaxD > skin^2 when
 * safe != 0.  If fresh != 0 the list is built and forces computed first
 * (MDcode/main.c:331-332).  The state arrays are updated in place.
 * chain = {eta[3], p_eta[3], q[3]}.  xold = x_u at the last list build.
 * nve != 0 replaces the thermostat by plain velocity Verlet with the same
 * U2-U1-F-U2 order (chain untouched).
 * out8 = {pe, virial, ke, inst_T, Heff, rebuilds, maxmaxD, drift}. */
int or_nhc_steps(int n, double L, double T, double f0, int nsteps, int fresh,
                 int safe, int nve, double *x, double *y, double *z,
                 double *xu, double *yu, double *zu, double *px, double *py,
                 double *pz, double *fx, double *fy, double *fz,
                 const int *type, const int *act_tag, const int *kx,
                 const int *ky, const int *kz, double *chain, double *xold,
                 double *yold, double *zold, double *out8);

/* MDcode/active.c:58-75 for one component: for i: j=(int)(nact*u[i]); swap */
void or_shuffle(int nact, int *k, const double *u);

/* MDcode/main.c:62-119 incl. shellSort (:17-41) on a[1..n] where a[n] is the
 * word just past the filled range (`tail`, zero for a fresh heap page).
 * out must hold cap entries; returns the frame count `current` and leaves the
 * entries the run loop may still read after the last listed one in
 * out[current], out[current+1] (MDcode/main.c:252-257). */
int or_schedule(unsigned long long run_length, int n_origins, double factor,
                unsigned long long tail, unsigned long long *out, int cap);

/* ------------------------------------------------------------ Analysis */
/* Analysis/read_all_data.c:36-48 */
int or_origin_gap(const int *fidx, int nframes);
/* Analysis/S4Qt.c:297-356; arrays of >= 162; returns 162 */
int or_getq(int *vx, int *vy, int *vz, int *qsq);

/* The (origin, lag) walk shared by Analysis/msd_pos.c:44-107,
 * overlap.c:42-106, Fsqt.c:84-256, S4Qt.c:82-250.
 * rule = 0: origins counted on frames k*gap (msd_pos.c:17-29);
 * rule = 1: on frames k*gap+1 (overlap.c:16-29).
 * Returns the number of pairs; pair_cur/pair_oth are frame positions,
 * pair_slot the 1-based time_index.  *maxint receives maximal_interval. */
int or_walk(const int *fidx, int nframes, int gap, double factor, int rule,
            int cap, int *pair_cur, int *pair_oth, int *pair_slot,
            int *maxint);

/* Analysis/msd_pos.c:5-125.  X/Y/Z/TY are [nframes][N].  slot arrays hold
 * nslots entries (nslots >= the reference's num_of_intervals).
 * pair_vals (NULL ok) receives 3 sums per pair in walk order. */
int or_msd(int N, double L, int nframes, const int *fidx, int gap,
           const double *X, const double *Y, const double *Z, const int *TY,
           double factor, int flag, int nslots, double *m0, double *m1,
           double *m2, int *counters, double *delta_t, double *pair_vals,
           int *num_intervals);
/* Analysis/overlap.c:4-128.  pair_count (NULL ok) = overlap count per pair. */
int or_overlap(int N, double L, int nframes, const int *fidx, int gap,
               const double *X, const double *Y, const double *Z,
               const int *TY, double factor, double a, int flag, int nslots,
               double *c, double *x4, int *counters, double *delta_t,
               int *pair_count, int *num_intervals);
/* Analysis/S4Qt.c:4-296.  Slot arrays are [nslots][162] (nslots >= nframes).
 * pair_cs (NULL ok) = {cos[162], sin[162]} per pair; pair_mask (NULL ok) =
 * masked-particle count per pair.  qval[162] receives Q[n]. */
int or_s4(int N, double L, int nframes, const int *fidx, int gap,
          const double *X, const double *Y, const double *Z, double factor,
          double a, int nslots, double *cs, double *ss, double *cs2,
          double *ss2, int *counters, double *delta_t, double *qval,
          double *pair_cs, int *pair_mask);
/* Analysis/Fsqt.c:4-303 (same layout as or_s4). */
int or_fsqt(int N, double L, int nframes, const int *fidx, int gap,
            const double *X, const double *Y, const double *Z, const int *TY,
            double factor, double a, int flag, int nslots, double *cs,
            double *ss, double *cs2, double *ss2, int *counters,
            double *delta_t, double *qval, double *pair_cs);

/* Output writers with the reference's exact formats. */
int or_write_msd(const char *path, int num_intervals, const double *m0,
                 const double *m1, const double *m2, const int *counters,
                 const double *delta_t);                 /* msd_pos.c:112-120 */
int or_write_overlap(const char *path, int num_intervals, const double *c,
                     const double *x4, const int *counters,
                     const double *delta_t);             /* overlap.c:108-121 */
int or_write_s4(const char *path, int N, int nslots, const double *cs,
                const double *ss, const double *cs2, const double *ss2,
                const int *counters, const double *delta_t,
                const double *qval);                     /* S4Qt.c:251-271 */
int or_write_fsqt(const char *path, double totpart, int nslots,
                  const double *cs, const double *ss, const double *cs2,
                  const double *ss2, const int *counters,
                  const double *delta_t, const double *qval); /* Fsqt.c:257-278 */

/* Whole Analysis program on one job directory (Analysis/main.c:3-80 with one
 * rank): reads time_index_file.dat, ActiveFileIndex.dat and data_coor.out
 * (read_all_data.c:5-102) and writes msd.dat, relax_a_0.30.dat and
 * S4qt_a_0.30.dat (+ Fsqt_a_0.30.dat when do_fs).  Returns 0 on success. */
int or_analyse_dir(const char *sub_dir, int N, int do_msd, int do_overlap,
                   int do_s4, int do_fs);

/* Ensemble chi4 (no reference counterpart) from the per-lag sums each
 * replica accumulates, in the form Analysis/overlap.c:89-94,114-118 keeps
 * them: sq[r*nlag+l] = sum over origins of Q, sq2 of Q^2, cnt the origins.
 * Pooled over the replicas (summed in index order) and their origins:
 * out3[3l] = Nbar*(s2/c - s*s/(c*c)), out3[3l+1] = s/c, out3[3l+2] = c, with
 * Nbar the mean of n_r; zeros where c = 0. */
void or_chi4_ensemble(int nrep, int nlag, const double *sq, const double *sq2,
                      const int *cnt, const int *n_r, double *out3);
/* Analysis_Block/{msd_pos,overlap,vanHove}.c for one block edge cell_cut:
 * writes MSDLb, RelaxLb, RelaxTotLb and vanHove files under
 * sub_dir/BlockAnalysis/ (the directories must exist).  param = a*a. */
int or_block_scale(int N, double L, int nframes, const int *fidx, int gap,
                   const double *X, const double *Y, const double *Z,
                   double factor, double param, double cell_cut,
                   double vh_time, const char *sub_dir);
/* Analysis_Block/main.c:45-70 for one job directory (scale 3..19). */
int or_block_analyse_dir(const char *sub_dir, int N, double vh_time);

/* Analysis/guu.c:3-300 (the guuTauAlphaAct.dat branch): writes path when
 * not NULL; hist3 (nullable) receives hist, hist1, hist2 [3][maxbin]. */
int or_guu(int N, double L, int nframes, const int *fidx, int gap,
           const double *X, const double *Y, const double *Z, const int *act,
           double factor, double tau, const char *path, double *hist3,
           int *maxbin, int *frame_num, int *n_origins);
/* GetList of Analysis/overlap_cage_rel.c:185-251 on wrapped positions:
 * nn[i] neighbours nb[i*cap + k] in traversal order; -1 on overflow. */
int or_cage_list(int N, double L, double cell_cut, const double *x0,
                 const double *y0, const double *z0, int cap, int *nn, int *nb);
/* Analysis/overlap_cage_rel.c:4-184: writes path (relaxMPICR1 format) when
 * not NULL; pair_counts (nullable) receives the count of every pair. */
int or_overlap_cage(int N, double L, int nframes, const int *fidx, int gap,
                    const double *X, const double *Y, const double *Z,
                    double factor, double a, double cell_cut, const char *path,
                    int *pair_counts, int cap_pairs, int *npairs);

/* glibc baselines of the frame text (MDcode/print.c:10,
 * Analysis/read_all_data.c:81): sprintf / strtod per line. */
long long or_frame_printf(int n, const double *x, const double *y,
                          const double *z, const int *t, char *buf);
long long or_frame_strtod(const char *buf, int n, double *x, double *y,
                          double *z, int *t);

#endif
