/*
 * ac4.h -- C ABI of the MI355X-native activeChi4 hot path (libac4.so).
 *
 * Plain pointers and sizes only.  Every entry point returns an int status:
 * AC4_OK (0) or a negative AC4_E* code; ac4_last_error() gives the message.
 * Nothing calls exit().  Host arrays are caller-owned and copied in/out;
 * device state is owned by the handle.  One host thread drives one handle.
 *
 * The reference has no plugin or FFI layer: its boundary is two executables
 * (MDcode `run`, Analysis `analyse`) plus text files, and inside them the
 * global-state functions cited below.  Each entry point names the reference
 * routine it replaces (paths relative to the reference checkout).
 */
#ifndef AC4_H
#define AC4_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define AC4_OK 0
#define AC4_EARG -1      /* bad argument */
#define AC4_EHIP -2      /* HIP runtime error */
#define AC4_ECAP -3      /* capacity exceeded (neighbour list, cell) */
#define AC4_ESTATE -4    /* call out of order / missing state */
#define AC4_EIO -5       /* file I/O or parse error */

#define AC4_FP64 0       /* plain mode: every force in FP64 */
#define AC4_FP32 1       /* pair terms in FP32, summed per particle in FP64 */

#define AC4_NQ 162       /* lattice q vectors of Analysis/S4Qt.c:297-356 */

/* Text of the last error reported by any entry point of this thread. */
const char *ac4_last_error(void);
/* Number of visible HIP devices (0 when none): lets callers fail loudly. */
int ac4_device_count(void);

/* ------------------------------------------------------------------ MD --
 * One handle holds R >= 1 independent replicas (R = 1 is the plain single
 * run).  Replicas are packed into every launch: one launch sequence per step
 * advances all of them.  A replica's results do not depend on which other
 * replicas share the handle (per-replica block partition, fixed-order
 * reductions, counter-based RNG keyed by the replica's seed).              */
typedef struct {
    int n;            /* particles (MDcode/global.h:15, compile-time there) */
    double density;   /* DENSITY, para_file line 4 (MDcode/main.c:421-422) */
    double T;         /* thermostat temperature, para_file line 4 */
    double f0;        /* active force magnitude, para_file line 6 (:423-425) */
    int n_act;        /* active particles; < 0 -> (int)(n*0.1) (global.h:55-56) */
    int tau_per;      /* steps between direction shuffles (global.h:58) */
    uint64_t seed;    /* Philox-4x32-10 key of this replica's RNG stream */
} ac4_replica_params;

typedef struct ac4_md ac4_md;

/* Allocates device state for all replicas.  precision: AC4_FP64 / AC4_FP32.
 * Replaces the set-up of MDcode/initialise.c:203-227 + Parallel.c:6-106
 * (the whole box of every replica is resident on one GPU). */
int ac4_md_create(int n_rep, const ac4_replica_params *p, int precision,
                  ac4_md **out);
int ac4_md_destroy(ac4_md *md);
/* Number of particles of every replica, summed. */
int ac4_md_total_particles(const ac4_md *md);

/* State of replica r in the reference's particle order (SPHERE fields of
 * MDcode/global.h:62-68): wrapped box-scaled x,y,z in [0,1), unwrapped
 * box-scaled x_u,y_u,z_u, momenta p (m = 1), type (0 = A, 1 = B), actTag
 * (-1 or active index), and the active directions kx,ky,kz (+-1, length
 * n_act; MDcode/active.c:30-41).  set_state also resets the neighbour-list
 * reference positions (xold = x_u), marks the list stale and restarts the
 * replica's step counter at 0 (the shuffle schedule counts from there).  Pass NULL for
 * act_tag/k* to keep the current ones. */
int ac4_md_set_state(ac4_md *md, int r, const double *x, const double *y,
                     const double *z, const double *xu, const double *yu,
                     const double *zu, const double *px, const double *py,
                     const double *pz, const int *type, const int *act_tag,
                     const int *kx, const int *ky, const int *kz);
int ac4_md_get_state(ac4_md *md, int r, double *x, double *y, double *z,
                     double *xu, double *yu, double *zu, double *px,
                     double *py, double *pz, int *kx, int *ky, int *kz);
/* Forces of replica r (real units) from the last force evaluation. */
int ac4_md_get_forces(ac4_md *md, int r, double *fx, double *fy, double *fz);
/* Nose-Hoover chain {eta[3], p_eta[3], Q[3]} of replica r
 * (MDcode/time_evolve.c:476-493; the caller draws p_eta). */
int ac4_md_set_chain(ac4_md *md, int r, const double *chain9);
int ac4_md_get_chain(ac4_md *md, int r, double *chain9);

/* update_cell_list() (MDcode/cell_list+forces.c:58-148): bins every replica
 * with the reference's cell grid (Parallel.c:72-80) and builds the neighbour
 * lists at ((dx*dx+dy*dy)+dz*dz)*L*L < 2.85^2.  Optional outputs for replica
 * r_out (reference particle order): cell index and the full neighbour count
 * (number of list partners, both directions). */
int ac4_md_update_cell_list(ac4_md *md, int r_out, int *cell_out,
                            int *full_count_out);
/* calculate_forces() (MDcode/cell_list+forces.c:153-227) on the current
 * lists; out2 per replica = {pe, virial/3} (reference sign conventions). */
int ac4_md_calculate_forces(ac4_md *md, double *out2);
/* nsteps x { ShuffleActDirec() when step % tau_per == 0 (MDcode/main.c:245-247);
 * AdvanceTimeNVTNoseHoover() (MDcode/time_evolve.c:495-526) } for every
 * replica.  Lists are rebuilt on device when 4*maxD^2 > skin^2 (safe form of
 * time_evolve.c:508); criterion = 1 selects the reference's 3*maxD^2.
 * Requires forces (ac4_md_calculate_forces or a previous step). */
int ac4_md_step(ac4_md *md, int nsteps);
/* Same as ac4_md_step but replays captured hipGraphs: 25 steps per graph
 * launch, then one-step graphs for the rest of nsteps (batched serving path);
 * the result equals ac4_md_step's bit for bit. */
int ac4_md_step_graph(ac4_md *md, int nsteps);
/* nsteps steps of each of nh distinct handles, graph replays interleaved on
 * their own streams so their kernels overlap on the device ("replicas per
 * launch" below the whole ensemble).  No reference counterpart. */
int ac4_md_step_graph_multi(ac4_md *const *mds, int nh, int nsteps);
/* Replicas split into `groups` launch groups of balanced particle count
 * (largest first, each to the least loaded group; ties to the lower group):
 * group_of[r] in [0, groups).  No reference counterpart. */
int ac4_group_plan(int n_rep, const ac4_replica_params *p, int groups, int *group_of);
/* Neighbour-list rebuild test of every replica (MDcode/time_evolve.c:508):
 * 0 (default) rebuilds when 4*max|x_u - x_old|^2*L^2 > skin^2 (no two
 * particles can close the skin before it fires); 1 is the reference's
 * 3*max|x_u - x_old|^2*L^2 > skin^2 (skin 0.35 there). */
int ac4_md_set_rebuild_criterion(ac4_md *md, int criterion);
/* coldstart()/breakxal() (MDcode/initialise.c:166-201): nsteps of velocity
 * Verlet with the momenta rescaled every step to a target temperature ramped
 * linearly from f_start*T to f_end*T (each replica its own T); no thermostat
 * chain, no direction shuffles, step counter unchanged.  Used to bring a
 * lattice start to temperature before Nose-Hoover dynamics. */
int ac4_md_relax(ac4_md *md, int nsteps, double f_start, double f_end);
/* NVE for replica r (on != 0): plain velocity Verlet, the reference's
 * AdvanceTimeNVE (MDcode/time_evolve.c:39-171) in the step's U2-U1-F-U2 order,
 * for energy-drift checks; Heff reported as KE + PE.  on == 0 returns to the
 * Nose-Hoover chain with its variables as they were. */
int ac4_md_set_nve(ac4_md *md, int r, int on);
/* ShuffleActDirec() (MDcode/active.c:45-85) now, for every replica, with the
 * device Philox stream; shuffle counter advances by one. */
int ac4_md_shuffle(ac4_md *md);
/* The uniforms the next shuffle of replica r will draw: 3*n_act doubles
 * (component-major), for checking the swap rule against the oracle. */
int ac4_md_shuffle_uniforms(ac4_md *md, int r, double *u);
/* print_stuff() quantities (MDcode/print.c:60-98) for replica r:
 * out8 = {pe, virial, ke, inst_T, pressure, drift (sum px), Heff, steps}. */
int ac4_md_stats(ac4_md *md, int r, double *out8);
/* Neighbour-list rebuilds performed by replica r so far. */
int ac4_md_rebuilds(ac4_md *md, int r);
/* Measurement hook (no reference counterpart): nsteps eager steps timed with
 * HIP events on the handle's stream around each launch group; ms (5 doubles)
 * = average milliseconds per step of {first half, rebuild group, force pass,
 * shuffle, force end (per-replica sums and chain update)}.  Where the handle
 * runs the shuffle inside the first-half kernel (small active sets), the
 * shuffle's time is part of the first entry. */
int ac4_md_kernel_times(ac4_md *md, int nsteps, double *ms);
/* Waits for all device work of the handle; reports deferred device errors
 * (list/cell overflow). */
int ac4_md_sync(ac4_md *md);
/* Raw Philox-4x32-10 blocks (known-answer tests): out4 = philox(ctr4, key2). */
int ac4_philox4x32(const uint32_t *ctr4, const uint32_t *key2, uint32_t *out4);

/* Kernel shape of handles created afterwards (set by the autotuner; no
 * reference counterpart).  tpp = lanes per particle in the force kernel
 * (1,2,4,8,16), or 0 = automatic from the handle's total capacity (default).
 * The team width fixes the summation order of each particle's force, so a
 * replica's trajectory is bitwise reproducible for a given handle shape.
 * nlist_mode = dynamics list builder: 0 wave per cell, 1 thread per particle,
 * 2/3/4 teams of 16/32/64 lanes per particle (3 is the default).  All
 * builders produce the same lists in the same order. */
int ac4_set_tuning(int tpp, int nlist_mode);
/* Verlet skin of handles created afterwards, in [0.2, 0.6] (default 0.35 =
 * 2.85 - 2.5, MDcode/global.h:25-31).  Sets the per-species list radii
 * 2.5 sigma_ab + skin, the rebuild criterion and the list-cell size. */
int ac4_set_skin(double skin);

/* ------------------------------------------------- device frame store --
 * A trajectory of one replica kept in HBM (x_u, box-scaled, reference
 * order), so analyses can pair any origin with any lag after the run.  No
 * reference counterpart: the reference writes data_coor.out and rereads it
 * (MDcode/print.c:4-16, Analysis/read_all_data.c:61-91).  A frame of N
 * particles takes 24*N bytes of device memory. */
typedef struct ac4_frames ac4_frames;
int ac4_frames_create(int n, double L, int cap, ac4_frames **out);
int ac4_frames_destroy(ac4_frames *f);
int ac4_frames_push(ac4_frames *f, const double *xu, const double *yu, const double *zu, long long step);
/* Appends replica r's current x_u (device to device; synchronises the MD stream). */
int ac4_frames_push_md(ac4_frames *f, ac4_md *md, int r, long long step);
/* Appends a frame from device arrays (e.g. one ac4_textframes_parse made). */
int ac4_frames_push_device(ac4_frames *f, const double *xu, const double *yu, const double *zu, long long step);
int ac4_frames_count(const ac4_frames *f);
/* Index of the frame pushed with this step number, or -1. */
int ac4_frames_find(const ac4_frames *f, long long step);
/* Overlap count between frames i0 (origin) and i1 with the arithmetic of
 * Analysis/overlap.c:61-84. */
int ac4_frames_overlap(ac4_frames *f, int i0, int i1, double a, int *count);
/* ndir unit directions (n3[3*d..3*d+2]) drawn from seed, uniform on the sphere. */
int ac4_shell_directions(int ndir, uint64_t seed, double *n3);
/* S4(q, t) on |q| shells, the overlap-weighted four-point structure factor
 * of Analysis/S4Qt.c:4-267 with seeded random directions in place of the
 * lattice vectors of GetQ.  Pair p = (origin frame orig[p], lag frame
 * lagf[p]); |q_k| = q0 + k*dq for k < nshell; ndir directions from seed
 * (ndir * ceil(nshell/27) <= 256).  s4_ref[k] = ndir*(Var_p Cbar + Var_p
 * Sbar)/N with Cbar the direction mean (S4Qt's structure, the directions as
 * its permutations); s4_dir[k] = mean_d (Var_p C + Var_p S)/N.  phases
 * (nullable): [npairs][ndir][nshell][2] sums C, S.  Each may be NULL. */
int ac4_frames_s4_shells(ac4_frames *f, int npairs, const int *orig, const int *lagf, double a, double q0,
                         double dq, int nshell, int ndir, uint64_t seed, double *s4_ref, double *s4_dir,
                         double *phases);
/* ac4_frames_s4_shells that allocates nothing, device or host: no reference
 * counterpart.  For callers that cannot allocate (tight analysis loops in a
 * serving process, code that must not call hipMalloc).  Its per-block
 * partials, per-pair sums and directions sit in one file-scope device array
 * of 26.7 MB that each GPU holds until the process ends, sized for the
 * largest shape ac4_frames_s4_shells accepts (ndir * ceil(nshell/27) <= 256,
 * so ndir * nshell <= 6912; larger shapes are refused with AC4_EARG, and a
 * shape over the array with AC4_ECAP); phases is required
 * (npairs*ndir*nshell*2 doubles, caller-owned).  Results equal
 * ac4_frames_s4_shells bit for bit.  Calls that overlap on one GPU share the
 * array and race. */
int ac4_frames_s4_shells_noalloc(ac4_frames *f, int npairs, const int *orig, const int *lagf, double a,
                                 double q0, double dq, int nshell, int ndir, uint64_t seed,
                                 double *s4_ref, double *s4_dir, double *phases);

/* Block analysis of the stored trajectory for one block edge: replaces
 * Analysis_Block/msd_pos.c:3-240 (mean_square_displacement), overlap.c:4-232
 * (overlap_correlation_function) and vanHove.c:3-183 (VanHove), which
 * Analysis_Block/main.c:53-63 calls for cellCutOff = L/scale, scale = 3..19.
 * The frames must have been pushed in the order of time_index_file.dat, with
 * their step numbers; gap is the origin spacing (ac4_origin_gap), factor the
 * lag ratio (1.4), a the overlap length (0.3), cell_cut the nominal block
 * edge (the grid has (int)(L/cell_cut) blocks per edge, at most 25), vh_time
 * the van Hove time (1500.0 in the reference).  Writes, under
 * sub_dir/BlockAnalysis/ (created when missing), MSD/MSDLb<Lb>.dat,
 * Relax/RelaxLb<Lb>.dat, Relax/RelaxTotLb<Lb>.dat and vanHove<Lb>.dat in the
 * reference's formats, Lb = L/blocks per edge as "%.3lf". */
int ac4_frames_block(ac4_frames *f, int gap, double factor, double a, double cell_cut, double vh_time,
                     const char *sub_dir);

/* Cage-relative overlap of the stored trajectory: replaces
 * Analysis/overlap_cage_rel.c:4-184 (overlap_cage; its call at
 * Analysis/main.c:57 is commented out and the file is not in
 * Analysis/makefile).  Each particle's displacement is taken relative to the
 * mean displacement of its neighbours within 1.3 sigma at the origin,
 * listed in the order of GetList (:185-251) on a grid of (int)(L/cell_cut)
 * cells per edge (at least 3; cell_cut 2.85 is MDcode's Cell_cut); the
 * overlap count uses a*a.  Origins and lags as ac4_frames_block.  Writes path
 * (nullable) in the relaxMPICR1_a_<a>.dat format (:156-167); pair_counts
 * (nullable, cap_pairs entries) receives the count of every (origin, lag)
 * pair in walk order, *npairs (nullable) their number. */
int ac4_frames_overlap_cage(ac4_frames *f, int gap, double factor, double a, double cell_cut, const char *path,
                            int *pair_counts, int cap_pairs, int *npairs);
/* Displacement-displacement correlation g_uu(r) of the active particles
 * (act_tag[i] == 1, ActiveFileIndex.dat) at the lag nearest tau on the first
 * frame's ladder: replaces Analysis/guu.c:3-300 (the guuTauAlphaAct.dat
 * branch; its call at Analysis/main.c:64 is commented out).  Bins of 0.005
 * up to L/2: maxbin = (int)(0.5*L/0.005) with L the store's box edge
 * (guu.c:17).  Writes path (nullable) in the reference's format; hist3
 * (nullable, 3*maxbin doubles) receives hist, hist1, hist2; maxbin,
 * frame_num (the lag in steps) and n_origins (origins used) are returned
 * when not NULL. */
int ac4_frames_guu(ac4_frames *f, const int *act_tag, int gap, double factor, double tau, const char *path,
                   double *hist3, int *maxbin, int *frame_num, int *n_origins);

/* The data_coor.out text of replica r's current frame (print_data,
 * MDcode/print.c:4-16: x_u y_u z_u type in reference order), formatted on the
 * device: byte for byte the text ac4_format_frame writes for the same values
 * (a frame holding values of 1e12 or more is formatted by it, with its
 * AC4_ECAP).  out needs 80*N bytes; returns the bytes written or a negative
 * AC4_E* code. */
long long ac4_md_format_frame(ac4_md *md, int r, char *out, long long cap);

/* fixDrift (MDcode/initialise.c:104-138): removes the mean momentum of
 * replica r (summed in reference particle order, divided by N) from every
 * particle; mean3 (nullable) receives the mean removed.  The reference calls
 * it every 1e5 steps of its run loops (MDcode/main.c:188,261). */
int ac4_md_fix_drift(ac4_md *md, int r, double *mean3);

/* ------------------------------------------------------ exact checkpoints --
 * A superset of MDcode's restart_file.dat (main.c:272-334), which keeps only
 * positions, velocities and tags: the complete device state of replica r
 * (particles in cell order, forces, last-rebuild positions, cell tables,
 * neighbour list, active directions, Nose-Hoover chain and step/shuffle
 * counters), so a run resumed from it is bitwise the uninterrupted run (the
 * reference's restart is not exact: it redraws the chain momenta).  Restore
 * into a slot configured with the same replica parameters in a handle of the
 * same shape (precision, team width, skin); other replicas of the handle are
 * not touched.  Sizes in bytes. */
long long ac4_md_checkpoint_size(ac4_md *md, int r);
int ac4_md_checkpoint(ac4_md *md, int r, void *buf, long long cap);
int ac4_md_restore(ac4_md *md, int r, const void *buf, long long len);

/* ------------------------------------------------------ slots (serving) --
 * A handle can be created with idle slots of fixed capacity; jobs are
 * assigned to slots between launches and released when done.  No reference
 * counterpart (the reference runs one job per process). */
int ac4_md_create_slots(int nslots, const int *cap_n, int precision, ac4_md **out);
/* Configures slot `slot` for a job of p->n <= capacity particles (state
 * follows with set_state/load_slot). */
int ac4_md_assign(ac4_md *md, int slot, const ac4_replica_params *p);
int ac4_md_release(ac4_md *md, int slot);
/* Particles of the job in a slot (0 when idle). */
int ac4_md_slot_n(const ac4_md *md, int slot);
/* First global index of a slot in the export buffers of ac4_md_export_all. */
int ac4_md_slot_offset(const ac4_md *md, int slot);
/* set_state for one slot without touching the others: its forces start at
 * zero and its list is built inside its first step. */
int ac4_md_load_slot(ac4_md *md, int r, const double *x, const double *y,
                     const double *z, const double *xu, const double *yu,
                     const double *zu, const double *px, const double *py,
                     const double *pz, const int *type, const int *act_tag,
                     const int *kx, const int *ky, const int *kz);
/* ac4_md_relax for one slot, applied during its next nsteps steps. */
int ac4_md_relax_slot(ac4_md *md, int r, int nsteps, double f_start, double f_end);
/* Device bytes one slot of capacity cap_n adds to a handle (with in-situ
 * overlap of n_origin origins and n_lags lags; 0, 0 without). */
long long ac4_md_slot_bytes(int cap_n, int n_origin, int n_lags);
/* Error bits of every slot: err[r] bit 0 list or cell over capacity, bit 1
 * non-finite or runaway coordinate (0 for idle slots). */
int ac4_md_errors(ac4_md *md, int *err);
/* Sets error bits on replica r as the kernels do on an overflow or a
 * runaway coordinate, with none of those happening (tests of the callers'
 * failure handling). */
int ac4_md_raise_error(ac4_md *md, int r, int bits);
/* x_u of every slot in reference order (arrays of ac4_md_total_particles). */
int ac4_md_export_all(ac4_md *md, double *xu, double *yu, double *zu);

/* In-situ overlap Q(t) for every replica with device-resident accumulators
 * (running Q/chi4 accumulators; the arithmetic of each count is
 * Analysis/overlap.c:61-84).  Origins every origin_every steps (a ring of
 * n_origin snapshots per replica), lags l*sample_every for l = 0..n_lags.
 * Call ac4_md_overlap_sample whenever the step counter is a multiple of
 * sample_every (e.g. after ac4_md_step_graph(md, sample_every)). */
int ac4_md_overlap_enable(ac4_md *md, int sample_every, int origin_every, int n_origin,
                          int n_lags, double a);
int ac4_md_overlap_sample(ac4_md *md);
/* sq[l] = sum over origins of Q, sq2[l] of Q^2, cnt[l] origins (n_lags+1). */
int ac4_md_overlap_get(ac4_md *md, int r, double *sq, double *sq2, int *cnt);
/* Ensemble chi4 on the device in FP64, pooled over the nlist replicas of
 * `list` (summed in list order) and their origins: for l = 0..n_lags,
 * out3[3l] = Nbar*(<Q^2> - <Q>^2), out3[3l+1] = <Q>, out3[3l+2] = origins,
 * with <Q^2> - <Q>^2 formed as Analysis/overlap.c:114-118 forms X4 (which
 * has no factor N) and Nbar the mean N of the listed replicas.  chi4 as an
 * ensemble fluctuation over independent replicas has no reference
 * counterpart; the reference's commented-out ensemble mean is
 * Analysis/avg_and_print.c:30-41. */
int ac4_md_overlap_chi4(ac4_md *md, int nlist, const int *list, double *out3);
/* Clears replica r's accumulators and origins (a slot starting a new job). */
int ac4_md_overlap_reset(ac4_md *md, int r);

/* Request loop (continuous batching of MD jobs on one GPU). */
typedef struct {
    int n;           /* particles (80:20 KA at density 1.2) */
    double T, f0;    /* temperature, active force */
    double tau_p;    /* persistence time (time units; tau_per = tau_p/dt) */
    long steps;      /* production steps */
    uint64_t seed;   /* initial state + RNG key */
    double t_arrive; /* arrival time (seconds, caller's clock) */
} ac4_job;
typedef struct ac4_server ac4_server;
/* Slots of the given capacities; chunk = steps between frames;
 * relax_steps = velocity-rescale ramp before each job's production steps
 * (rounded up to whole chunks).  The slots are split into launch groups of
 * balanced capacity, one MD handle and stream each, stepped together (three
 * groups when the slots hold at least 3 x 16384 particles, else one); every
 * group uses the kernel shape of the server's total capacity, so a job's
 * results do not depend on its group. */
int ac4_server_create(int nslots, const int *slot_cap, int precision, int chunk,
                      int relax_steps, ac4_server **out);
/* ac4_server_create with `groups` launch groups (0 = as ac4_server_create
 * chooses; at most 64, and at most one per slot). */
int ac4_server_create_groups(int nslots, const int *slot_cap, int precision, int chunk,
                             int relax_steps, int groups, ac4_server **out);
/* Slot capacities from an HBM byte budget for requests with N in
 * [n_min, n_max]: classes a factor 2 apart, slot counts per class in
 * proportion to the share of a log-uniform N each class takes (at least one),
 * as many as hbm_bytes holds at ac4_md_slot_bytes(cap, n_origin, n_lags) per
 * slot.  Writes at most caps_cap capacities; returns the slot count. */
int ac4_server_plan(double hbm_bytes, int n_min, int n_max, int n_origin, int n_lags,
                    int *caps, int caps_cap);
/* ac4_server_plan + ac4_server_create (+ ac4_server_enable_overlap when
 * n_origin > 0). */
int ac4_server_create_budget(double hbm_bytes, int n_min, int n_max, int precision,
                             int chunk, int relax_steps, int origin_every, int n_origin,
                             int n_lags, double a, ac4_server **out);
/* Every job's overlap Q(t) accumulated on the device from its frames (one
 * per chunk; ac4_md_overlap_enable with sample_every = chunk), from its own
 * step counter; call before the first submit. */
int ac4_server_enable_overlap(ac4_server *s, int origin_every, int n_origin, int n_lags,
                              double a);
/* Slot count; caps (may be NULL) receives the capacities. */
int ac4_server_slots(const ac4_server *s, int *caps);
int ac4_server_destroy(ac4_server *s);
/* Queues a job; returns its id (>= 0) or an error. */
int ac4_server_submit(ac4_server *s, const ac4_job *job);
/* Admits arrived jobs (t_arrive <= now) into free slots, advances all slots
 * one chunk, counts each running job's frame into its Q(t) accumulators,
 * fails the jobs whose replica raised an error flag (the others go on) and
 * retires finished jobs.  Returns the number of jobs that ran in the chunk. */
int ac4_server_tick(ac4_server *s, double now);
int ac4_server_pending(const ac4_server *s);
/* Preemption (no reference counterpart): takes running job `id` off the GPU
 * with an exact checkpoint (ac4_md_checkpoint) held in host memory and frees
 * its slot; the job is queued again behind the jobs that have not started
 * and, when admitted, continues bitwise as if it had never been interrupted. */
int ac4_server_preempt(ac4_server *s, int id);
/* state: 0 queued, 1 running, 2 done, 3 failed (ac4_server_job_error gives
 * its error bits: 1 list or cell over capacity, 2 runaway coordinate). */
int ac4_server_job(const ac4_server *s, int id, double *t_start, double *t_done,
                   long *steps_done, int *frames, int *state);
int ac4_server_job_error(const ac4_server *s, int id);
/* Final unwrapped positions (reference order) of a finished job. */
int ac4_server_result(const ac4_server *s, int id, double *xu, double *yu, double *zu);
/* Per-lag overlap sums of a finished job (as ac4_md_overlap_get). */
int ac4_server_job_overlap(const ac4_server *s, int id, double *sq, double *sq2, int *cnt);
long long ac4_server_particle_steps(const ac4_server *s);
/* The first launch group's MD handle, the handle of the group holding a slot
 * and the slot's index in it (error injection in tests), the number of launch
 * groups, and the slot a running job occupies (-1 when it is not running). */
ac4_md *ac4_server_handle(ac4_server *s);
ac4_md *ac4_server_slot_handle(ac4_server *s, int slot, int *local);
int ac4_server_groups(const ac4_server *s);
int ac4_server_job_slot(const ac4_server *s, int id);

/* ------------------------------------------------------------ Analysis --
 * Streaming per-frame analysis over all time origins (Analysis/main.c:53-62).
 * The (origin, lag) walk of Analysis/msd_pos.c:44-107 / overlap.c:42-106 is
 * computed on the host from the frame index list; frames are pushed in file
 * order, origin frames stay resident, every pushed frame is compared on the
 * GPU with all origins whose walk it completes.                            */
typedef struct ac4_ana ac4_ana;

#define AC4_OBS_MSD 1
#define AC4_OBS_OVERLAP 2
#define AC4_OBS_S4 4
#define AC4_OBS_FS 8

/* fidx: frame step indices (time_index_file.dat, read_all_data.c:14-31);
 * gap: origin spacing (0 -> detected as read_all_data.c:36-48);
 * factor: lag growth (1.4, Analysis/main.c:4); a: overlap length (0.3);
 * obs: AC4_OBS_* mask; type_flag: the reference's `flag` argument
 * (1 -> exclude type-1 particles, msd_pos.c:62, overlap.c:59, Fsqt.c:104). */
int ac4_ana_create(int n, double L, int nframes, const int *fidx, int gap,
                   double factor, double a, int obs, int type_flag,
                   ac4_ana **out);
int ac4_ana_destroy(ac4_ana *an);
int ac4_ana_gap(const ac4_ana *an);
/* Frame at position `pos` of the index list (LoadFramePP, read_all_data.c:129-150),
 * unwrapped box-scaled coordinates, host arrays of n. Frames must arrive in
 * increasing pos. */
int ac4_ana_push_frame(ac4_ana *an, int pos, const double *x, const double *y,
                       const double *z, const int *type);
/* Same from device arrays (in-situ analysis of a running MD handle). */
int ac4_ana_push_frame_device(ac4_ana *an, int pos, const double *dx,
                              const double *dy, const double *dz,
                              const int *dtype);
/* Pushes replica r's current x_u of an MD handle as frame `pos` (in-situ,
 * no host round trip). */
int ac4_ana_push_md(ac4_ana *an, int pos, ac4_md *md, int r);
/* Accumulates the per-(origin,lag) results into slots in the reference's
 * origin order (msd_pos.c:87-93, overlap.c:89-94, S4Qt.c:228-236). */
int ac4_ana_finalize(ac4_ana *an);
/* Slot arrays (length nslots; slot 0 unused as in the reference):
 * MSD: m0=<r^2>, m1=<sum_a da^4>, m2=<r^4> sums, counters, delta_t. */
int ac4_ana_nslots(const ac4_ana *an, int obs);
int ac4_ana_get_msd(ac4_ana *an, double *m0, double *m1, double *m2,
                    int *counters, double *delta_t, int *num_intervals);
int ac4_ana_get_overlap(ac4_ana *an, double *c, double *x4, int *counters,
                        double *delta_t, int *num_intervals);
/* S4 / Fs slot arrays are [nslots][162]. qval[162] = Q column (|q|^2). */
int ac4_ana_get_phase(ac4_ana *an, int obs, double *cs, double *ss,
                      double *cs2, double *ss2, int *counters,
                      double *delta_t, double *qval);
/* Per-pair raw results in walk order (rule of `obs`): MSD 3 sums, overlap
 * count, S4 masked count, S4/Fs {cos[162], sin[162]}.  S4's phase sums are
 * formed per origin, over all of its lags at once: an origin's pairs hold
 * them once its last lag frame has been pushed (or after finalize); until
 * then they read zero. */
int ac4_ana_npairs(const ac4_ana *an, int obs);
int ac4_ana_get_pairs(ac4_ana *an, int obs, double *vals, int *ivals);
/* Test hook (no reference counterpart): the host build of the sin/cos the
 * S4 and Fs kernels evaluate, s[i], c[i] = sin, cos(2*pi*f[i]) for phases f
 * in turns (|f| < 2^28); the device code is the same source. */
int ac4_sincos_turns_host(const double *f, int n, double *s, double *c);
/* Writes msd.dat, relax_a_%.2lf.dat, S4qt_a_%.2lf.dat, Fsqt_a_%.2lf.dat into
 * sub_dir with the reference's formats (msd_pos.c:112-120, overlap.c:108-121,
 * S4Qt.c:251-271, Fsqt.c:257-278). */
int ac4_ana_write(ac4_ana *an, const char *sub_dir);

/* ------------------------------------------------------ host-side logic */
/* prepareSavingArray (MDcode/main.c:62-119) incl. its shellSort (:17-41):
 * returns the frame count; out[0..count] as the run loop reads it. */
int ac4_schedule(uint64_t run_length, int n_origins, double factor,
                 uint64_t *out, int cap);
/* Seeded initial state in the reference's SPHERE layout: simple-cubic sites
 * chosen at random (MDcode/initialise.c:34-62), 80:20 species (:63-72),
 * Maxwell-Boltzmann momenta at T with zero drift, n_act active particles with
 * the initial +-(1,1,1) directions of MDcode/active.c:24-41.  Replaces the
 * time()-seeded GetFccGrid/selectRandomActiveParticles; xu/act_tag/k* may be
 * NULL. */
int ac4_init_lattice(int n, double T, int n_act, uint64_t seed, double *x, double *y, double *z, double *xu,
                     double *yu, double *zu, double *px, double *py, double *pz, int *type, int *act_tag, int *kx,
                     int *ky, int *kz);
/* n lines of data_coor.out, "%.5lf\t%.5lf\t%.5lf\t%d\n" (MDcode/print.c:4-16),
 * byte-identical to glibc's printf; out needs 80*n bytes, which holds every
 * frame whose values are below 1e12 in magnitude.  A frame holding larger or
 * non-finite values is measured first and gives AC4_ECAP when its text does
 * not fit cap.  Returns the bytes written or a negative error. */
long long ac4_format_frame(int n, const double *x, const double *y, const double *z, const int *type, char *out,
                           long long cap);
/* Reads n such lines (Analysis/read_all_data.c:61-91); every number equals
 * strtod's.  Returns the bytes consumed or a negative error. */
long long ac4_parse_frame(const char *text, long long len, int n, double *x, double *y, double *z, int *type);
/* The same reading of data_coor.out on the GPU (replaces the fscanf loop of
 * LoadFramePP, Analysis/read_all_data.c:65-100, and its Analysis_Block
 * twin): the caller reads the file into the handle's pinned chunk
 * (ac4_textframes_buffer, chunk_bytes of it; 0 -> 64 MiB, at least 128 bytes
 * per line of a frame), then ac4_textframes_parse converts the whole frames
 * of n lines in its first len bytes, at most max_frames of them: *nframes
 * frames, *used bytes consumed (the caller keeps the rest for the next
 * call).  Values equal ac4_parse_frame's.  ac4_textframes_get gives frame k
 * of the last call as device arrays, valid until the next call. */
typedef struct ac4_textframes ac4_textframes;
int ac4_textframes_create(int n, long long chunk_bytes, ac4_textframes **out);
char *ac4_textframes_buffer(ac4_textframes *t, long long *cap);
int ac4_textframes_parse(ac4_textframes *t, long long len, int max_frames, int *nframes, long long *used);
int ac4_textframes_get(ac4_textframes *t, int k, const double **x, const double **y, const double **z,
                       const int **type);
/* Frame k of the last call copied to host arrays of n. */
int ac4_textframes_copy(ac4_textframes *t, int k, double *x, double *y, double *z, int *type);
int ac4_textframes_destroy(ac4_textframes *t);
/* read_all_data.c:36-48 */
int ac4_origin_gap(const int *fidx, int nframes);
/* InitialiseNoseHoover (MDcode/time_evolve.c:476-493) with a seeded draw:
 * chain9 = {eta[3] = 0, p_eta[3] ~ N(0, T*Q), Q[3]}. */
void ac4_chain_init(int n, double T, uint64_t seed, double *chain9);
/* Autotuner (no reference counterpart): from one relaxed start of the given
 * replica set, times `steps` graph-replayed steps for every kernel shape
 * (force lanes per particle 1..16, i.e. 256..16 particles per workgroup;
 * neighbour-list builder), then the skin for the best shape, then the
 * replicas per launch (the set in 1..4 launch groups of balanced particle
 * count, ac4_group_plan, stepped together with ac4_md_step_graph_multi), and
 * keeps the fastest shape with ac4_set_tuning / ac4_set_skin.  best4 = {tpp,
 * nlist_mode, skin, groups}; table (rows of {tpp, nlist_mode, skin, groups,
 * particle-steps/s}) may be NULL. */
int ac4_autotune(int n_rep, const ac4_replica_params *p, int precision, int relax_steps, int steps,
                 double *best4, double *table, int table_cap, int *n_cand);

#ifdef __cplusplus
}
#endif
#endif
