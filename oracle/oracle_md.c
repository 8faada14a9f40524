/*
 * oracle_md.c -- CPU restatement of the reference MD hot path (TEST
 * INFRASTRUCTURE ONLY: the checker, never the product).
 *
 * Single process (npx=npy=npz=1, node=0, Block[i]=0, InMe[i]=true), so every
 * MPI collective of the reference reduces to the local value and factor=1 in
 * calculate_forces.  Expression order follows the reference token by token;
 * compile with -ffp-contract=off.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* MDcode/global.h:24-45 */
#define OR_SIGMA        1.00
#define OR_LCUT         (2.5 * OR_SIGMA)
#define OR_LCUTSQR      (6.25 * OR_SIGMA * OR_SIGMA)
#define OR_SKIN         (.35 * OR_SIGMA)
#define OR_CELL_CUT     (OR_LCUT + OR_SKIN)
#define OR_CELL_CUT_SQ  (OR_CELL_CUT * OR_CELL_CUT)
#define OR_DT           0.005
#define OR_HALF_DT      (OR_DT * 0.5)
#define OR_C0           0.0162665594880
#define OR_C2           -0.001949973872640
#define OR_M            3
#define OR_MAX_IN_CELL  600

/* MDcode/cell_list+forces.c:19-20 */
static const double or_sigma[3][3] = {{1.00, 0.80, 1.0}, {0.80, 0.88, 1.0}, {1.0, 1.0, 1.0}};
static const double or_eps[3][3] = {{1.0, 1.5, 1.0}, {1.5, 0.50, 1.0}, {1.0, 1.0, .50}};

static double dof_of(int n) { return 3.0 * (double)n - 3.0; } /* global.h:35 */

double or_md_box_length(int n, double density) {
    return pow((double)n / density, 0.33333333); /* MDcode/main.c:465 */
}

double or_ana_box_length(int n) {
    return pow((double)n / 1.2, 1.0 / 3.0); /* Analysis/main.c:15, RHO=1.2 */
}

void or_cells_setup(double L, or_cells *c) {
    /* Parallel.c:31-33: Lpp = L/(double)np with np = 1 */
    double Lpp = L / (double)1;
    /* Parallel.c:72-80: (int)Lpp/Cell_cut is ((int)Lpp)/2.85, truncated on
     * assignment to int. */
    int nc = (int)((double)((int)Lpp) / OR_CELL_CUT);
    double cs = Lpp / (double)nc;
    double inv = L / cs;
    c->ncx = c->ncy = c->ncz = nc;
    c->inv_cx = c->inv_cy = c->inv_cz = inv;
    c->ncell = nc * nc * nc; /* Parallel.c:82-87 with np = 1 */
}

int or_cell_of(const or_cells *c, double x, double y, double z) {
    /* cell_list+forces.c:73-74 (double arithmetic, then truncation) */
    double v = floor(x * c->inv_cx) + floor(y * c->inv_cy) * c->ncx +
               floor(z * c->inv_cz) * c->ncx * c->ncy;
    return (int)v;
}

static int getcell(const or_cells *c, int ix, int iy, int iz) {
    /* cell_list+forces.c:24-28 */
    return ((ix + c->ncx * 3) % c->ncx) + ((iy + c->ncy * 3) % c->ncy) * c->ncx +
           ((iz + c->ncz * 3) % c->ncz) * c->ncx * c->ncy;
}

void or_maps(const or_cells *c, int *map) {
    /* cell_list+forces.c:32-54 */
    for (int ix = 0; ix < c->ncx; ix++)
        for (int iy = 0; iy < c->ncy; iy++)
            for (int iz = 0; iz < c->ncz; iz++) {
                int *m = map + 13 * getcell(c, ix, iy, iz);
                m[0] = getcell(c, ix + 1, iy, iz);
                m[1] = getcell(c, ix + 1, iy + 1, iz);
                m[2] = getcell(c, ix, iy + 1, iz);
                m[3] = getcell(c, ix - 1, iy + 1, iz);
                m[4] = getcell(c, ix + 1, iy, iz - 1);
                m[5] = getcell(c, ix + 1, iy + 1, iz - 1);
                m[6] = getcell(c, ix, iy + 1, iz - 1);
                m[7] = getcell(c, ix - 1, iy + 1, iz - 1);
                m[8] = getcell(c, ix + 1, iy, iz + 1);
                m[9] = getcell(c, ix + 1, iy + 1, iz + 1);
                m[10] = getcell(c, ix, iy + 1, iz + 1);
                m[11] = getcell(c, ix - 1, iy + 1, iz + 1);
                m[12] = getcell(c, ix, iy, iz + 1);
            }
}

/* minimum image as cell_list+forces.c:92-97 */
static inline double minimg_md(double d) {
    if (d < -0.5) d += 1.0;
    if (d > 0.5) d -= 1.0;
    return d;
}

static inline double dr2_md(double dx, double dy, double dz, double L) {
    return (dx * dx + dy * dy + dz * dz) * L * L; /* cell_list+forces.c:98 */
}

int or_half_list(const or_cells *c, int n, const double *x, const double *y,
                 const double *z, double L, int maxneib, int *cnt, int *list,
                 int *cell_out) {
    int ncell = c->ncell;
    int *incnt = calloc((size_t)ncell, sizeof(int));
    int *incell = malloc(sizeof(int) * (size_t)ncell * OR_MAX_IN_CELL);
    int *map = malloc(sizeof(int) * 13 * (size_t)ncell);
    int rc = 0;
    or_maps(c, map);
    for (int i = 0; i < n; i++) cnt[i] = 0;
    /* binning, cell_list+forces.c:71-80 (particle index order) */
    for (int i = 0; i < n; i++) {
        int ci = or_cell_of(c, x[i], y[i], z[i]);
        if (cell_out) cell_out[i] = ci;
        if (ci < 0 || ci >= ncell || incnt[ci] >= OR_MAX_IN_CELL) { rc = -1; goto done; }
        incell[ci * OR_MAX_IN_CELL + incnt[ci]++] = i;
    }
    /* own-cell pairs, cell_list+forces.c:82-105 (CellInBlk order = 0..NCell-1) */
    for (int ci = 0; ci < ncell; ci++) {
        int count = incnt[ci];
        for (int j = 0; j < count - 1; j++) {
            int p1 = incell[ci * OR_MAX_IN_CELL + j];
            for (int k = j + 1; k < count; k++) {
                int p2 = incell[ci * OR_MAX_IN_CELL + k];
                double dx = minimg_md(x[p1] - x[p2]);
                double dy = minimg_md(y[p1] - y[p2]);
                double dz = minimg_md(z[p1] - z[p2]);
                if (dr2_md(dx, dy, dz, L) < OR_CELL_CUT_SQ) {
                    if (cnt[p1] >= maxneib) { rc = -1; goto done; }
                    list[(size_t)p1 * maxneib + cnt[p1]++] = p2;
                }
            }
        }
    }
    /* half-stencil pairs, cell_list+forces.c:107-140 */
    for (int ci = 0; ci < ncell; ci++) {
        if (incnt[ci] <= 0) continue;
        for (int s = 0; s < 13; s++) {
            int c2 = map[ci * 13 + s];
            if (incnt[c2] <= 0) continue;
            for (int i = 0; i < incnt[ci]; i++)
                for (int j = 0; j < incnt[c2]; j++) {
                    int p1 = incell[ci * OR_MAX_IN_CELL + i];
                    int p2 = incell[c2 * OR_MAX_IN_CELL + j];
                    double dx = minimg_md(x[p1] - x[p2]);
                    double dy = minimg_md(y[p1] - y[p2]);
                    double dz = minimg_md(z[p1] - z[p2]);
                    if (dr2_md(dx, dy, dz, L) < OR_CELL_CUT_SQ) {
                        if (cnt[p1] >= maxneib) { rc = -1; goto done; }
                        list[(size_t)p1 * maxneib + cnt[p1]++] = p2;
                    }
                }
        }
    }
done:
    free(incnt);
    free(incell);
    free(map);
    return rc;
}

void or_full_counts_brute(int n, const double *x, const double *y,
                          const double *z, double L, int *cnt) {
    for (int i = 0; i < n; i++) cnt[i] = 0;
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++) {
            double dx = minimg_md(x[i] - x[j]);
            double dy = minimg_md(y[i] - y[j]);
            double dz = minimg_md(z[i] - z[j]);
            if (dr2_md(dx, dy, dz, L) < OR_CELL_CUT_SQ) { cnt[i]++; cnt[j]++; }
        }
}

/* One pair of calculate_forces (cell_list+forces.c:174-203), factor = 1. */
static inline void pair_force(int i, int j, const double *x, const double *y,
                              const double *z, const int *type, double L,
                              double *fx, double *fy, double *fz, double *pe,
                              double *vir) {
    const double factor = 1.00;
    int A = type[i], B = type[j];
    double dx = minimg_md(x[i] - x[j]);
    double dy = minimg_md(y[i] - y[j]);
    double dz = minimg_md(z[i] - z[j]);
    double dr2 = dr2_md(dx, dy, dz, L);
    double eps = or_eps[A][B];
    double sig = or_sigma[A][B];
    double r_over_sig_2 = dr2 / (sig * sig);
    if (r_over_sig_2 < OR_LCUTSQR) {
        double s2 = 1.0 / r_over_sig_2;
        double s6 = s2 * s2 * s2;
        *pe += factor * 4.0 * eps * ((s6 * (s6 - 1.0)) + OR_C0 + OR_C2 * r_over_sig_2);
        double grad = 8.0 * eps * (6.0 * s6 * (s6 - 0.5) / dr2 - OR_C2 / (sig * sig));
        fx[i] += grad * dx * L;
        fy[i] += grad * dy * L;
        fz[i] += grad * dz * L;
        fx[j] -= grad * dx * L;
        fy[j] -= grad * dy * L;
        fz[j] -= grad * dz * L;
        *vir += factor * grad * dr2;
    }
}

void or_forces_list(int n, const double *x, const double *y, const double *z,
                    const double *xu, const double *yu, const double *zu,
                    const int *type, double L, const int *cnt, const int *list,
                    int maxneib, const int *act_tag, const int *kx,
                    const int *ky, const int *kz, double f0, double *fx,
                    double *fy, double *fz, double *out3) {
    double pe = 0.0, vir = 0.0, vir_act = 0.0;
    for (int i = 0; i < n; i++) { fx[i] = 0.0e0; fy[i] = 0.0e0; fz[i] = 0.0e0; }
    for (int i = 0; i < n; i++) /* InExtdBlk order = index order */
        for (int k = 0; k < cnt[i]; k++)
            pair_force(i, list[(size_t)i * maxneib + k], x, y, z, type, L, fx, fy, fz, &pe, &vir);
    /* active force, cell_list+forces.c:207-218 */
    for (int i = 0; i < n; i++) {
        if (act_tag[i] != -1) {
            int a = act_tag[i];
            fx[i] += f0 * (double)kx[a];
            fy[i] += f0 * (double)ky[a];
            fz[i] += f0 * (double)kz[a];
            if (xu)
                vir_act += f0 * ((double)kx[a] * xu[i] + (double)ky[a] * yu[i] +
                                 (double)kz[a] * zu[i]);
        }
    }
    vir = vir / (3.0e0); /* :220 */
    out3[0] = pe;
    out3[1] = vir;
    out3[2] = vir_act;
}

void or_forces_brute(int n, const double *x, const double *y, const double *z,
                     const int *type, double L, const int *act_tag,
                     const int *kx, const int *ky, const int *kz, double f0,
                     double *fx, double *fy, double *fz, double *out2) {
    double pe = 0.0, vir = 0.0;
    for (int i = 0; i < n; i++) { fx[i] = 0.0; fy[i] = 0.0; fz[i] = 0.0; }
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++)
            pair_force(i, j, x, y, z, type, L, fx, fy, fz, &pe, &vir);
    for (int i = 0; i < n; i++)
        if (act_tag && act_tag[i] != -1) {
            int a = act_tag[i];
            fx[i] += f0 * (double)kx[a];
            fy[i] += f0 * (double)ky[a];
            fz[i] += f0 * (double)kz[a];
        }
    out2[0] = pe;
    out2[1] = vir / (3.0e0);
}

/* Forces on m chosen particles idx[] by an O(N) pass over every partner
 * j != i in index order, with the arithmetic of pair_force (the
 * contribution to i only) plus the active force: a full-size check that
 * needs no neighbour list.  f3[3*k..3*k+2] = force on particle idx[k]. */
void or_forces_sample(int n, const double *x, const double *y, const double *z, const int *type, double L,
                      const int *act_tag, const int *kx, const int *ky, const int *kz, double f0, int m,
                      const int *idx, double *f3) {
    for (int k = 0; k < m; k++) {
        const int i = idx[k], A = type[i];
        double fx = 0.0, fy = 0.0, fz = 0.0;
        for (int j = 0; j < n; j++) {
            if (j == i) continue;
            const int B = type[j];
            double dx = minimg_md(x[i] - x[j]);
            double dy = minimg_md(y[i] - y[j]);
            double dz = minimg_md(z[i] - z[j]);
            double dr2 = dr2_md(dx, dy, dz, L);
            double eps = or_eps[A][B];
            double sig = or_sigma[A][B];
            double r_over_sig_2 = dr2 / (sig * sig);
            if (r_over_sig_2 < OR_LCUTSQR) {
                double s2 = 1.0 / r_over_sig_2;
                double s6 = s2 * s2 * s2;
                double grad = 8.0 * eps * (6.0 * s6 * (s6 - 0.5) / dr2 - OR_C2 / (sig * sig));
                fx += grad * dx * L;
                fy += grad * dy * L;
                fz += grad * dz * L;
            }
        }
        if (act_tag && act_tag[i] != -1) {
            const int a = act_tag[i];
            fx += f0 * (double)kx[a];
            fy += f0 * (double)ky[a];
            fz += f0 * (double)kz[a];
        }
        f3[3 * k] = fx;
        f3[3 * k + 1] = fy;
        f3[3 * k + 2] = fz;
    }
}

void or_nhc_masses(int n, double T, double *q3) {
    double tau = .20; /* time_evolve.c:477-482 */
    for (int i = 0; i < OR_M; i++) q3[i] = T * tau * tau;
    q3[0] = dof_of(n) * T * tau * tau;
}

/* ---- NHC propagators, MDcode/time_evolve.c:527-609 (one process) ---- */
typedef struct {
    int n;
    double L, inv_L, T, dof;
    double *x, *y, *z, *xu, *yu, *zu, *px, *py, *pz, *fx, *fy, *fz;
    double *eta, *peta, *q;
    double *xold, *yold, *zold;
    double ke, maxmaxD;
} nhc_t;

static void applypbc1(nhc_t *s) { /* time_evolve.c:15-26 */
    for (int i = 0; i < s->n; i++) {
        while (s->x[i] >= 1.0) s->x[i] -= 1.0;
        while (s->x[i] < 0.0) s->x[i] += 1.0;
        while (s->y[i] >= 1.0) s->y[i] -= 1.0;
        while (s->y[i] < 0.0) s->y[i] += 1.0;
        while (s->z[i] >= 1.0) s->z[i] -= 1.0;
        while (s->z[i] < 0.0) s->z[i] += 1.0;
    }
}

static void U1(nhc_t *s) { /* time_evolve.c:527-557 */
    for (int i = 0; i < s->n; i++) {
        double dx = s->px[i] * OR_DT, dy = s->py[i] * OR_DT, dz = s->pz[i] * OR_DT;
        s->x[i] += dx * s->inv_L;
        s->xu[i] += dx * s->inv_L;
        s->y[i] += dy * s->inv_L;
        s->yu[i] += dy * s->inv_L;
        s->z[i] += dz * s->inv_L;
        s->zu[i] += dz * s->inv_L;
    }
    applypbc1(s);
    double maxD = 0.0;
    for (int i = 0; i < s->n; i++) {
        double dx = s->xu[i] - s->xold[i];
        double dy = s->yu[i] - s->yold[i];
        double dz = s->zu[i] - s->zold[i];
        double temp = dx * dx + dy * dy + dz * dz;
        if (maxD < temp) maxD = temp;
    }
    maxD *= s->L * s->L;
    s->maxmaxD = maxD;
}

static void U2(nhc_t *s) { /* time_evolve.c:558-567 */
    for (int i = 0; i < s->n; i++) {
        s->px[i] += s->fx[i] * OR_HALF_DT;
        s->py[i] += s->fy[i] * OR_HALF_DT;
        s->pz[i] += s->fz[i] * OR_HALF_DT;
    }
}

static void U3(nhc_t *s, double t) { /* time_evolve.c:568-577 */
    for (int i = 0; i < s->n; i++) {
        s->px[i] = s->px[i] * exp(-t * s->peta[0] / s->q[0]);
        s->py[i] = s->py[i] * exp(-t * s->peta[0] / s->q[0]);
        s->pz[i] = s->pz[i] * exp(-t * s->peta[0] / s->q[0]);
    }
    for (int i = 0; i < OR_M; i++) s->eta[i] = s->eta[i] + t * s->peta[i] / s->q[i];
}

static double sum_p2(const nhc_t *s) {
    double keblk = 0;
    for (int i = 0; i < s->n; i++)
        keblk += s->px[i] * s->px[i] + s->py[i] * s->py[i] + s->pz[i] * s->pz[i];
    return keblk;
}

static void U4(nhc_t *s, double t, int jStart, int jEnd) { /* time_evolve.c:578-609 */
    int dj = 1;
    double gj = 0.0;
    if (jStart > jEnd) dj = -1;
    int j = jStart;
    do {
        if (j == 0) {
            gj = sum_p2(s);
            s->ke = gj;
            gj -= s->dof * s->T;
        } else {
            gj = (s->peta[j - 1] * s->peta[j - 1]) / s->q[j - 1] - s->T;
        }
        if (j == OR_M - 1) {
            s->peta[j] = s->peta[j] + t * gj;
        } else {
            double xx = t * s->peta[j + 1] / s->q[j + 1];
            s->peta[j] = s->peta[j] * exp(-xx) + t * gj * (1.0 - exp(-xx)) / xx;
        }
        j = j + dj;
    } while (j - dj != jEnd);
}

int or_nhc_steps(int n, double L, double T, double f0, int nsteps, int fresh,
                 int safe, int nve, double *x, double *y, double *z,
                 double *xu, double *yu, double *zu, double *px, double *py,
                 double *pz, double *fx, double *fy, double *fz,
                 const int *type, const int *act_tag, const int *kx,
                 const int *ky, const int *kz, double *chain, double *xold,
                 double *yold, double *zold, double *out8) {
    const int maxneib = 500; /* global.h:19 */
    nhc_t s = {n, L, 1.0 / L, T, dof_of(n), x, y, z, xu, yu, zu, px, py, pz,
               fx, fy, fz, chain, chain + 3, chain + 6, xold, yold, zold, 0, 0};
    or_cells c;
    or_cells_setup(L, &c);
    int *cnt = malloc(sizeof(int) * (size_t)n);
    int *list = malloc(sizeof(int) * (size_t)n * maxneib);
    double f3[3] = {0, 0, 0};
    int rebuilds = 0, rc = 0;
    /* the list always exists in the reference when the step loop starts;
     * here it is rebuilt from the current positions (xold is overwritten). */
    if (or_half_list(&c, n, x, y, z, L, maxneib, cnt, list, NULL)) { rc = -1; goto out; }
    if (fresh) {
        for (int i = 0; i < n; i++) { xold[i] = xu[i]; yold[i] = yu[i]; zold[i] = zu[i]; }
        or_forces_list(n, x, y, z, xu, yu, zu, type, L, cnt, list, maxneib,
                       act_tag, kx, ky, kz, f0, fx, fy, fz, f3);
    }
    const double skin2 = OR_SKIN * OR_SKIN;
    for (int st = 0; st < nsteps; st++) {
        if (!nve) {
            U4(&s, OR_DT * 0.25, OR_M - 1, 0);
            U3(&s, OR_DT * 0.50);
            U4(&s, OR_DT * 0.25, 0, OR_M - 1);
        }
        U2(&s);
        U1(&s);
        int rebuild = safe ? (4.0 * s.maxmaxD > skin2) : (3.0 * s.maxmaxD > skin2);
        if (rebuild) {
            if (or_half_list(&c, n, x, y, z, L, maxneib, cnt, list, NULL)) { rc = -1; goto out; }
            for (int i = 0; i < n; i++) { xold[i] = xu[i]; yold[i] = yu[i]; zold[i] = zu[i]; }
            rebuilds++;
        }
        or_forces_list(n, x, y, z, xu, yu, zu, type, L, cnt, list, maxneib,
                       act_tag, kx, ky, kz, f0, fx, fy, fz, f3);
        U2(&s);
        if (!nve) {
            U4(&s, OR_DT * 0.25, OR_M - 1, 0);
            U3(&s, OR_DT * 0.50);
            U4(&s, OR_DT * 0.25, 0, OR_M - 1);
        }
    }
out:;
    /* print_stuff quantities, MDcode/print.c:63-87 */
    double ke = 0.0, drift = 0.0;
    for (int i = 0; i < n; i++) {
        drift += px[i];
        ke += (px[i] * px[i] + py[i] * py[i] + pz[i] * pz[i]);
    }
    ke = ke * 0.5;
    double inst_T = 2.0 * ke / s.dof;
    double heff = ke + f3[0];
    if (!nve) {
        heff = ke + f3[0] + s.eta[0] * T * s.dof;
        for (int i = 0; i < OR_M; i++) heff = heff + s.peta[i] * s.peta[i] / (2.0 * s.q[i]);
        for (int i = 1; i < OR_M; i++) heff = heff + T * s.eta[i];
    }
    out8[0] = f3[0];
    out8[1] = f3[1];
    out8[2] = ke;
    out8[3] = inst_T;
    out8[4] = heff;
    out8[5] = rebuilds;
    out8[6] = s.maxmaxD;
    out8[7] = drift;
    free(cnt);
    free(list);
    return rc;
}

void or_shuffle(int nact, int *k, const double *u) {
    /* MDcode/active.c:58-63 with uniform(0,1) replaced by the supplied u[i] */
    for (int i = 0; i < nact; i++) {
        int j = (int)((double)nact * u[i]);
        int temp = k[i];
        k[i] = k[j];
        k[j] = temp;
    }
}

/* MDcode/main.c:17-41: Numerical-Recipes shell sort of a[1..n] */
static void shell_sort(int n, unsigned long long *a) {
    int i, j, inc;
    unsigned long long v;
    inc = 1;
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

int or_schedule(unsigned long long run_length, int n_origins, double factor,
                unsigned long long tail, unsigned long long *out, int cap) {
    /* MDcode/main.c:62-105 */
    unsigned long long current = 0, k, offset, lin, maxint, index;
    lin = run_length / (unsigned long long)n_origins;
    maxint = run_length;
    for (k = 0; k < (unsigned long long)n_origins; k++) {
        if (current + 1 >= (unsigned long long)cap) return -1;
        out[current++] = k * lin;
        offset = 1;
        while (offset < maxint) {
            index = k * lin + offset;
            if (index < run_length) {
                if (current + 1 >= (unsigned long long)cap) return -1;
                out[current++] = index;
            }
            if ((unsigned long long)(offset * factor) == offset)
                offset++;
            else
                offset = (unsigned long long)((double)offset * factor);
        }
    }
    /* the words past the filled range that shellSort and the dedupe read */
    for (unsigned long long t = current; t < (unsigned long long)cap; t++) out[t] = tail;
    shell_sort((int)current, out);
    unsigned long long i = 0, j = 0;
    while (j < current) {
        while (j + 1 < (unsigned long long)cap && out[j] == out[i]) j++;
        i++;
        out[i] = out[j];
    }
    return (int)i;
}
