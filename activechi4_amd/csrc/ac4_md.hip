// ac4_md.hip -- active Kob-Andersen MD step for MI355X (gfx950).
//
// Replaces MDcode/cell_list+forces.c, MDcode/time_evolve.c:476-609 and
// MDcode/active.c:45-85 (paths relative to the reference checkout).  The
// whole box of every replica is resident in HBM (no domain decomposition,
// MDcode/Parallel.c); replicas are packed into every launch through a block
// table (block -> replica, first particle), so one launch sequence per step
// advances all of them.
//
// Per step (every launch covers all replicas, kernels early-exit per replica):
//   k_shuffle      ShuffleActDirec when step % tau_per == 0 (swap chain of
//                  active.c:58-75 per component, Philox uniforms); in the
//                  first workgroups of k_first_half where its LDS is small
//   k_first_half   U4 U3 U4 (chain start, computed redundantly per block),
//                  U2 kick, U1 drift + wrap + max displacement (atomicMax)
//   k_bin .. k_nlist_team   counting-sort cell build + neighbour lists,
//                  executed only for replicas whose displacement criterion
//                  fired (flag read on device, so the sequence is graphable)
//   k_force        pair forces + active force + second U2 kick + block
//                  partial sums
//   k_force_end    per replica: the partials summed in block order, U4 U3 U4
//                  (chain end).
// The momentum factor of the last U3 is kept pending (p_scale) and applied at
// the next first half, so a step needs one sum of p^2 instead of four.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#include <vector>
#include <algorithm>
#include <type_traits>
#include "ac4_internal.h"

#define BLK 256
#define LPB 16  // particles per list-build workgroup

// MDcode/global.h:24-45 (Sigma = 1.00)
#define K_LCUTSQR (6.25 * 1.00 * 1.00)
#define K_SKIN (.35 * 1.00)
#define K_CELL_CUT ((2.5 * 1.00) + (.35 * 1.00))
#define K_CELL_CUT_SQ (K_CELL_CUT * K_CELL_CUT)
#define K_DT 0.005
#define K_HALF_DT (K_DT * 0.5)
#define K_C0 0.0162665594880
#define K_C2 -0.001949973872640

struct RepC {
    int n, off, nact, off_act;
    int nc, ncell, off_cell, tau;      // list grid: nc = floor(L/2.85) cells per axis (27-cell stencil)
    int nblk, blk_off, nc_ref, tpp;    // nc_ref: the reference's grid (Parallel.c:72-80)
    int fnblk, fblk_off, ltpp, cap_n;  // force-kernel block table (tpp = 1<<ltpp lanes per particle)
    int cap_cells, lblk_off;           // lblk: list-build blocks of LPB particles
    int lnblk, lc;                     // lc: log2 of the list chunk (below)
    double L, invL, inv_cell_ref, T, f0, dof, density;
    uint32_t key0, key1;
    double qscale;  // L / 2^32: real length of one unit of a fixed-point coordinate
    double skin;    // Verlet skin: list radii 2.5 sigma_ab + skin, rebuild criterion
    float qs2;      // qscale^2 in FP32
    int xcull;      // list teams skip the far x-half of the cells beside along x (configure_rep)
};

// Fixed-point copy of a wrapped position for the FP32 force path and the
// dynamics lists: each box fraction s in [0,1) as the 32-bit integer
// floor(s * 2^32) (quantum L/2^32, below 1e-8 sigma for L < 40), with the
// species (0 or 1) in the lowest bit of x.  The 32-bit difference of two such
// coordinates, read as a signed integer, is the minimum-image separation
// (cell_list+forces.c:92-97 up to the tie at exactly half a box), exact in y
// and z and within one quantum in x.  A neighbour costs one 12-byte load: 64
// lanes gathering 12-byte records touch fewer 128-byte lines than with
// 16-byte ones, and the gathers are what bound the force kernel.
struct QPos {
    unsigned int x, y, z;
};
__host__ __device__ __forceinline__ QPos qpos(double sx, double sy, double sz, int t) {
    QPos q;
    q.x = ((unsigned int)(unsigned long long)(sx * 4294967296.0) & ~1u) | (unsigned int)(t & 1);
    q.y = (unsigned int)(unsigned long long)(sy * 4294967296.0);
    q.z = (unsigned int)(unsigned long long)(sz * 4294967296.0);
    return q;
}
__host__ __device__ __forceinline__ int ptype(const QPos &q) { return (int)(q.x & 1u); }
__host__ __device__ __forceinline__ int ptype(const double4 &q) { return (int)q.w; }

struct ChainS {
    double eta[3], peta[3], q[3];
    double S_end, p_scale;
    double pe, vir, ke, inst_T, heff, drift;
    unsigned long long maxbits;
    long long step, shuffles;
    int rebuilds, force_rebuild, err, crit;
    unsigned int done_blocks, bin_done;  // no longer counted (kept for the checkpoint layout)
    int mode;                 // 0: Nose-Hoover chain, 1: velocity-rescale ramp, 2: NVE
    int ramp_n, ramp_k;
    double ramp0, ramp1;      // target T = T*(ramp0 + (ramp1-ramp0)*k/(n-1))
};

// MDcode/cell_list+forces.c:19-20 restricted to the two KA species.
__device__ __forceinline__ double sig_ab(int a, int b) { return (a + b) == 0 ? 1.00 : ((a + b) == 1 ? 0.80 : 0.88); }
__device__ __forceinline__ double eps_ab(int a, int b) { return (a + b) == 0 ? 1.0 : ((a + b) == 1 ? 1.5 : 0.50); }

__device__ __forceinline__ double minimg_md(double d) {  // cell_list+forces.c:92-97
    if (d < -0.5) d += 1.0;
    if (d > 0.5) d -= 1.0;
    return d;
}

// U4 propagator (MDcode/time_evolve.c:578-609) with the p^2 sum supplied.
__host__ __device__ inline void nhc_u4(double t, int jStart, int jEnd, double S, double T, double dof,
                                       const double *q, double *peta) {
    int dj = 1;
    double gj = 0.0;
    if (jStart > jEnd) dj = -1;
    int j = jStart;
    do {
        if (j == 0) {
            gj = S;
            gj -= dof * T;
        } else {
            gj = (peta[j - 1] * peta[j - 1]) / q[j - 1] - T;
        }
        if (j == 2) {
            peta[j] = peta[j] + t * gj;
        } else {
            double x = t * peta[j + 1] / q[j + 1];
            peta[j] = peta[j] * exp(-x) + t * gj * (1.0 - exp(-x)) / x;
        }
        j = j + dj;
    } while (j - dj != jEnd);
}

// U4(dt/4,2->0) U3(dt/2) U4(dt/4,0->2) on a copy of the chain (time_evolve.c:499-501
// and :517-519); S = sum p^2 entering; returns the U3 momentum factor.
__host__ __device__ inline double nhc_half(double S, double T, double dof, const double *q, double *eta,
                                           double *peta, double *S_out) {
    nhc_u4(K_DT * 0.25, 2, 0, S, T, dof, q, peta);
    double t = K_DT * 0.50;
    double s = exp(-t * peta[0] / q[0]);                        // time_evolve.c:571
    for (int i = 0; i < 3; i++) eta[i] = eta[i] + t * peta[i] / q[i];  // :575
    double S2 = S * s * s;
    nhc_u4(K_DT * 0.25, 0, 2, S2, T, dof, q, peta);
    *S_out = S2;
    return s;
}

__device__ __forceinline__ bool need_rebuild(const RepC &R, const ChainS &c) {
    if (c.force_rebuild) return true;
    double maxD = __longlong_as_double((long long)c.maxbits);
    maxD *= R.L * R.L;  // time_evolve.c:553
    return (c.crit == 1 ? 3.0 : 4.0) * maxD > R.skin * R.skin;  // :508 (crit 1, skin 0.35 there)
}

// Rebuild work of one step: the blocks of the replicas due for a rebuild,
// appended by k_bin_end for each such replica, so the kernels after
// k_bin run a fixed grid over this list instead of one workgroup per block of
// every replica (most of which would have nothing to do).  n[0]: particle
// blocks of BLK (global block ids); n[1]: list blocks of LPB particles
// (global ids) at item + off_l; n[2]: cell blocks (replica << 16 | block of
// BLK/64 cells) at item + off_c.  The counters are cleared by k_first_half
// (and k_cell_zero for a forced rebuild).
struct WorkList {
    int *item, *n;
    int off_l, off_c;
};

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup w runs on
// XCD w % 8).  Kernels that gather neighbour positions launch 8*ceil(nb/8)
// workgroups and map workgroup w to logical block (w % 8)*ceil(nb/8) + w/8, so
// every XCD works on a contiguous range of blocks (a few whole replicas) and
// finds their positions in its own L2.  Returns -1 for the padding blocks.
#define N_XCD 8
__device__ __forceinline__ int xcd_block(int nb) {
    const int per = (nb + N_XCD - 1) / N_XCD;
    const int lb = (int)(blockIdx.x % N_XCD) * per + (int)(blockIdx.x / N_XCD);
    return lb < nb ? lb : -1;
}
static inline int xcd_grid(int nb) { return ((nb + N_XCD - 1) / N_XCD) * N_XCD; }

// ---------------------------------------------------------------- kernels
// Swap shuffle of one direction component of one replica per workgroup
// (grid: replicas x 3).  The swap partners j_i = (int)(Nact*U_i) do not depend
// on the array, so the 256 lanes draw them SHUF_CH at a time into LDS and one
// lane applies the swaps in order.  The component lives in dynamic LDS when
// it fits (kcap bytes, sized from the device's LDS per workgroup), else in
// global memory.  The shuffle counter is advanced by k_force_end (or
// the host for an explicit shuffle), after every component has read it.
#define SHUF_CH 4096
#define SHUF_PAR_MIN 16384  // active sets above this take the parallel path (k_shuf_part, k_shuf_apply)
#define SHUF_TRACE_MAX 1024 // active sets up to this are shuffled by tracing every position back

// Lane t traces positions t, t+256, ... (C of them, independent chains side
// by side) back through the swaps na-1 ... 0; the partners come from LDS four
// at a time, the next four loaded before the current four are applied.
template <int C>
__device__ __forceinline__ void shuffle_trace(const unsigned int *s_j, const int8_t *s_k, int8_t *kg, int na) {
    int y[C];
#pragma unroll
    for (int c = 0; c < C; c++) y[c] = threadIdx.x + 256 * c;
    int q = na - 1;
    for (; q >= 0 && ((q + 1) & 3); q--) {  // down to a multiple of four
        const int jq = (int)s_j[q];
#pragma unroll
        for (int c = 0; c < C; c++) y[c] = y[c] == q ? jq : (y[c] == jq ? q : y[c]);
    }
    if (q >= 3) {
        uint4 nx = *(const uint4 *)(s_j + q - 3);
        for (; q >= 3; q -= 4) {
            const uint4 cur = nx;
            if (q >= 7) nx = *(const uint4 *)(s_j + q - 7);
            const int j4[4] = {(int)cur.w, (int)cur.z, (int)cur.y, (int)cur.x};  // partners of q, q-1, q-2, q-3
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int c = 0; c < C; c++) y[c] = y[c] == q - u ? j4[u] : (y[c] == j4[u] ? q - u : y[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < C; c++) {
        const int x = threadIdx.x + 256 * c;
        if (x < na) kg[x] = s_k[y[c]];
    }
}
// One component of one replica's shuffle by a workgroup of 256 threads (the
// branch is uniform over the workgroup); s_shuf is the dynamic LDS: SHUF_CH
// partners, then kcap bytes of directions.
__device__ __forceinline__ void shuffle_comp(const RepC &R, const ChainS &c, int comp, int8_t *kx, int8_t *ky,
                                             int8_t *kz, int force, int kcap, int skip_large, unsigned int *s_shuf) {
    unsigned int *s_j = s_shuf;
    int8_t *s_k = (int8_t *)(s_shuf + SHUF_CH);
    if (!force && (c.mode == 1 || R.tau <= 0 || (c.step % R.tau) != 0)) return;
    const uint64_t sh = (uint64_t)c.shuffles;
    const int na = R.nact;
    if (skip_large && na > SHUF_PAR_MIN) return;
    int8_t *kg = (comp == 0 ? kx : (comp == 1 ? ky : kz)) + R.off_act;
    const bool lds = na <= kcap;
    int8_t *k = lds ? s_k : kg;
    if (lds)
        for (int i = threadIdx.x; i < na; i += 256) s_k[i] = kg[i];
    if (lds && na <= SHUF_TRACE_MAX) {
        // Every position at once: the swaps s = 0..na-1 in order give the
        // value at x the one that started at the position found by tracing
        // x back through them (swap s maps s <-> j_s), so the new array is a
        // gather of the old one; exactly the serial loop's result, in about
        // na^2/64 lane operations instead of na dependent LDS round trips.
        for (int q = threadIdx.x; q < na; q += 256)
            s_j[q] = (unsigned int)(int)((double)na * ac4_shuffle_uniform(R.key0, R.key1, sh, comp, q));
        __syncthreads();
        const int nch = (na + 255) / 256;  // positions per lane, traced side by side
        if (nch == 1) shuffle_trace<1>(s_j, s_k, kg, na);
        else if (nch == 2) shuffle_trace<2>(s_j, s_k, kg, na);
        else if (nch == 3) shuffle_trace<3>(s_j, s_k, kg, na);
        else shuffle_trace<4>(s_j, s_k, kg, na);
        return;
    }
    for (int c0 = 0; c0 < na; c0 += SHUF_CH) {
        const int m = min(SHUF_CH, na - c0);
        __syncthreads();
        for (int q = threadIdx.x; q < m; q += 256)
            s_j[q] = (unsigned int)(int)((double)na * ac4_shuffle_uniform(R.key0, R.key1, sh, comp, c0 + q));
        __syncthreads();
        if (threadIdx.x == 0)  // MDcode/active.c:58-63: for i: j = (int)(Nact*U[0,1)); swap(k[i], k[j])
            for (int q = 0; q < m; q++) {
                const int i = c0 + q, jj = (int)s_j[q];
                const int8_t a = k[i], b = k[jj];
                k[i] = b;
                k[jj] = a;
            }
    }
    __syncthreads();
    if (lds)
        for (int i = threadIdx.x; i < na; i += 256) kg[i] = s_k[i];
}
__global__ void __launch_bounds__(256) k_shuffle(const RepC *__restrict__ reps, const ChainS *__restrict__ chain,
                                                 int8_t *kx, int8_t *ky, int8_t *kz, int force, int kcap, int skip_large) {
    extern __shared__ unsigned int s_shuf[];
    const int r = blockIdx.x, comp = blockIdx.y;
    const RepC R = reps[r];
    shuffle_comp(R, chain[r], comp, kx, ky, kz, force, kcap, skip_large, s_shuf);
}

// The same swap sequence in parallel for large active sets.  The swaps
// i = b*SHUF_M ... b*SHUF_M + SHUF_M - 1 of block b move values only among the
// positions they touch: the block's own SHUF_M positions and the partners j_i
// outside them.  k_shuf_part finds, for block b, where the value at each
// touched position comes from (a position of the state before the block),
// one wave per block replaying the swaps on that map (lane 0 swaps, the wave
// searches the list of outside partners); k_shuf_apply then applies the
// blocks in order, each as one gather over its touched positions.  The
// result is the serial chain's, value for value.
#define SHUF_M 256
__global__ void __launch_bounds__(64) k_shuf_part(const RepC *__restrict__ reps, const ChainS *__restrict__ chain,
                                                  int force, int bmax, int2 *__restrict__ ent, int *__restrict__ ent_n) {
    __shared__ int s_own[SHUF_M];            // source of own position b0 + q
    __shared__ int s_fp[SHUF_M], s_fs[SHUF_M];  // outside positions and their sources
    __shared__ unsigned int s_j[SHUF_M];
    __shared__ int s_nf;
    const int blk = blockIdx.x, comp = blockIdx.y, r = blockIdx.z;
    const RepC R = reps[r];
    const ChainS &c = chain[r];
    if (!force && (c.mode == 1 || R.tau <= 0 || (c.step % R.tau) != 0)) return;
    const int na = R.nact, b0 = blk * SHUF_M;
    if (na <= SHUF_PAR_MIN || b0 >= na) return;
    const int m = min(SHUF_M, na - b0), lane = threadIdx.x;
    const uint64_t sh = (uint64_t)c.shuffles;
    for (int q = lane; q < m; q += 64) {
        s_own[q] = b0 + q;
        s_j[q] = (unsigned int)(int)((double)na * ac4_shuffle_uniform(R.key0, R.key1, sh, comp, b0 + q));
    }
    if (lane == 0) s_nf = 0;
    __syncthreads();
    for (int q = 0; q < m; q++) {  // MDcode/active.c:58-63, swap(k[i], k[j]) on the source map
        const int j = (int)s_j[q];
        int f = -1;  // outside slot of j
        if (j < b0 || j >= b0 + m) {
            const int nf = s_nf;
            for (int k0 = 0; k0 < nf; k0 += 64) {
                const unsigned long long hit = __ballot(k0 + lane < nf && s_fp[k0 + lane] == j);
                if (hit) {
                    f = k0 + __ffsll((long long)hit) - 1;
                    break;
                }
            }
            if (f < 0) {  // first touch of j: its value is its own
                f = nf;
                if (lane == 0) { s_fp[nf] = j; s_fs[nf] = j; s_nf = nf + 1; }
            }
        }
        if (lane == 0) {
            const int a = s_own[q];
            if (f < 0) { s_own[q] = s_own[j - b0]; s_own[j - b0] = a; }
            else { s_own[q] = s_fs[f]; s_fs[f] = a; }
        }
        __syncthreads();
    }
    const size_t e0 = ((size_t)(r * 3 + comp) * bmax + blk) * (2 * SHUF_M);
    const int nf = s_nf;
    for (int q = lane; q < m; q += 64) ent[e0 + q] = make_int2(b0 + q, s_own[q]);
    for (int q = lane; q < nf; q += 64) ent[e0 + m + q] = make_int2(s_fp[q], s_fs[q]);
    if (lane == 0) ent_n[(size_t)(r * 3 + comp) * bmax + blk] = m + nf;
}

__global__ void __launch_bounds__(2 * SHUF_M) k_shuf_apply(const RepC *__restrict__ reps,
                                                         const ChainS *__restrict__ chain, int8_t *kx, int8_t *ky,
                                                         int8_t *kz, int force, int kcap, int bmax,
                                                         const int2 *__restrict__ ent, const int *__restrict__ ent_n) {
    extern __shared__ unsigned int s_shuf[];  // same layout as k_shuffle: directions after SHUF_CH words
    int8_t *s_k = (int8_t *)(s_shuf + SHUF_CH);
    const int r = blockIdx.x, comp = blockIdx.y;
    const RepC R = reps[r];
    const ChainS &c = chain[r];
    if (!force && (c.mode == 1 || R.tau <= 0 || (c.step % R.tau) != 0)) return;
    const int na = R.nact;
    if (na <= SHUF_PAR_MIN) return;
    int8_t *kg = (comp == 0 ? kx : (comp == 1 ? ky : kz)) + R.off_act;
    const bool lds = na <= kcap;
    int8_t *k = lds ? s_k : kg;
    if (lds)
        for (int i = threadIdx.x; i < na; i += 2 * SHUF_M) s_k[i] = kg[i];
    __syncthreads();
    const int nb = (na + SHUF_M - 1) / SHUF_M;
    for (int blk = 0; blk < nb; blk++) {  // blocks in swap order
        const size_t cb = (size_t)(r * 3 + comp) * bmax + blk;
        const int ne = ent_n[cb];
        int2 e = make_int2(0, 0);
        int8_t v = 0;
        if ((int)threadIdx.x < ne) {
            e = ent[cb * (2 * SHUF_M) + threadIdx.x];
            v = k[e.y];
        }
        __syncthreads();
        if ((int)threadIdx.x < ne) k[e.x] = v;
        __syncthreads();
    }
    if (lds)
        for (int i = threadIdx.x; i < na; i += 2 * SHUF_M) kg[i] = s_k[i];
}

// The step's shuffle as the first 3*R workgroups of k_first_half (nblk = 3*R;
// 0 when it runs as k_shuffle before it): the shuffle only has to finish
// before the force kernel reads the directions, so its workgroups run beside
// the half kick instead of as a launch of their own in the step's chain.
struct ShufArgs {
    int8_t *kx, *ky, *kz;
    int kcap, nblk;
};

__global__ void __launch_bounds__(BLK) k_first_half(const RepC *__restrict__ reps, const ChainS *chain,
                                                    const int *__restrict__ blk_rep, const int *__restrict__ blk_first,
                                                    double *x, double *y, double *z, double *xu, double *yu, double *zu,
                                                    double *px, double *py, double *pz, const double *__restrict__ fx,
                                                    const double *__restrict__ fy, const double *__restrict__ fz,
                                                    const double *__restrict__ xo, const double *__restrict__ yo,
                                                    const double *__restrict__ zo, const int *__restrict__ type,
                                                    double4 *pos4, QPos *posq, int q32, ChainS *chain_w,
                                                    const int *__restrict__ blk_idx, int *cnt, int *wl_n,
                                                    ShufArgs sa) {
    __shared__ double s_scale[2];
    __shared__ double s_max[BLK / 64];
    if ((int)blockIdx.x < sa.nblk) {  // the step's shuffle: replica b / 3, component b % 3
        extern __shared__ unsigned int s_shuf[];
        const int rs = blockIdx.x / 3;
        const RepC Rs = reps[rs];
        shuffle_comp(Rs, chain[rs], blockIdx.x % 3, sa.kx, sa.ky, sa.kz, 0, sa.kcap, 0, s_shuf);
        return;
    }
    const int b = blockIdx.x - sa.nblk, r = blk_rep[b];
    if (b == 0 && threadIdx.x < 3) wl_n[threadIdx.x] = 0;  // rebuild work of this step
    const RepC R = reps[r];
    if (threadIdx.x == 0) {
        const ChainS &c = chain[r];
        double eta[3] = {c.eta[0], c.eta[1], c.eta[2]}, peta[3] = {c.peta[0], c.peta[1], c.peta[2]}, S2;
        s_scale[0] = c.p_scale;
        s_scale[1] = (c.mode == 0 && R.n > 0) ? nhc_half(c.S_end, R.T, R.dof, c.q, eta, peta, &S2) : 1.0;
    }
    __syncthreads();
    // clear this replica's cell counts for a possible rebuild later in the step
    for (int c = blk_idx[b] * BLK + threadIdx.x; c < R.ncell; c += R.nblk * BLK) cnt[R.off_cell + c] = 0;
    const double ps = s_scale[0], s1 = s_scale[1];
    const int i = blk_first[b] + threadIdx.x;
    double d2 = 0.0;
    if (i < R.n) {
        const int g = R.off + i;
        double vx = px[g], vy = py[g], vz = pz[g];
        vx = vx * ps; vy = vy * ps; vz = vz * ps;   // end-of-step U3 of the previous step
        vx = vx * s1; vy = vy * s1; vz = vz * s1;   // U3, time_evolve.c:571-573
        vx += fx[g] * K_HALF_DT;                    // U2, :562-564
        vy += fy[g] * K_HALF_DT;
        vz += fz[g] * K_HALF_DT;
        px[g] = vx; py[g] = vy; pz[g] = vz;
        double dx = vx * K_DT, dy = vy * K_DT, dz = vz * K_DT;   // U1, :532-540
        double sx = x[g] + dx * R.invL, sy = y[g] + dy * R.invL, sz = z[g] + dz * R.invL;
        double ux = xu[g] + dx * R.invL, uy = yu[g] + dy * R.invL, uz = zu[g] + dz * R.invL;
        // applypbc1 (:15-26) loops until the coordinate is in [0,1); a
        // coordinate that is not finite or many boxes away means the run blew
        // up: flag it instead of looping
        if (!(fabs(sx) < 8.0 && fabs(sy) < 8.0 && fabs(sz) < 8.0)) {
            atomicOr(&chain_w[r].err, 2);
            sx = 0.0; sy = 0.0; sz = 0.0;
        }
        while (sx >= 1.0) sx -= 1.0;
        while (sx < 0.0) sx += 1.0;
        while (sy >= 1.0) sy -= 1.0;
        while (sy < 0.0) sy += 1.0;
        while (sz >= 1.0) sz -= 1.0;
        while (sz < 0.0) sz += 1.0;
        x[g] = sx; y[g] = sy; z[g] = sz;
        xu[g] = ux; yu[g] = uy; zu[g] = uz;
        if (q32)
            posq[g] = qpos(sx, sy, sz, type[g]);
        else
            pos4[g] = make_double4(sx, sy, sz, (double)type[g]);
        double ex = ux - xo[g], ey = uy - yo[g], ez = uz - zo[g];  // :547-551
        d2 = ex * ex + ey * ey + ez * ez;
    }
    d2 = ac4_wave_max(d2);
    if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = d2;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = s_max[0];
        for (int w = 1; w < BLK / 64; w++) m = fmax(m, s_max[w]);
        // The maximum only grows during this kernel (it was cleared by the
        // previous step's force epilogue), so a block whose maximum is not
        // above the value it reads needs no atomic: same-address atomics are
        // serialised across the device, one per block of a large replica.
        const unsigned long long mb = (unsigned long long)__double_as_longlong(m);
        if (mb > __hip_atomic_load(&chain_w[r].maxbits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(&chain_w[r].maxbits, mb);
    }
}

// ---- rebuild: counting sort into cells, canonical order, permutation of the
// particle arrays into cell order, neighbour lists from LDS-staged cells.
// Binning uses the finest grid with cells >= 2.85 (list cutoff); the
// reference's own grid (MDcode/Parallel.c:72-80) is evaluated beside it for
// the cell ids update_cell_list reports.
__global__ void __launch_bounds__(BLK) k_cell_zero(const RepC *__restrict__ reps, const ChainS *__restrict__ chain,
                                                   int *cnt, int *wl_n) {
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < 3) wl_n[threadIdx.x] = 0;
    const RepC R = reps[blockIdx.x];
    const int c = blockIdx.y * BLK + threadIdx.x;
    if (c >= R.ncell || !need_rebuild(R, chain[blockIdx.x])) return;
    cnt[R.off_cell + c] = 0;
}

__device__ __forceinline__ int clampc(int c, int n) { return min(max(c, 0), n - 1); }

// Binning: the reference cell id (bit-exact, for update_cell_list) and the
// list-grid cell of every particle, with the place within its cell taken by
// an atomic on the cell's count; the particle's (slot, key) is written to
// that place of its cell's row of CELL_CAP entries (the reference's
// InCell[cell][MaxPartInCell], MDcode/global.h:20,85; a fuller cell sets the
// capacity error).  key: the octant of the cell the particle is in (x bit
// highest where the list teams cull, else z), then its id, the order
// k_cell_sort puts the cell in.  k_bin_end
// then turns the cell counts into start offsets.
#define CELL_CAP 600
__global__ void __launch_bounds__(BLK) k_bin(const RepC *__restrict__ reps, ChainS *chain,
                                             const int *__restrict__ blk_rep, const int *__restrict__ blk_first,
                                             const double *__restrict__ x, const double *__restrict__ y,
                                             const double *__restrict__ z, const int *__restrict__ id, int *cnt,
                                             int *cell_of, int2 *cell_mem, int *cell_ref) {
    const int b = blockIdx.x, r = blk_rep[b];
    const RepC R = reps[r];
    if (!need_rebuild(R, chain[r])) return;  // uniform over the replica's blocks
    const int i = blk_first[b] + threadIdx.x, lane = threadIdx.x & 63;
    const bool valid = i < R.n;
    const int g = R.off + i;
    int c = -1, key = 0;
    if (valid) {
        const double sx = x[g], sy = y[g], sz = z[g];
        // reference cell id, cell_list+forces.c:73-74 (floor in double, then sum)
        const int rx = clampc((int)floor(sx * R.inv_cell_ref), R.nc_ref),
                  ry = clampc((int)floor(sy * R.inv_cell_ref), R.nc_ref),
                  rz = clampc((int)floor(sz * R.inv_cell_ref), R.nc_ref);
        cell_ref[g] = rx + ry * R.nc_ref + rz * R.nc_ref * R.nc_ref;
        const int cx = clampc((int)floor(sx * R.nc), R.nc), cy = clampc((int)floor(sy * R.nc), R.nc),
                  cz = clampc((int)floor(sz * R.nc), R.nc);
        c = cx + cy * R.nc + cz * R.nc * R.nc;
        cell_of[g] = c;
        // octant within the cell: the parity of the cell index on the grid
        // of half-width cells, clamped as the cell index is (s * 2nc is
        // exactly twice s * nc, so the half agrees with the cell).  Where the
        // list teams cull (RepC::xcull) the x bit is highest, so each x-half
        // of a cell is one slot range (k_cell_sort); elsewhere the z bit (one
        // N=10^6 box: 2.48e9 against 2.45e9 particle-steps/s with x highest).
        // n < 2^21, so the id fits below bit 21.
        const int ox = clampc((int)floor(sx * (2 * R.nc)), 2 * R.nc) & 1,
                  oy = clampc((int)floor(sy * (2 * R.nc)), 2 * R.nc) & 1,
                  oz = clampc((int)floor(sz * (2 * R.nc)), 2 * R.nc) & 1;
        key = ((R.xcull ? ((ox << 2) | (oy << 1) | oz) : (ox | (oy << 1) | (oz << 2))) << 21) | id[g];
    }
    // one returning atomic per particle, all in flight at once: the place
    // within the cell row depends on the order the atomics land in, which
    // k_cell_sort removes (it orders each cell by key)
    // Places within the cells: the wave's particles form runs of equal cells
    // (they are in cell order from the last rebuild, up to the few that
    // moved); the first lane of each run takes the run's places with one
    // atomic, all runs at once, and the run's lanes follow in lane order.  A
    // cell that reappears in another run of the wave gets a second atomic.
    const int cp = __shfl_up(c, 1, 64);
    const bool head = valid && (lane == 0 || cp != c);  // c == -1 on invalid lanes
    const unsigned long long heads = __ballot(head);
    const unsigned long long vmask = __ballot(valid);
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    const unsigned long long hu = heads & upto;
    const int lead = hu ? 63 - __clzll((long long)hu) : lane;  // this lane's run head
    int my_rank = 0;
    if (head) {
        // run: from this lane to the next head or the last valid lane
        const unsigned long long after = heads & ~upto;
        const int next = after ? __ffsll((long long)after) - 1 : 64 - __clzll((long long)vmask);
        my_rank = atomicAdd(&cnt[R.off_cell + c], next - lane);
    }
    const int base = __shfl(my_rank, valid ? lead : lane, 64);
    if (valid) my_rank = base + (lane - lead);
    if (valid) {
        if (my_rank < CELL_CAP) cell_mem[(size_t)(R.off_cell + c) * CELL_CAP + my_rank] = make_int2(i, key);
        else atomicOr(&chain[r].err, 1);
    }
}

// After k_bin, one workgroup per replica due for a rebuild: the cell counts
// turned into start offsets (exclusive scan) and the replica's share of the
// step's rebuild work appended to the work list.  A kernel of its own rather
// than the last k_bin block of each replica, which needs an agent-scope
// fence per block (on gfx950 a write-back and invalidation of the XCD's L2).
// BIN_END_T threads per replica where a slot holds more than BIN_END_T cells
// (fixed per handle at creation), else BLK.
#define BIN_END_T 1024
__global__ void __launch_bounds__(BIN_END_T) k_bin_end(const RepC *__restrict__ reps,
                                                       const ChainS *__restrict__ chain, const int *__restrict__ cnt,
                                                       int *start, WorkList wl) {
    __shared__ int s_w[BIN_END_T / 64], s_base[3];
    const int r = blockIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6, nt = blockDim.x;
    const RepC R = reps[r];
    if (!need_rebuild(R, chain[r])) return;  // uniform over the workgroup
    const int *vc = cnt + R.off_cell;
    const int per = (R.ncell + nt - 1) / nt;
    const int c0 = min((int)threadIdx.x * per, R.ncell), c1 = min(c0 + per, R.ncell);
    int own = 0;
    for (int q = c0; q < c1; q++) own += vc[q];
    int v = own;
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    if (lane == 63) s_w[w] = v;
    __syncthreads();
    int acc = v - own;
    for (int k = 0; k < w; k++) acc += s_w[k];
    for (int q = c0; q < c1; q++) {
        start[R.off_cell + q] = acc;
        acc += vc[q];
    }
    // this replica's share of the step's rebuild work
    const int nb = (R.n + BLK - 1) / BLK, nl = (R.n + LPB - 1) / LPB, ncb = (R.ncell + BLK / 64 - 1) / (BLK / 64);
    if (threadIdx.x == 0) {
        s_base[0] = atomicAdd(&wl.n[0], nb);
        s_base[1] = atomicAdd(&wl.n[1], nl);
        s_base[2] = atomicAdd(&wl.n[2], ncb);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < nb; q += nt) wl.item[s_base[0] + q] = R.blk_off + q;
    for (int q = threadIdx.x; q < nl; q += nt) wl.item[wl.off_l + s_base[1] + q] = R.lblk_off + q;
    for (int q = threadIdx.x; q < ncb; q += nt) wl.item[wl.off_c + s_base[2] + q] = (r << 16) | q;
}

// Distinct neighbour-cell coordinates along one axis (cells are at least one
// list radius wide); with fewer than 3 cells per axis every cell is a
// neighbour once (an aliased stencil counts pairs twice, cf.
// cell_list+forces.c:37-49 with two cells per axis).
__device__ __forceinline__ int stencil_axis(int c, int nc, int *out) {
    if (nc >= 3) { out[0] = (c + nc - 1) % nc; out[1] = c; out[2] = (c + 1) % nc; return 3; }
    for (int d = 0; d < nc; d++) out[d] = d;
    return nc;
}

// Stencil of a cell as contiguous slot ranges: the 9 (z, y) rows of the
// 27-cell stencil, each as at most two runs of consecutive cells along x (two
// where the row wraps), in stencil order.  Cells of one x-run are adjacent in
// cell order, so their members form one slot range.  NRANGE = 18 entries per
// cell, empty ones (s == e) where the grid has fewer than 3 cells per axis.
#define NRANGE 18
__device__ __forceinline__ void xruns(int cx, int nc, int *f0, int *l0, int *f1, int *l1) {
    if (nc >= 3) {
        if (cx == 0) { *f0 = nc - 1; *l0 = nc - 1; *f1 = 0; *l1 = 1; }
        else if (cx == nc - 1) { *f0 = nc - 2; *l0 = nc - 1; *f1 = 0; *l1 = 0; }
        else { *f0 = cx - 1; *l0 = cx + 1; *f1 = 0; *l1 = -1; }
    } else {
        *f0 = 0; *l0 = nc - 1; *f1 = 0; *l1 = -1;  // every cell once (stencil_axis)
    }
}

struct PartArrays {  // particle arrays that move with a particle
    double *x, *y, *z, *xu, *yu, *zu, *px, *py, *pz;
    int *type, *act, *id, *cell_ref, *cell_of;
};

// The particle arrays as a table of 2*PART_FIELDS device pointers: P's
// x y z xu yu zu px py pz (doubles) type act id cell_ref cell_of (ints), then
// the permuted copy A's in the same order.
#define PART_FIELDS 14

// the position copies the list builders and the force read, at a new slot
struct PosCopies {
    double4 *pos4;
    QPos *posq;
    double *xo, *yo, *zo;  // x_u at this rebuild (update_cell_list, cell_list+forces.c:141-146)
};

// new slot gn <- old slot go of every array that moves with a particle, into
// dst, and the position copies at gn
__device__ __forceinline__ void gather_one(const PartArrays &src, const PartArrays &dst, const PosCopies &pc, int gn,
                                           int go) {
    const double sx = src.x[go], sy = src.y[go], sz = src.z[go], ux = src.xu[go], uy = src.yu[go], uz = src.zu[go];
    const int t = src.type[go];
    dst.x[gn] = sx; dst.y[gn] = sy; dst.z[gn] = sz;
    dst.xu[gn] = ux; dst.yu[gn] = uy; dst.zu[gn] = uz;
    dst.px[gn] = src.px[go]; dst.py[gn] = src.py[go]; dst.pz[gn] = src.pz[go];
    dst.type[gn] = t; dst.act[gn] = src.act[go]; dst.id[gn] = src.id[go];
    dst.cell_ref[gn] = src.cell_ref[go]; dst.cell_of[gn] = src.cell_of[go];
    pc.pos4[gn] = make_double4(sx, sy, sz, (double)t);
    pc.posq[gn] = qpos(sx, sy, sz, t);
    pc.xo[gn] = ux;
    pc.yo[gn] = uy;
    pc.zo[gn] = uz;
}

// Members of a cell by octant of the cell (x bit highest where the list teams
// cull, else z; k_bin), and in increasing original index within an octant
// (the binning order of cell_list+forces.c:71-80): a canonical order, so lists and force sums do not
// depend on launch timing or on which replicas share the launch.  Consecutive
// slots are then nearer in space than in plain index order, so the 64 lanes of
// a force gather touch fewer cache lines.  One wave per cell: keys staged in
// LDS, each lane ranks its entries by counting and moves each of its particles
// to its new slot (start of the cell + rank) in dst, so consecutive ranks
// write consecutive slots.  xmid: the first slot of the cell's upper x-half
// (used only where the x bit is highest).
#define SORT_CAP 256
#define XHALF (4 << 21)  // keys from here on: the upper x-half of the cell
__global__ void __launch_bounds__(BLK) k_cell_sort(const RepC *__restrict__ reps, WorkList wl,
                                                   const int *__restrict__ cnt, const int *__restrict__ start,
                                                   int2 *__restrict__ cell_mem,
                                                   int *members, int2 *ranges, int *xmid, PartArrays src,
                                                   PartArrays dst, PosCopies pc) {
    __shared__ int s_m[BLK / 64][SORT_CAP];
    __shared__ int s_k[BLK / 64][SORT_CAP];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nit = wl.n[2];
    for (int it = blockIdx.x; it < nit; it += gridDim.x) {
        const int item = wl.item[wl.off_c + it], r = item >> 16;
        const RepC R = reps[r];
        const int c = (item & 0xffff) * (BLK / 64) + w;  // one cell per wave
        if (c >= R.ncell) continue;
        if (lane < NRANGE) {  // the cell's stencil as slot ranges, for the list builders
            const int nc = R.nc, cx = c % nc, cy = (c / nc) % nc, cz = c / (nc * nc);
            int ay[3], az[3];
            const int ny = stencil_axis(cy, nc, ay), nz = stencil_axis(cz, nc, az);
            const int row = lane >> 1, part = lane & 1, iz = row / 3, iy = row - 3 * iz;
            int2 rg = make_int2(0, 0);
            if (iz < nz && iy < ny) {
                int f0, l0, f1, l1;
                xruns(cx, nc, &f0, &l0, &f1, &l1);
                const int f = part ? f1 : f0, l = part ? l1 : l0;
                if (l >= f) {
                    const int rb = R.off_cell + ay[iy] * nc + az[iz] * nc * nc;
                    rg = make_int2(R.off + start[rb + f], R.off + start[rb + l] + cnt[rb + l]);
                }
            }
            ranges[(size_t)(R.off_cell + c) * NRANGE + lane] = rg;
        }
        const int first = start[R.off_cell + c];
        int *m = members + R.off + first;
        // the cell's members in binning order (a cell beyond CELL_CAP has set
        // the capacity error; only its first CELL_CAP are moved)
        int2 *row = cell_mem + (size_t)(R.off_cell + c) * CELL_CAP;
        const int n = min(cnt[R.off_cell + c], CELL_CAP);
        if (n <= SORT_CAP) {
            for (int k = lane; k < n; k += 64) { const int2 v = row[k]; s_m[w][k] = v.x; s_k[w][k] = v.y; }
            __syncwarp();
            int lo = 0;
            for (int k0 = 0; k0 < n; k0 += 64) lo += __popcll(__ballot(k0 + lane < n && s_k[w][k0 + lane] < XHALF));
            if (lane == 0) xmid[R.off_cell + c] = R.off + first + lo;
            for (int k = lane; k < n; k += 64) {
                const int v = s_k[w][k];
                int rank = 0;
                for (int q = 0; q < n; q++) rank += s_k[w][q] < v;
                m[rank] = s_m[w][k];
                gather_one(src, dst, pc, R.off + first + rank, R.off + s_m[w][k]);
            }
            __syncwarp();
        } else {
            if (lane == 0) {  // crowded cell: serial insertion sort of the row on the key
                for (int a = 1; a < n; a++) {
                    const int2 v = row[a];
                    int k = a - 1;
                    while (k >= 0 && row[k].y > v.y) { row[k + 1] = row[k]; k--; }
                    row[k + 1] = v;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");  // lane 0's order, read by every lane
            __syncwarp();
            int lo = 0;
            for (int k0 = 0; k0 < n; k0 += 64) lo += __popcll(__ballot(k0 + lane < n && row[k0 + lane].y < XHALF));
            if (lane == 0) xmid[R.off_cell + c] = R.off + first + lo;
            for (int k = lane; k < n; k += 64) {
                const int sl = row[k].x;
                m[k] = sl;
                gather_one(src, dst, pc, R.off + first + k, R.off + sl);
            }
        }
    }
}


// the permuted copy back to the primary arrays, to[g] <- from[g] (the list
// builders other than the teams, which move their own particles back)
__global__ void __launch_bounds__(BLK) k_commit(const RepC *__restrict__ reps, const int *__restrict__ blk_rep,
                                                const int *__restrict__ blk_first, WorkList wl, PartArrays to,
                                                PartArrays from) {
    const int nit = wl.n[0];
    for (int it = blockIdx.x; it < nit; it += gridDim.x) {
        const int b = wl.item[it], r = blk_rep[b];
        const RepC R = reps[r];
        const int i = blk_first[b] + threadIdx.x;
        if (i >= R.n) continue;
        const int g = R.off + i;
        to.x[g] = from.x[g]; to.y[g] = from.y[g]; to.z[g] = from.z[g];
        to.xu[g] = from.xu[g]; to.yu[g] = from.yu[g]; to.zu[g] = from.zu[g];
        to.px[g] = from.px[g]; to.py[g] = from.py[g]; to.pz[g] = from.pz[g];
        to.type[g] = from.type[g]; to.act[g] = from.act[g]; to.id[g] = from.id[g];
        to.cell_ref[g] = from.cell_ref[g]; to.cell_of[g] = from.cell_of[g];
    }
}


// List radius per species pair: its force cutoff 2.5*sigma_ab plus the skin
// (0.35 in the reference, MDcode/global.h:25-31, which applies 2.85 to every
// pair); pairs are stored when ((dx*dx+dy*dy)+dz*dz)*L*L < (2.5 sigma_ab +
// skin)^2.  The reference list (< 2.85^2 for every pair,
// cell_list+forces.c:129-130) is counted beside it.
__device__ __forceinline__ double list_cut2(int ab, double skin) {
    const double rc = (ab == 0 ? (2.5 * 1.00) : (ab == 1 ? (2.5 * 0.80) : (2.5 * 0.88))) + skin;
    return rc * rc;
}

// Each replica owns cap*n list entries from off*cap.  Neighbour k of local
// particle i lives at ((k/C)*n + i)*C + k%C in that region, C = 2^lc: with a
// team of tpp > 1 lanes per particle C = tpp, so the team reads tpp
// consecutive entries per chunk; with one lane per particle C = 8, so a lane
// reads its next 8 entries as two 16-byte loads and a wave a contiguous 2 KB.
__device__ __forceinline__ size_t nl_at(const RepC &R, int cap, int k, int i) {
    return (size_t)R.off * cap + (((size_t)(k >> R.lc) * R.n + i) << R.lc) + (k & ((1 << R.lc) - 1));
}

// Dynamics-list screen shared by every builder (so all of them keep exactly
// the same pairs): minimum image from the fixed-point positions, r^2 in FP32
// in fixed-point units (fused multiply-adds) against the list radius plus a
// margin of 1e-3 sigma^2 in those units, far above the FP32 error of r^2
// (below 1e-5 sigma^2 for the list radii), so no pair within its list radius
// is lost.  Which pairs beyond the force cutoff the margin lets in does not
// change any force: a pair contributes only inside 2.5 sigma_ab.
__device__ __forceinline__ bool qscreen(const QPos &pi, const QPos &q, float ua, float ub) {
    const float dx = (float)(int)(pi.x - q.x), dy = (float)(int)(pi.y - q.y), dz = (float)(int)(pi.z - q.z);
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)) < ((q.x & 1u) ? ub : ua);
}
// list radius^2 of species pair ab plus the screen's margin, in fixed-point units^2
__device__ __forceinline__ float qscreen_cut(int ab, const RepC &R) {
    return (float)((list_cut2(ab, R.skin) + 1e-3) / (double)R.qs2);
}

#define NL_UNROLL 8
// Thread per particle (slots are in cell order, so the lanes of a wave scan
// the same few cells and the position loads are shared in L1); candidate
// positions are loaded NL_UNROLL at a time to keep several loads in flight.
// Dynamics lists (EXACT=false) screen candidates in FP32 with a margin of
// 1e-3 sigma^2, far above the FP32 error of r^2 (< 1e-4 for L < 100), so no
// pair within its list radius is lost.  EXACT=true (update_cell_list) tests
// every candidate with the reference's FP64 expression and also counts the
// reference's 2.85 list.
template <bool EXACT>
__global__ void __launch_bounds__(BLK) k_nlist(const RepC *__restrict__ reps, ChainS *chain,
                                               const int *__restrict__ blk_rep, const int *__restrict__ blk_first,
                                               const double4 *__restrict__ pos4, const QPos *__restrict__ posq,
                                               const int *__restrict__ cell_of, const int *__restrict__ cnt,
                                               const int *__restrict__ start, int *nn, int *nn_ref, int *nl,
                                               int stride, int cap) {
    const int b = blockIdx.x, r = blk_rep[b];
    const RepC R = reps[r];
    if (!need_rebuild(R, chain[r])) return;
    const int i = blk_first[b] + threadIdx.x;
    if (i >= R.n) return;
    const int g = R.off + i, nc = R.nc;
    const double4 pi = pos4[g];
    const QPos pq = posq[g];
    const int ti = (int)pi.w;
    const int c = cell_of[g];
    const int cx = c % nc, cy = (c / nc) % nc, cz = c / (nc * nc);
    int ax[3], ay[3], az[3];
    const int nx = stencil_axis(cx, nc, ax), ny = stencil_axis(cy, nc, ay), nz = stencil_axis(cz, nc, az);
    const double L = R.L;
    const double cut_a = list_cut2(ti, R.skin), cut_b = list_cut2(ti + 1, R.skin);
    const float ua = qscreen_cut(ti, R), ub = qscreen_cut(ti + 1, R);
    int count = 0, cref = 0;
    for (int iz = 0; iz < nz; iz++)
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                const int c2 = R.off_cell + ax[ix] + ay[iy] * nc + az[iz] * nc * nc;
                const int s2 = R.off + start[c2], e2 = s2 + cnt[c2];
                for (int k = s2; k < e2; k += NL_UNROLL) {
                    if (EXACT) {
                        double4 q[NL_UNROLL];
#pragma unroll
                        for (int u = 0; u < NL_UNROLL; u++) q[u] = pos4[min(k + u, e2 - 1)];
#pragma unroll
                        for (int u = 0; u < NL_UNROLL; u++) {
                            const int j = k + u;
                            if (j >= e2 || j == g) continue;
                            // cell_list+forces.c:120-130, same operation order
                            double dx = minimg_md(pi.x - q[u].x);
                            double dy = minimg_md(pi.y - q[u].y);
                            double dz = minimg_md(pi.z - q[u].z);
                            double dr2 = (dx * dx + dy * dy + dz * dz) * L * L;
                            cref += dr2 < K_CELL_CUT_SQ;
                            if (dr2 < (q[u].w == 0.0 ? cut_a : cut_b)) {
                                if (count < cap) nl[nl_at(R, cap, count, i)] = j;
                                count++;
                            }
                        }
                    } else {
                        QPos q[NL_UNROLL];
#pragma unroll
                        for (int u = 0; u < NL_UNROLL; u++) q[u] = posq[min(k + u, e2 - 1)];
#pragma unroll
                        for (int u = 0; u < NL_UNROLL; u++) {
                            const int j = k + u;
                            if (j >= e2 || j == g) continue;
                            if (qscreen(pq, q[u], ua, ub)) {
                                if (count < cap) nl[nl_at(R, cap, count, i)] = j;
                                count++;
                            }
                        }
                    }
                }
            }
    if (count > cap) atomicOr(&chain[r].err, 1);
    nn[g] = min(count, cap);
    if (EXACT) nn_ref[g] = cref;
}

// Dynamics lists, a team of NLT lanes per particle (16, 32 or 64; LPB = 16
// particles per workgroup): the team tests NLT consecutive candidates at once
// and appends the hits in candidate order (team ballot + popcount of the
// lanes below), so the list is in the order of a serial scan (stencil cells
// z, y, x; slots within a cell) while the chain of dependent loads per
// particle is NLT times shorter.  Only the replicas due for a rebuild do work
// in a step, so wide teams are what keeps enough waves in flight.  The
// control flow is uniform within a team (one home particle per team), so
// every lane of a team is active at each ballot.  FP32 screen with the margin
// described at k_nlist.
//
// The stencil's slot ranges are cut into chunks of NLT candidates, listed by
// the team in its LDS row (ct, CH_CAP entries), and the team takes CH_PER
// chunks per round, with CH_PER loads in flight per lane across range
// boundaries: about half the dependent rounds of a range-by-range walk, the
// same candidates in the same order.  A stencil of more than CH_CAP chunks
// (crowded cells, narrow teams) is walked range by range.
//
// Where the replica allows it (RepC::xcull), the range holding the cell
// beside the home cell along x on the far side loses that cell's far x-half:
// its start moves to the cell's xmid for a particle in the upper x-half of
// its own cell (the cell below), its end for one in the lower half (the cell
// above).  Those candidates are at least one cell width away, beyond every
// list radius, so the lists are the same pairs in the same order; each of the
// 9 rows loses half a cell, 1/6 of the candidates.
#define CH_CAP 64
#define CH_PER 4
template <int NLT, bool CULL>
__device__ __forceinline__ void nlist_team_one(const RepC &R, int r, int i, ChainS *chain,
                                               const QPos *__restrict__ posq, const int *__restrict__ cell_of,
                                               const int2 *__restrict__ ranges,
                                               const int *__restrict__ xmid, int *nn, int *nl, int cap, int2 *ct) {
    const int lane = threadIdx.x & 63, t = lane & (NLT - 1);
    const unsigned long long tmask = (NLT == 64 ? ~0ull : ((1ull << (NLT & 63)) - 1ull)) << (lane & ~(NLT - 1));
    const unsigned long long below = (1ull << lane) - 1ull;
    const int g = R.off + i;
    const QPos pi = posq[g];
    const int ti = ptype(pi);
    const int cg = cell_of[g];
    const int2 *rgs = ranges + (size_t)(R.off_cell + cg) * NRANGE;
    const float ua = qscreen_cut(ti, R), ub = qscreen_cut(ti + 1, R);
    int count = 0;
    auto inlist = [&](const QPos &q) -> bool { return qscreen(pi, q, ua, ub); };
    // entry p of this particle's list at nlb[(p >> lc) * nC + (p & cm)] (nl_at
    // relative to entry 0; 24-bit products, n < 2^21 is checked at creation)
    int *const nlb = nl + nl_at(R, cap, 0, i);
    const unsigned nC = (unsigned)R.n << R.lc, cm = (1u << R.lc) - 1u;
    const int lc = R.lc;
    // a hit's place: the team's hits of lower lanes (the wave's lower hits,
    // mbcnt, less those of the teams below) after the team's count so far
    auto emit = [&](bool hit, int k) {
        const unsigned long long b = __ballot(hit);
        const unsigned lo = (unsigned)b, hi = (unsigned)(b >> 32);
        const int wr = (int)__builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u));
        int rank, tot;
        if (NLT == 64) {
            rank = wr;
            tot = __popcll(b);
        } else if (NLT == 32) {
            const bool up = lane >= 32;
            const int pl = __popc(lo), ph = __popc(hi);
            rank = wr - (up ? pl : 0);
            tot = up ? ph : pl;
        } else {
            const unsigned long long m = b & tmask;
            rank = __popcll(m & below);
            tot = __popcll(m);
        }
        if (hit) {
            const unsigned pos = (unsigned)(count + rank);
            if (pos < (unsigned)cap) nlb[__umul24(pos >> lc, nC) + (pos & cm)] = k;
        }
        count += tot;
    };
    // the home cell's stencil as slot ranges (k_cell_sort), in stencil order:
    // lane t holds ranges t and t + NLT (when NLT < NRANGE), and the team
    // lists their chunks in range order
    int2 ra = make_int2(0, 0), rb = make_int2(0, 0);
    if (t < NRANGE) ra = rgs[t];
    if (NLT < NRANGE && t + NLT < NRANGE) rb = rgs[t + NLT];
    if (CULL && R.xcull) {
        const int nc = R.nc, cx = cg % nc, cy = (cg / nc) % nc, cz = cg / (nc * nc);
        const bool up = g >= xmid[R.off_cell + cg];
        // the cell whose far half goes, and the part of its row's runs
        // holding it (xruns: part 1 where the home cell is at either end)
        const int xc = up ? (cx == 0 ? nc - 1 : cx - 1) : (cx == nc - 1 ? 0 : cx + 1);
        const int qp = (!up && (cx == 0 || cx == nc - 1)) ? 1 : 0;
        auto cull = [&](int q, int2 &rg) {
            if (q >= NRANGE || (q & 1) != qp) return;
            const int row = q >> 1, iz = row / 3, iy = row - 3 * iz;
            const int yy = (cy + iy + nc - 1) % nc, zz = (cz + iz + nc - 1) % nc;
            const int m = xmid[R.off_cell + xc + yy * nc + zz * nc * nc];
            if (up) rg.x = m;
            else rg.y = m;
        };
        cull(t, ra);
        if (NLT < NRANGE) cull(t + NLT, rb);
    }
    const int na = (ra.y - ra.x + NLT - 1) / NLT, nb = (rb.y - rb.x + NLT - 1) / NLT;
    int pa = na, pb = nb;  // inclusive scans over the team
    for (int o = 1; o < NLT; o <<= 1) {
        const int ua = __shfl_up(pa, o, NLT), ub = __shfl_up(pb, o, NLT);
        if (t >= o) { pa += ua; pb += ub; }
    }
    const int tot_a = __shfl(pa, NLT - 1, NLT), nch = tot_a + __shfl(pb, NLT - 1, NLT);
    if (nch <= CH_CAP) {
        for (int j = 0; j < na; j++) ct[pa - na + j] = make_int2(ra.x + j * NLT, min(NLT, ra.y - ra.x - j * NLT));
        for (int j = 0; j < nb; j++)
            ct[tot_a + pb - nb + j] = make_int2(rb.x + j * NLT, min(NLT, rb.y - rb.x - j * NLT));
        if (t < CH_PER) ct[nch + t] = make_int2(0, 0);  // empty chunks up to the next whole round
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // the row is read by the other lanes
        for (int c0 = 0; c0 < nch; c0 += CH_PER) {
            int2 ch[CH_PER];
            QPos qv[CH_PER];
#pragma unroll
            for (int u = 0; u < CH_PER; u++) ch[u] = ct[c0 + u];
#pragma unroll
            for (int u = 0; u < CH_PER; u++) qv[u] = posq[t < ch[u].y ? ch[u].x + t : g];
#pragma unroll
            for (int u = 0; u < CH_PER; u++) {
                const int k = ch[u].x + t;
                emit((t < ch[u].y) & (k != g) & inlist(qv[u]), k);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // reads done before the row is rewritten
    } else {
        for (int q = 0; q < NRANGE; q++) {
            const int2 rg = rgs[q];
            const int s2 = rg.x, e2 = rg.y;
            // two candidates per lane, both loaded before either is used
            // (indices clamped into [s2, e2), validity applied after)
            for (int k0 = s2; k0 < e2; k0 += 2 * NLT) {
                const int ka = k0 + t, kb = k0 + NLT + t;
                const QPos qa = posq[min(ka, e2 - 1)], qb = posq[min(kb, e2 - 1)];
                const bool ha = inlist(qa) && ka < e2 && ka != g, hb = inlist(qb) && kb < e2 && kb != g;
                emit(ha, ka);
                emit(hb, kb);
            }
        }
    }
    if (t == 0) {
        if (count > cap) atomicOr(&chain[r].err, 1);
        nn[g] = min(count, cap);
    }
}

// Fixed grid over the step's list blocks (LPB particles each, one team per
// particle); plain block order spreads the few rebuilding replicas over all
// XCDs.
// CULL: the x-half culling compiled in (handles with a replica that allows
// it); without it the kernel keeps the registers of the plain walk (N=10^6:
// 2.45e9 against 2.55e9 particle-steps/s with the culling code present but
// unused).
template <int NLT, bool MOVE, bool CULL>
__global__ void __launch_bounds__(NLT * LPB) k_nlist_team(const RepC *__restrict__ reps, ChainS *chain,
                                                          const int *__restrict__ lblk_rep,
                                                          const int *__restrict__ lblk_first, WorkList wl,
                                                          const QPos *__restrict__ posq,
                                                          const int *__restrict__ cell_of,
                                                          const int2 *__restrict__ ranges,
                                                          const int *__restrict__ xmid, int *nn, int *nl, int cap,
                                                          void *const *__restrict__ ftab) {
    __shared__ int2 s_ct[LPB][CH_CAP + CH_PER];  // each team's chunk list, then empty chunks
    const int nit = wl.n[1];
    for (int it = blockIdx.x; it < nit; it += gridDim.x) {
        const int b = wl.item[wl.off_l + it], r = lblk_rep[b];
        const RepC R = reps[r];
        // MOVE: the block's LPB particles go back from the permuted copy
        // (second half of ftab) to the primary arrays (first half) here
        // instead of in k_commit: thread t moves field t / LPB of particle
        // t % LPB, loaded here and stored after the lists, so each field's LPB
        // consecutive slots are one coalesced write.  The list build reads
        // cell_of of the copy, nothing else.  The held value costs a wave of
        // occupancy, so this pays where k_commit's launch dominates (small
        // ensembles), not for large boxes.
        const int cf = (int)threadIdx.x / LPB, cg = R.off + lblk_first[b] + (int)threadIdx.x % LPB;
        const bool cpy = MOVE && cf < PART_FIELDS && cg < R.off + R.n;
        long long cv = 0;
        if (cpy)
            cv = cf < 9 ? ((const long long *)ftab[PART_FIELDS + cf])[cg] : ((const int *)ftab[PART_FIELDS + cf])[cg];
        const int i = lblk_first[b] + (int)threadIdx.x / NLT;
        if (i < R.n)  // whole team
            nlist_team_one<NLT, CULL>(R, r, i, chain, posq, cell_of, ranges, xmid, nn, nl, cap, s_ct[threadIdx.x / NLT]);
        if (cpy) {
            if (cf < 9) ((long long *)ftab[cf])[cg] = cv;
            else ((int *)ftab[cf])[cg] = (int)cv;
        }
    }
}

// Dynamics lists, one wave per cell: the particles of the 27 neighbour cells
// (contiguous slot ranges) are staged in LDS in stencil order; then for each
// home particle the 64 lanes test 64 candidates at once and the hits are
// appended in candidate order by ballot + popcount, so every list is in the
// same canonical order as a serial scan.  FP32 screen with the margin
// described at k_nlist.
#define NLC_CAP 2304
__global__ void __launch_bounds__(64) k_nlist_cell(const RepC *__restrict__ reps, ChainS *chain,
                                                   const int *__restrict__ cblk_rep, const int *__restrict__ cblk_cell,
                                                   const QPos *__restrict__ posq, const int *__restrict__ cnt,
                                                   const int *__restrict__ start, int *nn, int *nl, int cap) {
    __shared__ QPos s_c[NLC_CAP];
    __shared__ int s_pre[28], s_st[27];
    const int r = cblk_rep[blockIdx.x], c = cblk_cell[blockIdx.x];
    const RepC R = reps[r];
    if (c >= R.ncell || !need_rebuild(R, chain[r])) return;
    const int lane = threadIdx.x, nc = R.nc;
    const int cx = c % nc, cy = (c / nc) % nc, cz = c / (nc * nc);
    int ax[3], ay[3], az[3];
    const int nx = stencil_axis(cx, nc, ax), ny = stencil_axis(cy, nc, ay), nz = stencil_axis(cz, nc, az);
    const int ns = nx * ny * nz;
    // lane q < ns reads stencil cell q (stencil order: x fastest, then y, z)
    int my_n = 0, my_s = 0, is_home = 0;
    if (lane < ns) {
        const int cc = ax[lane % nx] + ay[(lane / nx) % ny] * nc + az[lane / (nx * ny)] * nc * nc;
        my_n = cnt[R.off_cell + cc];
        my_s = start[R.off_cell + cc];
        is_home = cc == c;
    }
    int pre = my_n;  // inclusive scan over the stencil cells
    for (int o = 1; o < 32; o <<= 1) {
        const int v = __shfl_up(pre, o, 64);
        if (lane >= o) pre += v;
    }
    if (lane < ns) { s_pre[lane + 1] = pre; s_st[lane] = my_s; }
    if (lane == 0) s_pre[0] = 0;
    const unsigned long long hm = __ballot(is_home);
    const int hl = __ffsll((long long)hm) - 1;  // stencil index of the home cell
    const int total = __shfl(pre, ns - 1, 64);
    const int hn = __shfl(my_n, hl, 64), hs = __shfl(my_s, hl, 64), hoff = __shfl(pre - my_n, hl, 64);
    if (total > NLC_CAP) {
        if (lane == 0) atomicOr(&chain[r].err, 1);
        return;
    }
    __syncthreads();
    for (int k = lane; k < total; k += 64) {  // independent loads
        int q = 0;
        while (s_pre[q + 1] <= k) q++;
        s_c[k] = posq[R.off + s_st[q] + (k - s_pre[q])];
    }
    __syncthreads();
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int mine = 0;  // list length of home particle `lane` (and lane+64, ... kept below)
    for (int h = 0; h < hn; h++) {
        const QPos pi = s_c[hoff + h];
        const int ti = ptype(pi), kh = hoff + h;
        const float ua = qscreen_cut(ti, R), ub = qscreen_cut(ti + 1, R);
        int base = 0;
        for (int k0 = 0; k0 < total; k0 += 64) {
            const int k = k0 + lane;
            bool hit = false;
            if (k < total && k != kh) hit = qscreen(pi, s_c[k], ua, ub);
            const unsigned long long mask = __ballot(hit);
            if (hit) {
                const int pos = base + __popcll(mask & below);
                if (pos < cap) {
                    int q = 0;
                    while (s_pre[q + 1] <= k) q++;
                    nl[nl_at(R, cap, pos, hs + h)] = R.off + s_st[q] + (k - s_pre[q]);
                }
            }
            base += __popcll(mask);
        }
        if ((h & 63) == lane) mine = base;
        if ((h & 63) == 63 || h == hn - 1) {  // flush 64 finished counts
            const int h0 = h & ~63;
            if (lane <= h - h0) {
                if (mine > cap) atomicOr(&chain[r].err, 1);
                nn[R.off + hs + h0 + lane] = min(mine, cap);
            }
        }
    }
}

// ---- in-situ overlap Q(t) for every replica (device-resident accumulators)
// Origins every origin_every steps are kept as snapshots of x_u in reference
// order (ring of n_origin per replica); at a sample the overlap of every
// resident origin whose lag (step - origin step) is a multiple of
// sample_every below n_lags*sample_every is counted with the arithmetic of
// Analysis/overlap.c:61-84 (unwrapped difference, one minimum-image
// correction, L*L*((dx*dx+dy*dy)+dz*dz) < a*a).  Counts are integers, so the
// per-lag sums do not depend on the order of the threads.
struct OvCfg {
    int sample_every, origin_every, n_origin, n_lags;
    double a2;
};

__device__ __forceinline__ double minimg_ov(double d) {
    if (d >= 0.5)
        d -= 1.0;
    else if (d < -0.5)
        d += 1.0;
    return d;
}

// grid (particle blocks of all replicas, n_origin)
__global__ void __launch_bounds__(BLK) k_ov_count(const RepC *__restrict__ reps, const ChainS *__restrict__ chain,
                                                  const int *__restrict__ blk_rep, const int *__restrict__ blk_first,
                                                  OvCfg cfg, const int *__restrict__ id, const double *__restrict__ xu,
                                                  const double *__restrict__ yu, const double *__restrict__ zu,
                                                  double *ox, double *oy, double *oz, const long long *ostep,
                                                  unsigned int *ocnt) {
    const int b = blockIdx.x, r = blk_rep[b], o = blockIdx.y;
    const RepC R = reps[r];
    const long long step = chain[r].step;
    const int i = blk_first[b] + threadIdx.x;
    const size_t base = ((size_t)R.off * cfg.n_origin + (size_t)o * R.cap_n);  // snapshot o of replica r
    const long long s0 = ostep[(size_t)r * cfg.n_origin + o];
    int hit = 0;
    if (i < R.n && s0 >= 0 && step > s0 && (step - s0) % cfg.sample_every == 0 &&
        (step - s0) / cfg.sample_every <= cfg.n_lags) {
        const int g = R.off + i, k = id[g];
        double dx = xu[g] - ox[base + k], dy = yu[g] - oy[base + k], dz = zu[g] - oz[base + k];
        dz = minimg_ov(dz);
        dy = minimg_ov(dy);
        dx = minimg_ov(dx);
        hit = R.L * R.L * (dx * dx + dy * dy + dz * dz) < cfg.a2;
    }
    const int w = ac4_wave_sum_i(hit);
    if ((threadIdx.x & 63) == 0 && w) atomicAdd(&ocnt[(size_t)r * cfg.n_origin + o], (unsigned int)w);
}

// one thread per (replica, origin): fold the count into the lag slot, then
// store a new origin when the step is a multiple of origin_every
__global__ void k_ov_fold(const RepC *__restrict__ reps, const ChainS *__restrict__ chain, OvCfg cfg, int nrep,
                          long long *ostep, unsigned int *ocnt, double *sq, double *sq2, int *scnt) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrep * cfg.n_origin) return;
    const int r = t / cfg.n_origin;
    const RepC R = reps[r];
    const long long step = chain[r].step, s0 = ostep[t];
    if (R.n > 0 && s0 >= 0 && step > s0 && (step - s0) % cfg.sample_every == 0) {
        const long long lag = (step - s0) / cfg.sample_every;
        if (lag <= cfg.n_lags) {
            const double q = (double)ocnt[t] / (double)R.n;  // overlap.c:82
            const size_t k = (size_t)r * (cfg.n_lags + 1) + lag;
            sq[k] += q;        // overlap.c:89-94 (origin order: one origin per lag per step)
            sq2[k] += q * q;
            scnt[k] += 1;
        }
    }
    ocnt[t] = 0u;
}

__global__ void __launch_bounds__(BLK) k_ov_origin(const RepC *__restrict__ reps, const ChainS *__restrict__ chain,
                                                   const int *__restrict__ blk_rep, const int *__restrict__ blk_first,
                                                   OvCfg cfg, const int *__restrict__ id, const double *__restrict__ xu,
                                                   const double *__restrict__ yu, const double *__restrict__ zu,
                                                   double *ox, double *oy, double *oz, long long *ostep) {
    const int b = blockIdx.x, r = blk_rep[b];
    const RepC R = reps[r];
    const long long step = chain[r].step;
    if (R.n == 0 || step % cfg.origin_every != 0) return;
    const int o = (int)((step / cfg.origin_every) % cfg.n_origin);
    const int i = blk_first[b] + threadIdx.x;
    const size_t base = ((size_t)R.off * cfg.n_origin + (size_t)o * R.cap_n);
    if (i < R.n) {
        const int g = R.off + i, k = id[g];
        ox[base + k] = xu[g];
        oy[base + k] = yu[g];
        oz[base + k] = zu[g];
    }
    if (i == 0) ostep[(size_t)r * cfg.n_origin + o] = step;
}

// Ensemble chi4 from the per-lag accumulators, FP64 on the device: one thread
// per lag sums the listed replicas in list order (so the result does not
// depend on the launch), then Q = s/c and X4 = s2/c - s*s/(c*c) in the
// expression order of Analysis/overlap.c:114-118, chi4 = Nbar * X4 with Nbar
// the mean N of the listed replicas.  out[l] = {chi4, Q, origins}.
__global__ void k_ov_chi4(const RepC *__restrict__ reps, OvCfg cfg, const int *__restrict__ list, int nlist,
                          const double *__restrict__ sq, const double *__restrict__ sq2, const int *__restrict__ scnt,
                          double *out) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l > cfg.n_lags) return;
    double s = 0.0, s2 = 0.0, nbar = 0.0;
    long long c = 0;
    for (int k = 0; k < nlist; k++) {
        const int r = list[k];
        const size_t a = (size_t)r * (cfg.n_lags + 1) + l;
        s += sq[a];
        s2 += sq2[a];
        c += scnt[a];
        nbar += (double)reps[r].n;
    }
    nbar /= (double)nlist;
    double q = 0.0, x4 = 0.0;
    if (c > 0) {
        const double dc = (double)c;
        q = s / dc;
        x4 = s2 / dc - s * s / (dc * dc);
    }
    out[3 * l] = nbar * x4;
    out[3 * l + 1] = q;
    out[3 * l + 2] = (double)c;
}

// Clears replica r's accumulators and origins (a slot taking a new job).
__global__ void k_ov_reset(OvCfg cfg, int r, long long *ostep, unsigned int *ocnt, double *sq, double *sq2,
                           int *scnt) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < cfg.n_origin) {
        ostep[(size_t)r * cfg.n_origin + t] = -1;
        ocnt[(size_t)r * cfg.n_origin + t] = 0u;
    }
    if (t <= cfg.n_lags) {
        const size_t a = (size_t)r * (cfg.n_lags + 1) + t;
        sq[a] = 0.0;
        sq2[a] = 0.0;
        scnt[a] = 0;
    }
}

// Reference-order export of x_u and type for in-situ analysis.
__global__ void __launch_bounds__(BLK) k_export(const RepC *__restrict__ reps, const int *__restrict__ blk_rep,
                                                const int *__restrict__ blk_first, const int *__restrict__ id,
                                                const double *__restrict__ xu, const double *__restrict__ yu,
                                                const double *__restrict__ zu, const int *__restrict__ type,
                                                double *ex, double *ey, double *ez, int *et, int r_only) {
    const int b = blockIdx.x, r = blk_rep[b];
    if (r_only >= 0 && r != r_only) return;
    const RepC R = reps[r];
    const int i = blk_first[b] + threadIdx.x;
    if (i >= R.n) return;
    const int g = R.off + i, o = R.off + id[g];
    ex[o] = xu[g]; ey[o] = yu[g]; ez[o] = zu[g]; et[o] = type[g];
}

// ---- forces
template <bool FP32>
struct PairAcc;

template <>
struct PairAcc<false> {  // FP64: the reference's per-pair arithmetic
    double fx = 0, fy = 0, fz = 0, pe = 0, vir = 0;
    __device__ __forceinline__ void flush() {}
    __device__ __forceinline__ PairAcc(int, const RepC &) {}
    __device__ __forceinline__ void add(const double4 &pi, int ti, const double4 &pj, const RepC &R) {
        const double L = R.L;
        double dx = minimg_md(pi.x - pj.x);
        double dy = minimg_md(pi.y - pj.y);
        double dz = minimg_md(pi.z - pj.z);
        double dr2 = (dx * dx + dy * dy + dz * dz) * L * L;  // cell_list+forces.c:183
        const int tj = (int)pj.w;
        double eps = eps_ab(ti, tj), sig = sig_ab(ti, tj);
        double r_over_sig_2 = dr2 / (sig * sig);
        if (r_over_sig_2 < K_LCUTSQR) {  // :189-203
            double s2 = 1.0 / r_over_sig_2;
            double s6 = s2 * s2 * s2;
            pe += 4.0 * eps * ((s6 * (s6 - 1.0)) + K_C0 + K_C2 * r_over_sig_2);
            double grad = 8.0 * eps * (6.0 * s6 * (s6 - 0.5) / dr2 - K_C2 / (sig * sig));
            fx += grad * dx * L;
            fy += grad * dy * L;
            fz += grad * dz * L;
            vir += grad * dr2;
        }
    }
    __device__ __forceinline__ void result(double &ox, double &oy, double &oz, const RepC &) const {
        ox = fx; oy = fy; oz = fz;
    }
};

template <>
struct PairAcc<true> {
    // FP32 pair forces from fixed-point positions (qpos): exact minimum
    // image in integers, separations in units of L/2^32 rounded once to FP32.
    // The cutoff test runs in those units (r^2 in units^2 against the
    // cutoff divided by the unit^2), so a candidate outside costs no scaling.
    // Each FP32 pair force term is widened to FP64 and summed there (exactly
    // the compensation a Kahan sum aims at, and on CDNA4 a widening add costs
    // two FP64 instructions against five FP32 ones); the force is scaled to
    // real units at the end.  The pair energy and virial, which only feed the
    // reported PE, pressure and Heff and never the trajectory, are summed in
    // FP32 over each group of at most 8 list entries and widened once per
    // group (flush): k_force 33.4 -> 32.1 us on the 64-replica ensemble, the
    // trajectories bit for bit the same.  The pair constants of the two possible partner
    // species are set once per particle; one hardware reciprocal of r^2
    // (v_rcp_f32, within 1 ulp) replaces the two divisions, since
    // sigma^2/r^2 = (sigma/r)^2 and 1/r^2 appears in the force factor.  The
    // pair terms are formed with fused multiply-adds (each rounds once).
    double fx = 0, fy = 0, fz = 0, pe = 0, vir = 0;
    float gpe = 0.f, gvir = 0.f;  // FP32 partials of the energy and virial of up to 8 pairs
    float sg2[2], e4[2], e8[2], rcq[2], c2i[2], qs2;
    __device__ __forceinline__ PairAcc(int ti, const RepC &R) {
        // (ti, tj) -> ab = ti + tj: sigma {1, 0.8, 0.88}, epsilon {1, 1.5, 0.5}
        // (cell_list+forces.c:19-20); cutoff (r/sigma)^2 < 6.25 (:189)
        qs2 = R.qs2;
        for (int tj = 0; tj < 2; tj++) {
            const int ab = ti + tj;
            const double sig = ab == 0 ? 1.00 : (ab == 1 ? 0.80 : 0.88);
            const double eps = ab == 0 ? 1.00 : (ab == 1 ? 1.50 : 0.50);
            sg2[tj] = (float)(sig * sig);
            e4[tj] = (float)(4.0 * eps);
            e8[tj] = (float)(8.0 * eps);
            rcq[tj] = (float)(K_LCUTSQR * sig * sig / (double)R.qs2);
            c2i[tj] = (float)(K_C2 / (sig * sig));  // C2 / sigma^2
        }
    }
    __device__ __forceinline__ void add(const QPos &pi, int, const QPos &pj, const RepC &) {
        const float dx = (float)(int)(pi.x - pj.x);
        const float dy = (float)(int)(pi.y - pj.y);
        const float dz = (float)(int)(pi.z - pj.z);
        const bool b = (pj.x & 1u) != 0u;
        const float rq = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));  // (L/2^32)^2 units
        if (rq < (b ? rcq[1] : rcq[0])) {
            const float r2 = rq * qs2;  // real units
            const float c2 = b ? c2i[1] : c2i[0];
            const float inv = __builtin_amdgcn_rcpf(r2);
            const float s2 = (b ? sg2[1] : sg2[0]) * inv;  // (sigma/r)^2
            const float s6 = s2 * s2 * s2;
            // cell_list+forces.c:194-203: 4 eps [s6 (s6 - 1) + C0 + C2 (r/sigma)^2],
            // 8 eps [6 s6 (s6 - 1/2) / r^2 - C2 / sigma^2]
            const float pterm = (b ? e4[1] : e4[0]) * __builtin_fmaf(s6, s6 - 1.0f, __builtin_fmaf(c2, r2, (float)K_C0));
            const float grad = (b ? e8[1] : e8[0]) * __builtin_fmaf(6.0f * s6 * inv, s6 - 0.5f, -c2);
            fx += (double)(grad * dx);
            fy += (double)(grad * dy);
            fz += (double)(grad * dz);
            gpe += pterm;
            gvir = __builtin_fmaf(grad, r2, gvir);
        }
    }
    __device__ __forceinline__ void flush() {
        pe += (double)gpe; vir += (double)gvir;
        gpe = gvir = 0.f;
    }
    __device__ __forceinline__ void result(double &ox, double &oy, double &oz, const RepC &R) {
        flush();
        ox = fx * R.qscale;
        oy = fy * R.qscale;
        oz = fz * R.qscale;
    }
};

// Four consecutive 16-byte-aligned list entries as one 16-byte load.  (A
// nontemporal load here measured 79 against 56 us per k_force launch on the
// 64-replica ensemble: the list stream is better served by the caches.)
typedef int ac4_v4i __attribute__((ext_vector_type(4)));
__device__ __forceinline__ int4 ld_list4(const int *p) {
    const ac4_v4i v = *(const ac4_v4i *)p;
    return make_int4(v.x, v.y, v.z, v.w);
}

// One lane per particle: the list walk in groups of C entries (C/8 list
// chunks of two 16-byte index loads each), the indices of the next group
// loaded while this group's positions are gathered; entries at or past m are
// replaced by the particle itself before they are used as an index.
template <int C>
__device__ __forceinline__ void ld_idx(const int *lp, size_t cst, int k, int m, int g, int (&j)[C]) {
#pragma unroll
    for (int c = 0; c < C / 8; c++) {
        if (k + 8 * c < m) {  // this list chunk holds entries
            const int *p = lp + (size_t)(k / 8 + c) * cst;
            const int4 a = ld_list4(p), b = ld_list4(p + 4);
            j[8 * c + 0] = a.x; j[8 * c + 1] = a.y; j[8 * c + 2] = a.z; j[8 * c + 3] = a.w;
            j[8 * c + 4] = b.x; j[8 * c + 5] = b.y; j[8 * c + 6] = b.z; j[8 * c + 7] = b.w;
        } else {
#pragma unroll
            for (int u = 0; u < 8; u++) j[8 * c + u] = g;
        }
    }
}
template <int C, class PosT, class Acc>
__device__ __forceinline__ void walk_list1(const RepC &R, const int *lp, int m, int g, const PosT *__restrict__ pos,
                                           const PosT &pi, int ti, Acc &acc) {
    const size_t cst = (size_t)R.n * 8;
    if (m <= 0) return;
    int j[C];
    PosT q[C];
    ld_idx<C>(lp, cst, 0, m, g, j);
    for (int k = 0; k < m; k += C) {
#pragma unroll
        for (int u = 0; u < C; u++) q[u] = pos[k + u < m ? j[u] : g];
        if (k + C < m) ld_idx<C>(lp, cst, k + C, m, g, j);
#pragma unroll
        for (int u = 0; u < C; u++)
            if (k + u < m) acc.add(pi, ti, q[u], R);
        acc.flush();
    }
}

// Pair forces: a team of tpp lanes per particle (tpp is fixed per handle at
// creation, so a replica's results do not depend on what else is loaded) walks the particle's
// list with stride tpp, 8 independent index/position loads in flight per lane;
// the team's partial forces are combined by a fixed butterfly.  Each block
// writes its partial sums of p^2, PE, virial and p_x; k_force_end reduces
// them per replica.
template <bool FP32, bool KICK>
__global__ void __launch_bounds__(BLK) k_force(const RepC *__restrict__ reps,
                                               const int *__restrict__ fblk_rep, const int *__restrict__ fblk_first,
                                               const double4 *__restrict__ pos4, const QPos *__restrict__ posq,
                                               const int *__restrict__ nn,
                                               const int *__restrict__ nl, int cap, const int *__restrict__ act,
                                               const int8_t *__restrict__ kx, const int8_t *__restrict__ ky,
                                               const int8_t *__restrict__ kz, double *fx, double *fy, double *fz,
                                               double *px, double *py, double *pz, double *partial, int nb) {
    __shared__ double s_red[4][BLK / 64];
    const int b = xcd_block(nb);
    if (b < 0) return;
    const int r = fblk_rep[b];
    const RepC R = reps[r];
    const int tpp = R.tpp, t = threadIdx.x & (tpp - 1);
    const int i = fblk_first[b] + threadIdx.x / tpp;
    double v_p2 = 0.0, v_pe = 0.0, v_vir = 0.0, v_px = 0.0;
    double Fx = 0.0, Fy = 0.0, Fz = 0.0;
    const bool valid = i < R.n;
    const int g = R.off + (valid ? i : 0);
    // FP64: the wrapped box fractions; FP32: their fixed-point copy
    using PosT = typename std::conditional<FP32, QPos, double4>::type;
    const PosT *__restrict__ pos = (const PosT *)(FP32 ? (const void *)posq : (const void *)pos4);
    if (valid) {
        const PosT pi = pos[g];
        const int ti = ptype(pi);
        PairAcc<FP32> acc(ti, R);
        const int m = nn[g];
        if (R.tpp == 1) {
            walk_list1<8>(R, nl + nl_at(R, cap, 0, i), m, g, pos, pi, ti, acc);
            acc.result(Fx, Fy, Fz, R);
            v_pe = acc.pe;
            v_vir = acc.vir;
        } else {
        // entry t + c*tpp of particle i: chunk c of the replica's region
        const int *lp = nl + nl_at(R, cap, t, i);
        const size_t cstep = (size_t)R.n << R.ltpp;
        int k = t;
        // groups of 8 entries; the indices of the next group are loaded
        // while the positions of the current one are in flight
        int jn[8];
        if (k + 7 * tpp < m) {
#pragma unroll
            for (int u = 0; u < 8; u++) jn[u] = lp[u * cstep];
        }
        while (k + 7 * tpp < m) {
            PosT q[8];
#pragma unroll
            for (int u = 0; u < 8; u++) q[u] = pos[jn[u]];
            k += 8 * tpp;
            lp += 8 * cstep;
            if (k + 7 * tpp < m) {
#pragma unroll
                for (int u = 0; u < 8; u++) jn[u] = lp[u * cstep];
            }
#pragma unroll
            for (int u = 0; u < 8; u++) acc.add(pi, ti, q[u], R);
            acc.flush();
        }
        if (k < m) {  // the last, partial group: its loads issued together, the adds in list order
            int jt[8];
#pragma unroll
            for (int u = 0; u < 8; u++) jt[u] = k + u * tpp < m ? lp[u * cstep] : g;
            PosT q[8];
#pragma unroll
            for (int u = 0; u < 8; u++) q[u] = pos[jt[u]];
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (k + u * tpp < m) acc.add(pi, ti, q[u], R);
        }
        acc.result(Fx, Fy, Fz, R);
        v_pe = acc.pe;
        v_vir = acc.vir;
        }
    }
    for (int o = 1; o < tpp; o <<= 1) {  // team butterfly (teams are aligned lane groups)
        Fx += __shfl_xor(Fx, o, 64);
        Fy += __shfl_xor(Fy, o, 64);
        Fz += __shfl_xor(Fz, o, 64);
    }
    if (valid && t == 0) {
        const int a = act[g];
        if (a != -1) {  // cell_list+forces.c:207-213
            Fx += R.f0 * (double)kx[R.off_act + a];
            Fy += R.f0 * (double)ky[R.off_act + a];
            Fz += R.f0 * (double)kz[R.off_act + a];
        }
        fx[g] = Fx; fy[g] = Fy; fz[g] = Fz;
        if (KICK) {  // U2 after the forces, time_evolve.c:516
            double vx = px[g] + Fx * K_HALF_DT, vy = py[g] + Fy * K_HALF_DT, vz = pz[g] + Fz * K_HALF_DT;
            px[g] = vx; py[g] = vy; pz[g] = vz;
            v_p2 = vx * vx + vy * vy + vz * vz;
            v_px = vx;
        }
    }
    double v[4] = {v_p2, v_pe, v_vir, v_px};
    for (int q = 0; q < 4; q++) {
        double w = ac4_wave_sum(v[q]);
        if ((threadIdx.x & 63) == 0) s_red[q][threadIdx.x >> 6] = w;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < 4; q++) {
            double s = 0.0;
            for (int w = 0; w < BLK / 64; w++) s += s_red[q][w];
            partial[4 * (size_t)b + q] = s;
        }
    }
}

// End of the force pass, one workgroup per replica after k_force: the block
// partials summed in a fixed order (thread t of nt takes blocks t, t+nt, ...;
// a fixed butterfly per wave; the waves in order), then the end-of-step chain
// update.  nt = 64, or 256 where a replica has more than 64 force blocks
// (fixed per handle at creation).  A kernel of its own
// rather than the last force block of each replica: that pattern needs an
// agent-scope release fence in every block, which on gfx950 writes back and
// invalidates the L2 of the block's XCD (buffer_wbl2 / buffer_inv), and the
// invalidations cost the neighbour gathers of every other block their L2
// hits (k_force 55.6 -> 42.3 us on the 64-replica ensemble without them).
template <bool KICK>
__global__ void __launch_bounds__(256) k_force_end(const RepC *__restrict__ reps, ChainS *chain,
                                                   const double *__restrict__ partial) {
    __shared__ double s_w[4][4];
    const int r = blockIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6, nt = blockDim.x;
    const RepC R = reps[r];
    double a4[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = threadIdx.x; k < R.fnblk; k += nt) {  // thread t: blocks t, t+nt, ...
        const double *pp = partial + 4 * (size_t)(R.fblk_off + k);
        a4[0] += pp[0]; a4[1] += pp[1]; a4[2] += pp[2]; a4[3] += pp[3];
    }
    for (int q = 0; q < 4; q++)
        for (int o = 1; o < 64; o <<= 1) a4[q] += __shfl_xor(a4[q], o, 64);
    if (lane == 0)
        for (int q = 0; q < 4; q++) s_w[w][q] = a4[q];
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int q = 0; q < 4; q++) {
        a4[q] = s_w[0][q];
        for (int k = 1; k < nt / 64; k++) a4[q] += s_w[k][q];
    }
    ChainS &c = chain[r];
    if (R.n == 0) return;  // idle slot
    double S = a4[0], pe = a4[1], vir = a4[2], spx = a4[3];
    pe = 0.5 * pe;                  // every pair seen from both sides
    vir = (0.5 * vir) / (3.0e0);   // cell_list+forces.c:220
    c.pe = pe;
    c.vir = vir;
    if (!KICK) return;
    const bool rebuilt = need_rebuild(R, c);
    if (c.mode == 2) {
        // NVE: plain velocity Verlet (AdvanceTimeNVE, MDcode/time_evolve.c:39-171,
        // same U2-U1-F-U2 order as the thermostatted step), Heff = KE + PE
        c.p_scale = 1.0;
        c.S_end = S;
        c.ke = S * 0.5;
        c.inst_T = 2.0 * c.ke / R.dof;
        c.drift = spx;
        c.heff = c.ke + pe;
        if (rebuilt) c.rebuilds += 1;
        c.maxbits = 0ull;
        c.force_rebuild = 0;
        if (R.tau > 0 && c.step % R.tau == 0) c.shuffles += 1;  // k_shuffle ran this step
        c.step += 1;
        return;
    }
    if (c.mode == 1) {
        // coldstart/breakxal of MDcode/initialise.c:166-201: momenta rescaled
        // to a ramped target temperature every step, no thermostat chain
        const double f = c.ramp_n > 1 ? c.ramp0 + (c.ramp1 - c.ramp0) * (double)c.ramp_k / (double)(c.ramp_n - 1)
                                      : c.ramp1;
        const double s2 = S > 0.0 ? sqrt(R.T * f * R.dof / S) : 1.0;
        c.p_scale = s2;
        c.S_end = S * s2 * s2;
        c.ke = c.S_end * 0.5;
        c.inst_T = 2.0 * c.ke / R.dof;
        c.drift = spx * s2;
        c.heff = c.ke + pe;
        if (++c.ramp_k >= c.ramp_n) c.mode = 0;
        if (rebuilt) c.rebuilds += 1;
        c.maxbits = 0ull;
        c.force_rebuild = 0;
        return;
    }
    double eta[3] = {c.eta[0], c.eta[1], c.eta[2]}, peta[3] = {c.peta[0], c.peta[1], c.peta[2]}, S2;
    nhc_half(c.S_end, R.T, R.dof, c.q, eta, peta, &S2);   // start of step (same values as k_first_half)
    double s2 = nhc_half(S, R.T, R.dof, c.q, eta, peta, &S2);  // end of step, time_evolve.c:517-519
    for (int k = 0; k < 3; k++) { c.eta[k] = eta[k]; c.peta[k] = peta[k]; }
    c.S_end = S2;
    c.p_scale = s2;
    c.ke = S2 * 0.5;                 // time_evolve.c:520
    c.inst_T = 2.0 * c.ke / R.dof;   // :524
    c.drift = spx * s2;
    double heff = c.ke + pe + eta[0] * R.T * R.dof;  // print.c:85-87
    for (int k = 0; k < 3; k++) heff = heff + peta[k] * peta[k] / (2.0 * c.q[k]);
    for (int k = 1; k < 3; k++) heff = heff + R.T * eta[k];
    c.heff = heff;
    if (rebuilt) c.rebuilds += 1;
    c.maxbits = 0ull;
    c.force_rebuild = 0;
    if (R.tau > 0 && c.step % R.tau == 0) c.shuffles += 1;  // k_shuffle ran this step
    c.step += 1;
}

// ------------------------------------------------------------- host side
struct ac4_md {
    int R = 0, prec = 0, ntot = 0, nblk = 0, ncell_tot = 0, nact_tot = 0, cap = 192, stride = 0, max_ncell = 0;
    int max_fnblk = 0;  // force blocks of the largest slot
    std::vector<RepC> reps_h;
    RepC *reps = nullptr;
    ChainS *chain = nullptr;
    int *blk_rep = nullptr, *blk_first = nullptr, *blk_idx = nullptr, *fblk_rep = nullptr, *fblk_first = nullptr;
    int *cblk_rep = nullptr, *cblk_cell = nullptr;
    int fnblk = 0, lnblk = 0, ncb = 0;
    int shuf_kcap = 0;     // largest Nact whose directions k_shuffle keeps in LDS
    int shuf_bmax = 0;     // swap blocks of the largest slot, 0 without the parallel shuffle path
    int2 *shuf_ent = nullptr;  // its per-block source maps
    int *shuf_ent_n = nullptr;
    size_t shuf_lds = 0;   // dynamic LDS of k_shuffle
    bool shuf_fused = false;  // the shuffle runs in k_first_half's first workgroups
    int *wl_item = nullptr, *wl_n = nullptr;  // rebuild work list (WorkList)
    int2 *ranges = nullptr;                    // per cell: stencil slot ranges (NRANGE)
    int *xmid = nullptr;                       // per cell: first slot of its upper x-half
    int *lblk_rep = nullptr, *lblk_first = nullptr;  // list-build block table (LPB particles per block)
    double *partial = nullptr;
    PartArrays P{}, A{};  // primary and permutation scratch
    void **ftab = nullptr;  // their pointers on the device (PART_FIELDS of P, then of A)
    double *fx = nullptr, *fy = nullptr, *fz = nullptr, *xo = nullptr, *yo = nullptr, *zo = nullptr;
    double4 *pos4 = nullptr;
    QPos *posq = nullptr;
    int2 *cell_mem = nullptr;  // per cell: CELL_CAP (slot, key) of its members in binning order
    int *nn = nullptr, *nn_ref = nullptr, *nl = nullptr, *cnt = nullptr, *start = nullptr, *members = nullptr;
    int8_t *kx = nullptr, *ky = nullptr, *kz = nullptr;
    double *ex = nullptr, *ey = nullptr, *ez = nullptr;  // reference-order export
    // in-situ overlap accumulators
    OvCfg ov{0, 0, 0, 0, 0.0};
    double *ox = nullptr, *oy = nullptr, *oz = nullptr, *ov_q = nullptr, *ov_q2 = nullptr;
    long long *ov_step = nullptr;
    // block tables of every slot, in slot order, with each slot's first
    // entry: the launch grids cover only the occupied slots (refresh_tables)
    std::vector<int> t_brep, t_bfirst, t_bidx, t_frep, t_ffirst, t_lrep, t_lfirst, t_b0, t_f0, t_l0;
    int *ov_list = nullptr;     // replica list of ac4_md_overlap_chi4
    double *ov_out = nullptr;   // its per-lag results
    // frame text on the device (ac4_md_format_frame): 80-byte line slots in
    // reference order, line lengths, the joined text, per-block byte offsets
    // (the total after the last block) and a flag
    char *fmt_dev = nullptr, *fmt_txt = nullptr;
    unsigned char *fmt_len_dev = nullptr;
    long long *fmt_boff = nullptr;
    int *fmt_bad = nullptr;
    size_t fmt_cap = 0;
    unsigned int *ov_cnt = nullptr;
    int *ov_n = nullptr;
    int *et = nullptr;
    hipStream_t stream = nullptr;
    // captured steps: [0] one step, [1] GRAPH_CHUNK steps in one graph
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    hipGraph_t graph[2] = {nullptr, nullptr};
    bool have_forces = false;
    int nlist_mode = 0;
};

#define TRY(e) do { int rc_ = (e); if (rc_ != AC4_OK) return rc_; } while (0)

template <class T>
static int dalloc(T **p, size_t n) {
    if (n == 0) n = 1;
    AC4_HIP(hipMalloc((void **)p, n * sizeof(T)));
    AC4_HIP(hipMemset(*p, 0, n * sizeof(T)));
    return AC4_OK;
}

static int alloc_parts(PartArrays &a, size_t n) {
    int rc;
    double **d[] = {&a.x, &a.y, &a.z, &a.xu, &a.yu, &a.zu, &a.px, &a.py, &a.pz};
    for (double **q : d)
        if ((rc = dalloc(q, n)) != AC4_OK) return rc;
    int **i[] = {&a.type, &a.act, &a.id, &a.cell_ref, &a.cell_of};
    for (int **q : i)
        if ((rc = dalloc(q, n)) != AC4_OK) return rc;
    return AC4_OK;
}
static void free_parts(PartArrays &a) {
    void *q[] = {a.x, a.y, a.z, a.xu, a.yu, a.zu, a.px, a.py, a.pz, a.type, a.act, a.id, a.cell_ref, a.cell_of};
    for (void *v : q)
        if (v) (void)hipFree(v);
    a = PartArrays{};
}

// Kernel shape of handles created afterwards.  tpp: lanes per particle in
// the force kernel, 1..16, or 0 = automatic (the widest team with at most
// TPP_LANES lanes over the handle's total capacity, so small handles still
// fill the device and large ones keep one thread per particle).  The team
// width fixes the order of each particle's force sum, so results are bitwise
// reproducible for a given handle shape.  nlist_mode: dynamics list builder,
// 0 = one wave per cell, 1 = one thread per particle, 2/3/4 = teams of
// 16/32/64 lanes per particle.
#define TPP_LANES 65536.0
static int g_tpp = 0, g_nlist_mode = 3;
static double g_skin = K_SKIN;

// Verlet skin of handles created afterwards (0.2 ... 0.6; the reference's is
// 0.35).  A larger skin rebuilds less often and gives longer lists.
extern "C" int ac4_set_skin(double skin) {
    AC4_REQUIRE(skin >= 0.2 && skin <= 0.6, AC4_EARG, "ac4_set_skin: skin must be in [0.2, 0.6]");
    g_skin = skin;
    return AC4_OK;
}

extern "C" int ac4_set_tuning(int tpp, int nlist_mode) {
    AC4_REQUIRE((tpp == 0 || tpp == 1 || tpp == 2 || tpp == 4 || tpp == 8 || tpp == 16) &&
                    (nlist_mode >= 0 && nlist_mode <= 4),
                AC4_EARG, "ac4_set_tuning: tpp in {0,1,2,4,8,16}, nlist_mode in {0,...,4}");
    g_tpp = tpp;
    g_nlist_mode = nlist_mode;
    return AC4_OK;
}

// Job-dependent fields of a slot (the capacity-dependent ones are fixed at
// creation).  Returns 0 or an error message.
static const char *configure_rep(RepC &R, const ac4_replica_params &q) {
    if (q.n < 2 || !(q.density > 0) || !(q.T > 0)) return "bad n/density/T";
    if (q.n > R.cap_n) return "n exceeds the slot capacity";
    const int nact = q.n_act < 0 ? (int)((double)q.n * 0.1000) : q.n_act;
    if (nact > q.n) return "n_act larger than n";
    const double L = pow((double)q.n / q.density, 0.33333333);  // MDcode/main.c:465
    const double cell_cut = std::max(K_CELL_CUT, 2.5 + R.skin);  // a cell spans the largest list radius
    if (L <= 2.0 * cell_cut) return "box too small for the list cutoff";
    R.n = q.n;
    R.nact = nact;
    R.L = L;
    R.invL = 1.0 / R.L;  // :466
    // MDcode/Parallel.c:72-80 with one process (Lpp = L): the reference grid
    const double Lpp = R.L / (double)1;
    R.nc_ref = (int)((double)((int)Lpp) / K_CELL_CUT);
    if (R.nc_ref < 1) R.nc_ref = 1;
    R.inv_cell_ref = R.L / (Lpp / (double)R.nc_ref);
    R.nc = (int)(R.L / cell_cut);  // list grid: cells at least one list radius (and 2.85) wide
    // A particle in the upper (lower) x-half of its cell is at least one cell
    // width from the lower (upper) x-half of the cell below (above) it along
    // x, for any grid of three or more cells: beyond every list radius and the
    // screen's margin once the width exceeds the largest radius by 0.01 sigma.
    R.xcull = R.nc >= 3 && R.L / R.nc >= 2.5 + R.skin + 0.01;
    R.ncell = R.nc * R.nc * R.nc;
    if (R.ncell > R.cap_cells) return "density below the slot's cell capacity (0.5)";
    R.qscale = R.L / 4294967296.0;
    R.qs2 = (float)(R.qscale * R.qscale);
    R.tau = q.tau_per;
    R.T = q.T;
    R.f0 = q.f0;
    R.density = q.density;
    R.dof = 3.0 * (double)q.n - 3.0;  // global.h:35
    R.key0 = (uint32_t)q.seed;
    R.key1 = (uint32_t)(q.seed >> 32);
    return nullptr;
}

static void fresh_chain(ChainS &c, const RepC &R) {
    memset(&c, 0, sizeof(ChainS));
    const double tau = .20, T = R.T;  // time_evolve.c:477-482
    for (int k = 0; k < 3; k++) c.q[k] = T * tau * tau;
    c.q[0] = R.dof * T * tau * tau;
    c.p_scale = 1.0;
    c.force_rebuild = R.n > 0 ? 1 : 0;
}

// cap_rule > 0: the automatic force team width is taken for that total
// capacity instead of the handle's own (a server's launch groups share one
// kernel shape, so a job's arithmetic does not depend on its group)
extern "C" int ac4_md_create_slots_rule(int nslots, const int *cap_n, int precision, double cap_rule,
                                        ac4_md **out);
extern "C" int ac4_md_create_slots(int nslots, const int *cap_n, int precision, ac4_md **out) {
    return ac4_md_create_slots_rule(nslots, cap_n, precision, 0.0, out);
}
extern "C" int ac4_md_create_slots_rule(int nslots, const int *cap_n, int precision, double cap_rule,
                                        ac4_md **out) {
    AC4_REQUIRE(nslots >= 1 && cap_n && out, AC4_EARG, "ac4_md_create_slots: bad arguments");
    AC4_REQUIRE(precision == AC4_FP64 || precision == AC4_FP32, AC4_EARG, "precision must be 0 or 1");
    int ndev = 0;
    AC4_REQUIRE(hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0, AC4_EHIP,
                "ac4_md_create: no HIP device visible (libac4 has no CPU path)");
    ac4_md *md = new ac4_md();
    md->R = nslots;
    md->prec = precision;
    md->nlist_mode = g_nlist_mode;
    int off = 0, off_act = 0, off_cell = 0, blk_off = 0, fblk_off = 0;
    std::vector<int> brep, bfirst, bidx, frep, ffirst, crep, ccell, lrep, lfirst;
    int ncb = 0;
    AC4_REQUIRE(nslots < 32768, AC4_EARG, "at most 32767 slots");
    double cap_tot = 0.0;
    for (int r = 0; r < nslots; r++) cap_tot += cap_n[r] > 0 ? cap_n[r] : 0;
    int tpp = g_tpp, ltpp = 0;
    if (cap_rule > 0.0) cap_tot = cap_rule;
    if (tpp == 0) {
        tpp = 1;
        while (tpp < 16 && cap_tot * tpp * 2 <= TPP_LANES) tpp *= 2;
    }
    while ((1 << ltpp) < tpp) ltpp++;
    for (int r = 0; r < nslots; r++) {
        RepC R;
        memset(&R, 0, sizeof R);
        const int cn = cap_n[r];
        if (cn < 2) { delete md; ac4_set_error("slot %d: capacity %d < 2", r, cn); return AC4_EARG; }
        // list addressing in the team builders uses 24-bit products of n*8
        if (cn > (1 << 21)) { delete md; ac4_set_error("slot %d: capacity %d > 2^21", r, cn); return AC4_EARG; }
        R.cap_n = cn;
        R.skin = g_skin;
        R.off = off;
        R.off_act = off_act;
        R.off_cell = off_cell;
        // the densest box a slot can hold sets its cell capacity (rho >= 0.5)
        const int cap_cells = (int)pow((int)(pow((double)cn / 0.5, 1.0 / 3.0) / K_CELL_CUT) + 1, 3);
        R.cap_cells = cap_cells;
        R.tpp = tpp;
        R.ltpp = ltpp;
        R.lc = tpp == 1 ? 3 : ltpp;
        R.nblk = (cn + BLK - 1) / BLK;
        R.blk_off = blk_off;
        for (int k = 0; k < R.nblk; k++) { brep.push_back(r); bfirst.push_back(k * BLK); bidx.push_back(k); }
        const int ppb = BLK / R.tpp;  // particles per force block
        R.fnblk = (cn + ppb - 1) / ppb;
        R.fblk_off = fblk_off;
        for (int k = 0; k < R.fnblk; k++) { frep.push_back(r); ffirst.push_back(k * ppb); }
        R.lblk_off = (int)lrep.size();
        R.lnblk = (cn + LPB - 1) / LPB;
        for (int k = 0; k < R.lnblk; k++) { lrep.push_back(r); lfirst.push_back(k * LPB); }
        if ((cap_cells + BLK / 64 - 1) / (BLK / 64) >= 65536) { delete md; ac4_set_error("slot %d: too many cells", r); return AC4_EARG; }
        ncb += (cap_cells + BLK / 64 - 1) / (BLK / 64);
        for (int c = 0; c < cap_cells; c++) { crep.push_back(r); ccell.push_back(c); }
        off += cn;
        off_act += cn;  // room for n_act up to n
        off_cell += cap_cells;
        blk_off += R.nblk;
        fblk_off += R.fnblk;
        md->max_ncell = std::max(md->max_ncell, cap_cells);
        md->max_fnblk = std::max(md->max_fnblk, R.fnblk);
        md->reps_h.push_back(R);  // idle: n = 0
    }
    md->t_brep = brep; md->t_bfirst = bfirst; md->t_bidx = bidx;
    md->t_frep = frep; md->t_ffirst = ffirst; md->t_lrep = lrep; md->t_lfirst = lfirst;
    for (const RepC &R : md->reps_h) {
        md->t_b0.push_back(R.blk_off);
        md->t_f0.push_back(R.fblk_off);
        md->t_l0.push_back(R.lblk_off);
    }
    md->ntot = off;
    md->nact_tot = off_act;
    md->ncell_tot = off_cell;
    md->nblk = blk_off;
    md->stride = off;
    md->fnblk = fblk_off;
    md->lnblk = (int)lrep.size();
    md->ncb = ncb;
    {  // k_shuffle: the partner chunk plus one component in LDS when it fits
        int dev = 0, lds = 65536, cap_max = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) lds = 65536;
        for (int r = 0; r < nslots; r++) cap_max = std::max(cap_max, cap_n[r]);
        const int room = lds - (int)(SHUF_CH * sizeof(unsigned int));
        md->shuf_kcap = std::max(0, std::min(room, (cap_max + 15) & ~15));
        md->shuf_lds = SHUF_CH * sizeof(unsigned int) + (size_t)md->shuf_kcap;
        if (md->shuf_lds > 65536 &&
            (hipFuncSetAttribute((const void *)k_shuffle, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)md->shuf_lds) != hipSuccess ||
             hipFuncSetAttribute((const void *)k_shuf_apply, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)md->shuf_lds) != hipSuccess)) {
            (void)hipGetLastError();
            md->shuf_kcap = 65536 - (int)(SHUF_CH * sizeof(unsigned int));
            md->shuf_lds = 65536;
        }
        md->shuf_bmax = cap_max > SHUF_PAR_MIN ? (cap_max + SHUF_M - 1) / SHUF_M : 0;
        // in k_first_half when its LDS leaves that kernel at least 5 workgroups per CU
        md->shuf_fused = md->shuf_bmax == 0 && md->shuf_lds <= 32768;
    }
    int rc = AC4_OK;
#define ALLOC(ptr, n) if ((rc = dalloc(&md->ptr, (size_t)(n))) != AC4_OK) { ac4_md_destroy(md); return rc; }
    if (md->shuf_bmax > 0) {
        ALLOC(shuf_ent, (size_t)nslots * 3 * md->shuf_bmax * 2 * SHUF_M)
        ALLOC(shuf_ent_n, (size_t)nslots * 3 * md->shuf_bmax)
    }
    ALLOC(reps, nslots) ALLOC(chain, nslots) ALLOC(blk_rep, md->nblk) ALLOC(blk_first, md->nblk) ALLOC(blk_idx, md->nblk)
    ALLOC(fblk_rep, md->fnblk) ALLOC(fblk_first, md->fnblk) ALLOC(lblk_rep, md->lnblk) ALLOC(lblk_first, md->lnblk)
    ALLOC(wl_item, (size_t)md->nblk + md->lnblk + md->ncb) ALLOC(wl_n, 4) ALLOC(ranges, (size_t)off_cell * NRANGE) ALLOC(xmid, off_cell) ALLOC(cblk_rep, off_cell) ALLOC(cblk_cell, off_cell) ALLOC(partial, 4 * (size_t)md->fnblk)
    ALLOC(fx, off) ALLOC(fy, off) ALLOC(fz, off) ALLOC(xo, off) ALLOC(yo, off) ALLOC(zo, off) ALLOC(pos4, off) ALLOC(posq, off)
    ALLOC(cell_mem, (size_t)off_cell * CELL_CAP) ALLOC(nn, off) ALLOC(nn_ref, off) ALLOC(nl, (size_t)md->cap * off) ALLOC(cnt, off_cell) ALLOC(start, off_cell)
    ALLOC(members, off) ALLOC(kx, off_act) ALLOC(ky, off_act) ALLOC(kz, off_act)
#undef ALLOC
    if ((rc = alloc_parts(md->P, off)) != AC4_OK || (rc = alloc_parts(md->A, off)) != AC4_OK) { ac4_md_destroy(md); return rc; }
    {
        void *tab[2 * PART_FIELDS];
        for (int k = 0; k < 2; k++) {
            const PartArrays &a = k ? md->A : md->P;
            void *f[PART_FIELDS] = {a.x, a.y, a.z, a.xu, a.yu, a.zu, a.px, a.py, a.pz,
                                    a.type, a.act, a.id, a.cell_ref, a.cell_of};
            for (int q = 0; q < PART_FIELDS; q++) tab[k * PART_FIELDS + q] = f[q];
        }
        if ((rc = dalloc(&md->ftab, 2 * PART_FIELDS)) != AC4_OK) { ac4_md_destroy(md); return rc; }
        if (hipMemcpy(md->ftab, tab, sizeof tab, hipMemcpyHostToDevice) != hipSuccess) {
            ac4_md_destroy(md);
            ac4_set_error("ac4_md_create: field table upload failed");
            return AC4_EHIP;
        }
    }
    if (hipStreamCreate(&md->stream) != hipSuccess) {
        ac4_md_destroy(md);
        ac4_set_error("hipStreamCreate failed");
        return AC4_EHIP;
    }
    AC4_HIP(hipMemcpy(md->reps, md->reps_h.data(), sizeof(RepC) * nslots, hipMemcpyHostToDevice));
    AC4_HIP(hipMemcpy(md->blk_rep, brep.data(), sizeof(int) * md->nblk, hipMemcpyHostToDevice));
    AC4_HIP(hipMemcpy(md->blk_first, bfirst.data(), sizeof(int) * md->nblk, hipMemcpyHostToDevice));
    AC4_HIP(hipMemcpy(md->blk_idx, bidx.data(), sizeof(int) * md->nblk, hipMemcpyHostToDevice));
    AC4_HIP(hipMemcpy(md->fblk_rep, frep.data(), sizeof(int) * md->fnblk, hipMemcpyHostToDevice));
    AC4_HIP(hipMemcpy(md->fblk_first, ffirst.data(), sizeof(int) * md->fnblk, hipMemcpyHostToDevice));
    AC4_HIP(hipMemcpy(md->lblk_rep, lrep.data(), sizeof(int) * md->lnblk, hipMemcpyHostToDevice));
    AC4_HIP(hipMemcpy(md->lblk_first, lfirst.data(), sizeof(int) * md->lnblk, hipMemcpyHostToDevice));
    AC4_HIP(hipMemcpy(md->cblk_rep, crep.data(), sizeof(int) * off_cell, hipMemcpyHostToDevice));
    AC4_HIP(hipMemcpy(md->cblk_cell, ccell.data(), sizeof(int) * off_cell, hipMemcpyHostToDevice));
    std::vector<ChainS> ch(nslots);
    for (int r = 0; r < nslots; r++) fresh_chain(ch[r], md->reps_h[r]);
    AC4_HIP(hipMemcpy(md->chain, ch.data(), sizeof(ChainS) * nslots, hipMemcpyHostToDevice));
    std::vector<int> neg(off, -1), ids(off);
    for (int r = 0; r < nslots; r++)
        for (int i = 0; i < md->reps_h[r].cap_n; i++) ids[md->reps_h[r].off + i] = i;
    AC4_HIP(hipMemcpy(md->P.act, neg.data(), sizeof(int) * off, hipMemcpyHostToDevice));
    AC4_HIP(hipMemcpy(md->P.id, ids.data(), sizeof(int) * off, hipMemcpyHostToDevice));
    md->have_forces = true;  // nothing loaded yet; set_state asks for a force evaluation
    *out = md;
    return AC4_OK;
}

template <class T>
static int h2d(T *dst, const T *src, size_t n, hipStream_t s);

static void drop_graphs(ac4_md *md) {
    for (int k = 0; k < 2; k++) {
        if (md->gexec[k]) { (void)hipGraphExecDestroy(md->gexec[k]); md->gexec[k] = nullptr; }
        if (md->graph[k]) { (void)hipGraphDestroy(md->graph[k]); md->graph[k] = nullptr; }
    }
}

// Block tables of the occupied slots only (all slots while none is
// occupied), in slot order: the step's launch grids skip idle slots, whose
// blocks would only exit.  A replica's blocks, and so its arithmetic (every
// per-replica sum runs over its own blocks in their order), do not change.
// The captured step graphs are dropped: their grids change.
static int refresh_tables(ac4_md *md) {
    bool any = false;
    for (const RepC &R : md->reps_h) any |= R.n > 0;
    std::vector<int> br, bf, bi, fr, ff, lr, lf;
    for (int r = 0; r < md->R; r++) {
        RepC &R = md->reps_h[r];
        if (any && R.n == 0) continue;
        R.blk_off = (int)br.size();
        br.insert(br.end(), md->t_brep.begin() + md->t_b0[r], md->t_brep.begin() + md->t_b0[r] + R.nblk);
        bf.insert(bf.end(), md->t_bfirst.begin() + md->t_b0[r], md->t_bfirst.begin() + md->t_b0[r] + R.nblk);
        bi.insert(bi.end(), md->t_bidx.begin() + md->t_b0[r], md->t_bidx.begin() + md->t_b0[r] + R.nblk);
        R.fblk_off = (int)fr.size();
        fr.insert(fr.end(), md->t_frep.begin() + md->t_f0[r], md->t_frep.begin() + md->t_f0[r] + R.fnblk);
        ff.insert(ff.end(), md->t_ffirst.begin() + md->t_f0[r], md->t_ffirst.begin() + md->t_f0[r] + R.fnblk);
        R.lblk_off = (int)lr.size();
        lr.insert(lr.end(), md->t_lrep.begin() + md->t_l0[r], md->t_lrep.begin() + md->t_l0[r] + R.lnblk);
        lf.insert(lf.end(), md->t_lfirst.begin() + md->t_l0[r], md->t_lfirst.begin() + md->t_l0[r] + R.lnblk);
    }
    hipStream_t s = md->stream;
    AC4_HIP(hipStreamSynchronize(s));
    drop_graphs(md);
    TRY(h2d(md->blk_rep, br.data(), br.size(), s)); TRY(h2d(md->blk_first, bf.data(), bf.size(), s));
    TRY(h2d(md->blk_idx, bi.data(), bi.size(), s));
    TRY(h2d(md->fblk_rep, fr.data(), fr.size(), s)); TRY(h2d(md->fblk_first, ff.data(), ff.size(), s));
    TRY(h2d(md->lblk_rep, lr.data(), lr.size(), s)); TRY(h2d(md->lblk_first, lf.data(), lf.size(), s));
    TRY(h2d(md->reps, md->reps_h.data(), md->reps_h.size(), s));
    AC4_HIP(hipStreamSynchronize(s));
    md->nblk = (int)br.size();
    md->fnblk = (int)fr.size();
    md->lnblk = (int)lr.size();
    return AC4_OK;
}

extern "C" int ac4_md_assign(ac4_md *md, int slot, const ac4_replica_params *p) {
    AC4_REQUIRE(md && slot >= 0 && slot < md->R && p, AC4_EARG, "ac4_md_assign: bad arguments");
    RepC R = md->reps_h[slot];
    const char *why = configure_rep(R, *p);
    AC4_REQUIRE(!why, AC4_EARG, "slot %d: %s", slot, why ? why : "");
    const bool was_idle = md->reps_h[slot].n == 0;
    md->reps_h[slot] = R;
    ChainS c;
    fresh_chain(c, R);
    AC4_HIP(hipMemcpyAsync(md->reps + slot, &md->reps_h[slot], sizeof(RepC), hipMemcpyHostToDevice, md->stream));
    AC4_HIP(hipMemcpyAsync(md->chain + slot, &c, sizeof(ChainS), hipMemcpyHostToDevice, md->stream));
    AC4_HIP(hipStreamSynchronize(md->stream));
    return was_idle ? refresh_tables(md) : AC4_OK;
}

extern "C" int ac4_md_release(ac4_md *md, int slot) {
    AC4_REQUIRE(md && slot >= 0 && slot < md->R, AC4_EARG, "ac4_md_release: bad arguments");
    RepC &R = md->reps_h[slot];
    R.n = 0; R.nact = 0; R.ncell = 0; R.tau = 0; R.nc = 0;
    ChainS c;
    fresh_chain(c, R);
    AC4_HIP(hipMemcpyAsync(md->reps + slot, &R, sizeof(RepC), hipMemcpyHostToDevice, md->stream));
    AC4_HIP(hipMemcpyAsync(md->chain + slot, &c, sizeof(ChainS), hipMemcpyHostToDevice, md->stream));
    AC4_HIP(hipStreamSynchronize(md->stream));
    return refresh_tables(md);
}

extern "C" int ac4_md_slot_n(const ac4_md *md, int slot) {
    if (!md || slot < 0 || slot >= md->R) return AC4_EARG;
    return md->reps_h[slot].n;
}

extern "C" int ac4_md_create(int n_rep, const ac4_replica_params *p, int precision, ac4_md **out) {
    AC4_REQUIRE(n_rep >= 1 && p && out, AC4_EARG, "ac4_md_create: bad arguments");
    std::vector<int> caps(n_rep);
    for (int r = 0; r < n_rep; r++) caps[r] = p[r].n;
    ac4_md *md = nullptr;
    TRY(ac4_md_create_slots(n_rep, caps.data(), precision, &md));
    for (int r = 0; r < n_rep; r++) {
        int rc = ac4_md_assign(md, r, &p[r]);
        if (rc != AC4_OK) { ac4_md_destroy(md); return rc; }
    }
    *out = md;
    return AC4_OK;
}

extern "C" int ac4_md_destroy(ac4_md *md) {
    if (!md) return AC4_OK;
    if (md->stream) (void)hipStreamSynchronize(md->stream);
    drop_graphs(md);
    void *ptrs[] = {md->reps, md->chain, md->blk_rep, md->blk_first, md->blk_idx, md->fblk_rep, md->fblk_first,
                    md->lblk_rep, md->lblk_first, md->wl_item, md->wl_n, md->ranges, md->xmid,
                    md->cblk_rep, md->cblk_cell, md->ox, md->oy, md->oz, md->ov_q, md->ov_q2,
                    md->ov_step, md->ov_cnt, md->ov_n, md->ov_list, md->ov_out,
                    md->partial, md->fx, md->fy, md->fz, md->xo, md->yo, md->zo, md->pos4, md->posq, md->cell_mem, md->shuf_ent, md->shuf_ent_n, md->nn,
                    md->nn_ref, md->nl,
                    md->cnt, md->start, md->members, md->kx, md->ky, md->kz, md->ex, md->ey, md->ez, md->et};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    for (void *q : {(void *)md->fmt_dev, (void *)md->fmt_txt, (void *)md->fmt_len_dev, (void *)md->fmt_boff,
                    (void *)md->fmt_bad})
        if (q) (void)hipFree(q);
    free_parts(md->P);
    free_parts(md->A);
    if (md->ftab) (void)hipFree(md->ftab);
    if (md->stream) (void)hipStreamDestroy(md->stream);
    delete md;
    return AC4_OK;
}

extern "C" int ac4_md_total_particles(const ac4_md *md) { return md ? md->ntot : 0; }

// Device bytes a slot of capacity cap_n adds to a handle, from the sizes
// ac4_md_create_slots and ac4_md_overlap_enable allocate (the reference-order
// export buffers included; the parallel-shuffle tables counted at their
// largest, as if this slot were the largest of its handle).
extern "C" long long ac4_md_slot_bytes(int cap_n, int n_origin, int n_lags) {
    if (cap_n < 2 || n_origin < 0 || n_lags < 0) return AC4_EARG;
    const long long n = cap_n;
    const long long cells = (long long)pow((int)(pow((double)cap_n / 0.5, 1.0 / 3.0) / K_CELL_CUT) + 1, 3);
    const long long nblk = (n + BLK - 1) / BLK, lnblk = (n + LPB - 1) / LPB;
    const long long fnb = (n * 16 + BLK - 1) / BLK;  // force blocks at the widest team (16 lanes a particle)
    const long long ncb = (cells + BLK / 64 - 1) / (BLK / 64);
    long long b = 0;
    b += n * (6 * 8 + 32 + 12 + 2 * 4 + 192LL * 4 + 4 + 3);         // forces, xold, pos copies, counts, list, members, k
    b += 2 * n * (9 * 8 + 5 * 4);                                    // particle arrays and their scratch copy
    b += n * (3 * 8 + 4);                                            // reference-order export
    b += cells * ((long long)NRANGE * 8 + 3 * 4 + (long long)CELL_CAP * 8 + 2 * 4);
    b += nblk * 3 * 4 + fnb * 2 * 4 + lnblk * 2 * 4;                // block tables
    b += (nblk + lnblk + ncb) * 4 + fnb * 4 * 8;                     // work list, force partials
    b += (long long)sizeof(RepC) + (long long)sizeof(ChainS);
    if (cap_n > SHUF_PAR_MIN) b += 3LL * 2 * ((n + SHUF_M - 1) / SHUF_M) * SHUF_M * 4 + 3LL * ((n + SHUF_M - 1) / SHUF_M) * 4;
    if (n_origin > 0) b += n * 24LL * n_origin + (long long)n_origin * 12 + (long long)(n_lags + 1) * 20;
    return b;
}


static int check_rep(const ac4_md *md, int r) {
    AC4_REQUIRE(md && r >= 0 && r < md->R, AC4_EARG, "bad handle or replica index %d", r);
    AC4_REQUIRE(md->reps_h[r].n > 0, AC4_ESTATE, "slot %d is idle (no job assigned)", r);
    return AC4_OK;
}

template <class T>
static int h2d(T *dst, const T *src, size_t n, hipStream_t s) {
    AC4_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, s));
    return AC4_OK;
}
template <class T>
static int d2h(T *dst, const T *src, size_t n, hipStream_t s) {
    AC4_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, s));
    return AC4_OK;
}
static int get_chain(ac4_md *md, int r, ChainS *c) {
    TRY(d2h(c, md->chain + r, 1, md->stream));
    AC4_HIP(hipStreamSynchronize(md->stream));
    return AC4_OK;
}
static int put_chain(ac4_md *md, int r, const ChainS *c) {
    TRY(h2d(md->chain + r, c, 1, md->stream));
    AC4_HIP(hipStreamSynchronize(md->stream));
    return AC4_OK;
}
static int get_chains(ac4_md *md, std::vector<ChainS> &c) {
    c.resize(md->R);
    TRY(d2h(c.data(), md->chain, md->R, md->stream));
    AC4_HIP(hipStreamSynchronize(md->stream));
    return AC4_OK;
}
static int put_chains(ac4_md *md, const std::vector<ChainS> &c) {
    TRY(h2d(md->chain, c.data(), md->R, md->stream));
    AC4_HIP(hipStreamSynchronize(md->stream));
    return AC4_OK;
}

// Sets error bits on replica r as the kernels do when a list or cell
// overflows or a coordinate runs away, without any of those happening: the
// way tests exercise the failure handling of the callers.
extern "C" int ac4_md_raise_error(ac4_md *md, int r, int bits) {
    TRY(check_rep(md, r));
    AC4_REQUIRE(bits > 0 && bits < 4, AC4_EARG, "raise_error: bits must be 1, 2 or 3");
    ChainS c;
    TRY(get_chain(md, r, &c));
    c.err |= bits;
    return put_chain(md, r, &c);
}

// Error bits of every replica (bit 0: neighbour list or cell over capacity,
// bit 1: non-finite or runaway coordinate); err[r] for r < R.
extern "C" int ac4_md_errors(ac4_md *md, int *err) {
    AC4_REQUIRE(md && err, AC4_EARG, "errors: bad arguments");
    std::vector<ChainS> c;
    TRY(get_chains(md, c));
    for (int r = 0; r < md->R; r++) err[r] = md->reps_h[r].n > 0 ? c[r].err : 0;
    return AC4_OK;
}

// slot -> reference index map of replica r
static int get_ids(ac4_md *md, int r, std::vector<int> &id) {
    const RepC &R = md->reps_h[r];
    id.resize(R.n);
    TRY(d2h(id.data(), md->P.id + R.off, R.n, md->stream));
    AC4_HIP(hipStreamSynchronize(md->stream));
    return AC4_OK;
}
template <class T>
static int fetch_unpermuted(ac4_md *md, int r, const T *dev, T *out, const std::vector<int> &id) {
    const RepC &R = md->reps_h[r];
    std::vector<T> tmp(R.n);
    TRY(d2h(tmp.data(), dev + R.off, R.n, md->stream));
    AC4_HIP(hipStreamSynchronize(md->stream));
    for (int i = 0; i < R.n; i++) out[id[i]] = tmp[i];
    return AC4_OK;
}

static int set_state_impl(ac4_md *md, int r, const double *x, const double *y, const double *z, const double *xu,
                          const double *yu, const double *zu, const double *px, const double *py, const double *pz,
                          const int *type, const int *act_tag, const int *kx, const int *ky, const int *kz,
                          bool slot_local);

extern "C" int ac4_md_set_state(ac4_md *md, int r, const double *x, const double *y, const double *z,
                                const double *xu, const double *yu, const double *zu, const double *px,
                                const double *py, const double *pz, const int *type, const int *act_tag,
                                const int *kx, const int *ky, const int *kz) {
    return set_state_impl(md, r, x, y, z, xu, yu, zu, px, py, pz, type, act_tag, kx, ky, kz, false);
}

// Loads a job into a slot without touching the other slots: its forces start
// at zero and its list is built inside its first step (used by the server, so
// admitting a job never changes the running replicas' arithmetic).
extern "C" int ac4_md_load_slot(ac4_md *md, int r, const double *x, const double *y, const double *z,
                                const double *xu, const double *yu, const double *zu, const double *px,
                                const double *py, const double *pz, const int *type, const int *act_tag,
                                const int *kx, const int *ky, const int *kz) {
    return set_state_impl(md, r, x, y, z, xu, yu, zu, px, py, pz, type, act_tag, kx, ky, kz, true);
}

static int set_state_impl(ac4_md *md, int r, const double *x, const double *y, const double *z, const double *xu,
                          const double *yu, const double *zu, const double *px, const double *py, const double *pz,
                          const int *type, const int *act_tag, const int *kx, const int *ky, const int *kz,
                          bool slot_local) {
    TRY(check_rep(md, r));
    AC4_REQUIRE(x && y && z && xu && yu && zu && px && py && pz && type, AC4_EARG, "set_state: NULL array");
    const RepC &R = md->reps_h[r];
    const int n = R.n, g = R.off;
    std::vector<double4> p4(n);
    std::vector<QPos> q4(n);
    std::vector<int> ids(n);
    double keblk = 0.0;  // time_evolve.c:586-589 order
    for (int i = 0; i < n; i++) {
        AC4_REQUIRE(type[i] == 0 || type[i] == 1, AC4_EARG, "set_state: type[%d]=%d not in {0,1}", i, type[i]);
        AC4_REQUIRE(x[i] >= 0.0 && x[i] < 1.0 && y[i] >= 0.0 && y[i] < 1.0 && z[i] >= 0.0 && z[i] < 1.0, AC4_EARG,
                    "set_state: wrapped coordinate of particle %d outside [0,1)", i);
        p4[i] = make_double4(x[i], y[i], z[i], (double)type[i]);
        q4[i] = qpos(x[i], y[i], z[i], type[i]);
        keblk += px[i] * px[i] + py[i] * py[i] + pz[i] * pz[i];
        ids[i] = i;
    }
    hipStream_t s = md->stream;
    // a loaded state is stored in reference order (slot = index)
    TRY(h2d(md->P.x + g, x, n, s)); TRY(h2d(md->P.y + g, y, n, s)); TRY(h2d(md->P.z + g, z, n, s));
    TRY(h2d(md->P.xu + g, xu, n, s)); TRY(h2d(md->P.yu + g, yu, n, s)); TRY(h2d(md->P.zu + g, zu, n, s));
    TRY(h2d(md->xo + g, xu, n, s)); TRY(h2d(md->yo + g, yu, n, s)); TRY(h2d(md->zo + g, zu, n, s));
    TRY(h2d(md->P.px + g, px, n, s)); TRY(h2d(md->P.py + g, py, n, s)); TRY(h2d(md->P.pz + g, pz, n, s));
    TRY(h2d(md->P.type + g, type, n, s));
    TRY(h2d(md->P.id + g, ids.data(), n, s));
    TRY(h2d(md->pos4 + g, p4.data(), n, s));
    TRY(h2d(md->posq + g, q4.data(), n, s));
    std::vector<int> at;
    if (act_tag) {
        for (int i = 0; i < n; i++)
            AC4_REQUIRE(act_tag[i] >= -1 && act_tag[i] < R.nact, AC4_EARG, "set_state: actTag[%d]=%d out of range", i, act_tag[i]);
        at.assign(act_tag, act_tag + n);
    } else {  // keep the current tags, brought back to reference order
        std::vector<int> cur;
        TRY(get_ids(md, r, cur));
        at.resize(n);
        TRY(fetch_unpermuted(md, r, (const int *)md->P.act, at.data(), cur));
    }
    TRY(h2d(md->P.act + g, at.data(), n, s));
    if (kx && ky && kz && R.nact > 0) {
        std::vector<int8_t> a(R.nact), b(R.nact), c(R.nact);
        for (int i = 0; i < R.nact; i++) { a[i] = (int8_t)kx[i]; b[i] = (int8_t)ky[i]; c[i] = (int8_t)kz[i]; }
        TRY(h2d(md->kx + R.off_act, a.data(), R.nact, s));
        TRY(h2d(md->ky + R.off_act, b.data(), R.nact, s));
        TRY(h2d(md->kz + R.off_act, c.data(), R.nact, s));
        AC4_HIP(hipStreamSynchronize(s));
    }
    ChainS c;
    TRY(get_chain(md, r, &c));  // also completes the copies above
    c.S_end = keblk;
    c.p_scale = 1.0;
    c.force_rebuild = 1;
    c.maxbits = 0ull;
    c.step = 0;  // a loaded state starts a new run (MDcode/main.c:238-240: start = 0)
    c.mode = 0;
    c.err = 0;
    TRY(put_chain(md, r, &c));
    if (slot_local) {
        AC4_HIP(hipMemsetAsync(md->fx + g, 0, sizeof(double) * n, s));
        AC4_HIP(hipMemsetAsync(md->fy + g, 0, sizeof(double) * n, s));
        AC4_HIP(hipMemsetAsync(md->fz + g, 0, sizeof(double) * n, s));
        AC4_HIP(hipStreamSynchronize(s));
    } else {
        md->have_forces = false;
    }
    return AC4_OK;
}

// NVE (on != 0) or back to the Nose-Hoover chain (on == 0) for replica r; the
// chain variables are kept and resume where they were.
extern "C" int ac4_md_set_nve(ac4_md *md, int r, int on) {
    TRY(check_rep(md, r));
    ChainS c;
    TRY(get_chain(md, r, &c));
    AC4_REQUIRE(c.mode != 1, AC4_ESTATE, "set_nve: replica %d is in a relaxation ramp", r);
    c.mode = on ? 2 : 0;
    return put_chain(md, r, &c);
}

// Velocity-rescale ramp for one slot only, applied by its next nsteps steps.
extern "C" int ac4_md_relax_slot(ac4_md *md, int r, int nsteps, double f_start, double f_end) {
    TRY(check_rep(md, r));
    AC4_REQUIRE(nsteps > 0 && f_start > 0 && f_end > 0, AC4_EARG, "relax_slot: bad arguments");
    ChainS c;
    TRY(get_chain(md, r, &c));
    c.mode = 1; c.ramp_n = nsteps; c.ramp_k = 0; c.ramp0 = f_start; c.ramp1 = f_end;
    return put_chain(md, r, &c);
}

extern "C" int ac4_md_get_state(ac4_md *md, int r, double *x, double *y, double *z, double *xu, double *yu,
                                double *zu, double *px, double *py, double *pz, int *kx, int *ky, int *kz) {
    TRY(check_rep(md, r));
    const RepC &R = md->reps_h[r];
    std::vector<int> id;
    TRY(get_ids(md, r, id));
    double *outs[] = {x, y, z, xu, yu, zu, px, py, pz};
    const double *devs[] = {md->P.x, md->P.y, md->P.z, md->P.xu, md->P.yu, md->P.zu, md->P.px, md->P.py, md->P.pz};
    for (int k = 0; k < 9; k++)
        if (outs[k]) TRY(fetch_unpermuted(md, r, devs[k], outs[k], id));
    ChainS ch;
    TRY(get_chain(md, r, &ch));
    const double ps = ch.p_scale;  // pending end-of-step U3 factor
    if (ps != 1.0)
        for (int i = 0; i < R.n; i++) {
            if (px) px[i] = px[i] * ps;
            if (py) py[i] = py[i] * ps;
            if (pz) pz[i] = pz[i] * ps;
        }
    if ((kx || ky || kz) && R.nact > 0) {
        std::vector<int8_t> a(R.nact), b(R.nact), c(R.nact);
        TRY(d2h(a.data(), md->kx + R.off_act, R.nact, md->stream));
        TRY(d2h(b.data(), md->ky + R.off_act, R.nact, md->stream));
        TRY(d2h(c.data(), md->kz + R.off_act, R.nact, md->stream));
        AC4_HIP(hipStreamSynchronize(md->stream));
        for (int i = 0; i < R.nact; i++) {
            if (kx) kx[i] = a[i];
            if (ky) ky[i] = b[i];
            if (kz) kz[i] = c[i];
        }
    }
    return AC4_OK;
}

extern "C" int ac4_md_get_forces(ac4_md *md, int r, double *fx, double *fy, double *fz) {
    TRY(check_rep(md, r));
    std::vector<int> id;
    TRY(get_ids(md, r, id));
    if (fx) TRY(fetch_unpermuted(md, r, (const double *)md->fx, fx, id));
    if (fy) TRY(fetch_unpermuted(md, r, (const double *)md->fy, fy, id));
    if (fz) TRY(fetch_unpermuted(md, r, (const double *)md->fz, fz, id));
    return AC4_OK;
}

extern "C" int ac4_md_set_chain(ac4_md *md, int r, const double *c9) {
    TRY(check_rep(md, r));
    AC4_REQUIRE(c9, AC4_EARG, "set_chain: NULL");
    ChainS c;
    TRY(get_chain(md, r, &c));
    for (int k = 0; k < 3; k++) { c.eta[k] = c9[k]; c.peta[k] = c9[3 + k]; c.q[k] = c9[6 + k]; }
    return put_chain(md, r, &c);
}

// fixDrift (MDcode/initialise.c:104-138) for replica r: the momenta summed in
// reference particle order and divided by N are subtracted from every
// particle.  The pending end-of-step U3 factor is applied first (the
// reference's momenta are final after a step), and the sum of p^2 the next
// step's thermostat starts from is recomputed.  mean3 (nullable) receives
// the mean that was removed.
// ---- frame text on the device (MDcode/print.c:4-16 line format)
// One number as glibc's %.5lf: v = m*2^e exactly, q = round-half-even of
// m*10^5*2^e (the 70-bit product kept as two 64-bit words), printed as
// q/10^5 "." q%10^5 with the sign of v.  Values with |v| >= 1e12 or not
// finite set *bad (the host formats those frames).
__device__ __forceinline__ char *fmt_put_u64(char *o, unsigned long long v) {
    char t[24];
    int k = 0;
    do {
        t[k++] = (char)('0' + v % 10ull);
        v /= 10ull;
    } while (v);
    while (k) *o++ = t[--k];
    return o;
}
__device__ char *fmt5_dev(char *o, double v, int *bad) {
    const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
    const int neg = (int)(bits >> 63), ex = (int)((bits >> 52) & 0x7ff);
    const unsigned long long frac = bits & ((1ull << 52) - 1);
    if (ex == 0x7ff || fabs(v) >= 1e12) {
        *bad = 1;
        return o;
    }
    const unsigned long long m = ex ? (frac | (1ull << 52)) : frac;
    const int e = ex ? ex - 1075 : -1074;  // v = m * 2^e, e < 0 for |v| < 1e12 unless v is 0
    unsigned long long q = 0;
    if (m != 0 && e < 0 && -e <= 71) {
        const unsigned long long lo = m * 100000ull, hi = __umul64hi(m, 100000ull);  // t = hi:lo < 2^70
        const int sh = -e;
        // q = t >> sh; rounding bit sh-1 and the sticky bits below it
        q = sh >= 64 ? hi >> (sh - 64) : ((lo >> sh) | (hi << (64 - sh)));
        const int rb = sh - 1;
        const unsigned long long rbit = rb >= 64 ? (hi >> (rb - 64)) & 1ull : (lo >> rb) & 1ull;
        bool sticky;
        if (rb >= 64)
            sticky = lo != 0ull || (rb > 64 && (hi & ((1ull << (rb - 64)) - 1ull)) != 0ull);
        else
            sticky = rb > 0 && (lo & ((1ull << rb) - 1ull)) != 0ull;
        if (rbit && (sticky || (q & 1ull))) q++;
    } else if (m != 0 && e >= 0) {
        *bad = 1;  // an integer of 2^52 or more: not below 1e12
        return o;
    }
    if (neg) *o++ = '-';
    o = fmt_put_u64(o, q / 100000ull);
    *o++ = '.';
    unsigned long long fp = q % 100000ull;
    for (int d = 4; d >= 0; d--) {
        o[d] = (char)('0' + fp % 10ull);
        fp /= 10ull;
    }
    return o + 5;
}

// line of the particle in slot s of replica R, written to slot id[s] (its
// reference index) of out (80 bytes each)
__global__ void __launch_bounds__(BLK) k_fmt_frame(const RepC *__restrict__ reps, int r, const double *__restrict__ xu,
                                                   const double *__restrict__ yu, const double *__restrict__ zu,
                                                   const int *__restrict__ type, const int *__restrict__ id,
                                                   char *out, unsigned char *len, int *bad) {
    const RepC R = reps[r];
    const int s = blockIdx.x * BLK + threadIdx.x;
    if (s >= R.n) return;
    const int g = R.off + s, i = id[g];
    char line[80];
    int b = 0;
    char *o = fmt5_dev(line, xu[g], &b);
    *o++ = '\t';
    o = fmt5_dev(o, yu[g], &b);
    *o++ = '\t';
    o = fmt5_dev(o, zu[g], &b);
    *o++ = '\t';
    const int t = type[g];
    if (t < 0) {
        *o++ = '-';
        o = fmt_put_u64(o, (unsigned long long)(-(long long)t));
    } else {
        o = fmt_put_u64(o, (unsigned long long)t);
    }
    *o++ = '\n';
    const int k = (int)(o - line);
    char *dst = out + (size_t)i * 80;
    for (int c = 0; c < k; c++) dst[c] = line[c];
    len[i] = (unsigned char)k;
    if (b) atomicOr(bad, 1);
}

// bytes of each block of BLK lines (reference order)
__global__ void __launch_bounds__(BLK) k_fmt_bsum(int n, const unsigned char *__restrict__ len,
                                                  long long *__restrict__ boff) {
    __shared__ int sum;
    if (threadIdx.x == 0) sum = 0;
    __syncthreads();
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i < n) atomicAdd(&sum, (int)len[i]);
    __syncthreads();
    if (threadIdx.x == 0) boff[blockIdx.x] = sum;
}

// exclusive prefix of the nb block sums in place, the total to boff[nb]; one
// workgroup of FMT_SCAN threads, each scanning a contiguous chunk
#define FMT_SCAN 1024
__global__ void __launch_bounds__(FMT_SCAN) k_fmt_scan(int nb, long long *__restrict__ boff) {
    __shared__ long long part[FMT_SCAN];
    const int t = threadIdx.x, chunk = (nb + FMT_SCAN - 1) / FMT_SCAN;
    const int lo = min(nb, t * chunk), hi = min(nb, lo + chunk);
    long long acc = 0;
    for (int b = lo; b < hi; b++) acc += boff[b];
    part[t] = acc;
    __syncthreads();
    for (int d = 1; d < FMT_SCAN; d <<= 1) {  // inclusive Hillis-Steele scan
        const long long v = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    long long run = part[t] - acc;
    for (int b = lo; b < hi; b++) {
        const long long v = boff[b];
        boff[b] = run;
        run += v;
    }
    if (t == FMT_SCAN - 1) boff[nb] = part[t];
}

// line i from its slot to its place in the joined text
__global__ void __launch_bounds__(BLK) k_fmt_pack(int n, const char *__restrict__ slots,
                                                  const unsigned char *__restrict__ len,
                                                  const long long *__restrict__ boff, char *__restrict__ txt) {
    __shared__ int sc[BLK];
    const int t = threadIdx.x, i = blockIdx.x * BLK + t;
    const int k = i < n ? (int)len[i] : 0;
    sc[t] = k;
    __syncthreads();
    for (int d = 1; d < BLK; d <<= 1) {
        const int v = t >= d ? sc[t - d] : 0;
        __syncthreads();
        sc[t] += v;
        __syncthreads();
    }
    if (i >= n) return;
    char *dst = txt + boff[blockIdx.x] + (sc[t] - k);
    const char *src = slots + (size_t)i * 80;
    for (int c = 0; c < k; c++) dst[c] = src[c];
}

// data_coor.out text of replica r's current frame (x_u, y_u, z_u, type in
// reference order), formatted and joined on the device; byte for byte the
// text of ac4_format_frame on the same values, which formats a frame holding
// a value the device formatter leaves to the host.
extern "C" long long ac4_md_format_frame(ac4_md *md, int r, char *out, long long cap) {
    TRY(check_rep(md, r));
    const RepC &R = md->reps_h[r];
    const int n = R.n;
    AC4_REQUIRE(out && cap >= 80LL * n, AC4_EARG, "format_frame: buffer needs 80 bytes a line");
    const int nb = (n + BLK - 1) / BLK;
    if ((size_t)n > md->fmt_cap) {
        for (void *q : {(void *)md->fmt_dev, (void *)md->fmt_txt, (void *)md->fmt_len_dev, (void *)md->fmt_boff,
                        (void *)md->fmt_bad})
            if (q) (void)hipFree(q);
        md->fmt_dev = md->fmt_txt = nullptr;
        md->fmt_len_dev = nullptr;
        md->fmt_boff = nullptr;
        md->fmt_bad = nullptr;
        md->fmt_cap = 0;
        const size_t cn = (size_t)n;
        AC4_HIP(hipMalloc(&md->fmt_dev, 80 * cn));
        AC4_HIP(hipMalloc(&md->fmt_txt, 80 * cn));
        AC4_HIP(hipMalloc(&md->fmt_len_dev, cn));
        AC4_HIP(hipMalloc(&md->fmt_boff, sizeof(long long) * (nb + 1)));
        AC4_HIP(hipMalloc(&md->fmt_bad, sizeof(int)));
        md->fmt_cap = cn;
    }
    if (n == 0) return 0;
    hipStream_t s = md->stream;
    AC4_HIP(hipMemsetAsync(md->fmt_bad, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_fmt_frame, dim3(nb), dim3(BLK), 0, s, md->reps, r, md->P.xu, md->P.yu, md->P.zu, md->P.type,
                       md->P.id, md->fmt_dev, md->fmt_len_dev, md->fmt_bad);
    AC4_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_fmt_bsum, dim3(nb), dim3(BLK), 0, s, n, md->fmt_len_dev, md->fmt_boff);
    AC4_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_fmt_scan, dim3(1), dim3(FMT_SCAN), 0, s, nb, md->fmt_boff);
    AC4_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_fmt_pack, dim3(nb), dim3(BLK), 0, s, n, (const char *)md->fmt_dev,
                       (const unsigned char *)md->fmt_len_dev, (const long long *)md->fmt_boff, md->fmt_txt);
    AC4_HIP(hipGetLastError());
    int bad = 0;
    long long total = 0;
    AC4_HIP(hipMemcpyAsync(&bad, md->fmt_bad, sizeof(int), hipMemcpyDeviceToHost, s));
    AC4_HIP(hipMemcpyAsync(&total, md->fmt_boff + nb, sizeof(long long), hipMemcpyDeviceToHost, s));
    AC4_HIP(hipStreamSynchronize(s));
    if (bad) {  // a value the device formatter does not cover: the host formats the frame
        std::vector<double> x(n), y(n), z(n);
        std::vector<int> id, ty(n);
        TRY(get_ids(md, r, id));
        TRY(fetch_unpermuted(md, r, (const double *)md->P.xu, x.data(), id));
        TRY(fetch_unpermuted(md, r, (const double *)md->P.yu, y.data(), id));
        TRY(fetch_unpermuted(md, r, (const double *)md->P.zu, z.data(), id));
        TRY(fetch_unpermuted(md, r, (const int *)md->P.type, ty.data(), id));
        return ac4_format_frame(n, x.data(), y.data(), z.data(), ty.data(), out, cap);
    }
    AC4_REQUIRE(total > 0 && total <= 80LL * n, AC4_ESTATE, "format_frame: joined text of %lld bytes", total);
    AC4_HIP(hipMemcpyAsync(out, md->fmt_txt, (size_t)total, hipMemcpyDeviceToHost, s));
    AC4_HIP(hipStreamSynchronize(s));
    return total;
}

extern "C" int ac4_md_fix_drift(ac4_md *md, int r, double *mean3) {
    TRY(check_rep(md, r));
    const RepC &R = md->reps_h[r];
    const int n = R.n;
    AC4_REQUIRE(n > 0, AC4_ESTATE, "fix_drift: replica %d holds no particles", r);
    std::vector<int> id;
    TRY(get_ids(md, r, id));
    ChainS c;
    TRY(get_chain(md, r, &c));
    double *dev[3] = {md->P.px, md->P.py, md->P.pz};
    std::vector<double> p[3], tmp(n);
    for (int a = 0; a < 3; a++) {
        p[a].resize(n);
        TRY(d2h(tmp.data(), dev[a] + R.off, n, md->stream));
        AC4_HIP(hipStreamSynchronize(md->stream));
        for (int k = 0; k < n; k++) p[a][id[k]] = tmp[k] * c.p_scale;
    }
    double m[3];
    for (int a = 0; a < 3; a++) {
        double sum = 0.0;
        for (int i = 0; i < n; i++) sum += p[a][i];
        m[a] = sum / (double)n;
    }
    double keblk = 0.0;  // time_evolve.c:586-589 order
    for (int i = 0; i < n; i++) {
        p[0][i] -= m[0];
        p[1][i] -= m[1];
        p[2][i] -= m[2];
        keblk += p[0][i] * p[0][i] + p[1][i] * p[1][i] + p[2][i] * p[2][i];
    }
    for (int a = 0; a < 3; a++) {
        for (int k = 0; k < n; k++) tmp[k] = p[a][id[k]];
        TRY(h2d(dev[a] + R.off, tmp.data(), n, md->stream));
        AC4_HIP(hipStreamSynchronize(md->stream));
    }
    c.p_scale = 1.0;
    c.S_end = keblk;
    TRY(put_chain(md, r, &c));
    if (mean3)
        for (int a = 0; a < 3; a++) mean3[a] = m[a];
    return AC4_OK;
}

extern "C" int ac4_md_get_chain(ac4_md *md, int r, double *c9) {
    TRY(check_rep(md, r));
    ChainS c;
    TRY(get_chain(md, r, &c));
    for (int k = 0; k < 3; k++) { c9[k] = c.eta[k]; c9[3 + k] = c.peta[k]; c9[6 + k] = c.q[k]; }
    return AC4_OK;
}

extern "C" int ac4_md_set_rebuild_criterion(ac4_md *md, int criterion) {
    AC4_REQUIRE(md && (criterion == 0 || criterion == 1), AC4_EARG, "criterion must be 0 (safe) or 1 (reference)");
    std::vector<ChainS> c;
    TRY(get_chains(md, c));
    for (auto &e : c) e.crit = criterion;
    return put_chains(md, c);
}

// Cell counts are cleared by k_first_half inside a step; outside a step
// (update_cell_list) with_zero clears them first.
// fixed grid of the work-list kernels: enough workgroups to fill the device
#define WL_GRID 1024
#define TEAM_MOVE_N 400000
static int launch_rebuild(ac4_md *md, bool with_zero = false) {
    hipStream_t s = md->stream;
    const int cy = (md->max_ncell + BLK - 1) / BLK;
    const WorkList wl{md->wl_item, md->wl_n, md->nblk, md->nblk + md->lnblk};
    const int gb = std::min(md->nblk, WL_GRID), gl = std::min(md->lnblk, WL_GRID), gc = std::min(md->ncb, WL_GRID);
    if (with_zero)
        hipLaunchKernelGGL(k_cell_zero, dim3(md->R, cy), dim3(BLK), 0, s, md->reps, md->chain, md->cnt, md->wl_n);
    hipLaunchKernelGGL(k_bin, dim3(md->nblk), dim3(BLK), 0, s, md->reps, md->chain, md->blk_rep, md->blk_first,
                       md->P.x, md->P.y, md->P.z, md->P.id, md->cnt, md->P.cell_of, md->cell_mem, md->P.cell_ref);
    hipLaunchKernelGGL(k_bin_end, dim3(md->R), dim3(md->max_ncell > BIN_END_T ? BIN_END_T : BLK), 0, s, md->reps, md->chain,
                       md->cnt, md->start, wl);
    hipLaunchKernelGGL(k_cell_sort, dim3(gc), dim3(BLK), 0, s, md->reps, wl, md->cnt, md->start,
                       md->cell_mem, md->members, md->ranges, md->xmid, md->P, md->A,
                       PosCopies{md->pos4, md->posq, md->xo, md->yo, md->zo});
    // 16- and 32-lane list teams of handles up to TEAM_MOVE_N particles move
    // the permuted arrays back themselves (measured: 64 replicas, 1.5e5
    // particles, 134.5 -> 130.0 us per step; one box of 1e6, 686 -> 701 us)
    const bool team_moves = !with_zero && (md->nlist_mode == 2 || md->nlist_mode == 3) && md->ntot <= TEAM_MOVE_N;
    if (!team_moves)
        hipLaunchKernelGGL(k_commit, dim3(gb), dim3(BLK), 0, s, md->reps, md->blk_rep, md->blk_first, wl, md->P,
                           md->A);
    if (with_zero)  // update_cell_list: the reference's FP64 test and its 2.85 count
        hipLaunchKernelGGL(k_nlist<true>, dim3(md->nblk), dim3(BLK), 0, s, md->reps, md->chain, md->blk_rep,
                           md->blk_first, md->pos4, md->posq, md->P.cell_of, md->cnt, md->start, md->nn, md->nn_ref,
                           md->nl, md->stride, md->cap);
    else if (md->nlist_mode >= 2) {
        const int w = md->nlist_mode == 2 ? 16 : (md->nlist_mode == 3 ? 32 : 64);
        bool cull = false;  // any replica whose list teams may cull (the kernel checks each replica)
        for (int r = 0; r < md->R; r++) cull |= md->reps_h[r].n > 0 && md->reps_h[r].xcull;
        auto kern = cull ? (w == 16 ? (team_moves ? k_nlist_team<16, true, true> : k_nlist_team<16, false, true>)
                                    : (w == 32 ? (team_moves ? k_nlist_team<32, true, true> : k_nlist_team<32, false, true>)
                                               : k_nlist_team<64, false, true>))
                         : (w == 16 ? (team_moves ? k_nlist_team<16, true, false> : k_nlist_team<16, false, false>)
                                    : (w == 32 ? (team_moves ? k_nlist_team<32, true, false> : k_nlist_team<32, false, false>)
                                               : k_nlist_team<64, false, false>));
        hipLaunchKernelGGL(kern, dim3(gl), dim3(w * LPB), 0, s, md->reps, md->chain, md->lblk_rep, md->lblk_first, wl,
                           md->posq, md->A.cell_of, md->ranges, md->xmid, md->nn, md->nl, md->cap, md->ftab);
    }
    else if (md->nlist_mode == 0)
        hipLaunchKernelGGL(k_nlist_cell, dim3(md->ncell_tot), dim3(64), 0, s, md->reps, md->chain, md->cblk_rep,
                           md->cblk_cell, md->posq, md->cnt, md->start, md->nn, md->nl, md->cap);
    else
        hipLaunchKernelGGL(k_nlist<false>, dim3(md->nblk), dim3(BLK), 0, s, md->reps, md->chain, md->blk_rep,
                           md->blk_first, md->pos4, md->posq, md->P.cell_of, md->cnt, md->start, md->nn, md->nn_ref,
                           md->nl, md->stride, md->cap);
    AC4_HIP(hipGetLastError());
    return AC4_OK;
}

template <bool KICK>
static int launch_force_pass(ac4_md *md) {
    hipStream_t s = md->stream;
    if (md->prec == AC4_FP32)
        hipLaunchKernelGGL((k_force<true, KICK>), dim3(xcd_grid(md->fnblk)), dim3(BLK), 0, s, md->reps,
                           md->fblk_rep, md->fblk_first, md->pos4, md->posq, md->nn, md->nl, md->cap, md->P.act, md->kx,
                           md->ky, md->kz, md->fx, md->fy, md->fz, md->P.px, md->P.py, md->P.pz, md->partial,
                           md->fnblk);
    else
        hipLaunchKernelGGL((k_force<false, KICK>), dim3(xcd_grid(md->fnblk)), dim3(BLK), 0, s, md->reps,
                           md->fblk_rep, md->fblk_first, md->pos4, md->posq, md->nn, md->nl, md->cap, md->P.act, md->kx,
                           md->ky, md->kz, md->fx, md->fy, md->fz, md->P.px, md->P.py, md->P.pz, md->partial,
                           md->fnblk);
    AC4_HIP(hipGetLastError());
    return AC4_OK;
}
template <bool KICK>
static int launch_force_end(ac4_md *md) {
    hipLaunchKernelGGL(k_force_end<KICK>, dim3(md->R), dim3(md->max_fnblk > 64 ? 256 : 64), 0, md->stream, md->reps,
                       md->chain, md->partial);
    AC4_HIP(hipGetLastError());
    return AC4_OK;
}
template <bool KICK>
static int launch_force(ac4_md *md) {
    TRY(launch_force_pass<KICK>(md));
    return launch_force_end<KICK>(md);
}

static void launch_shuffle(ac4_md *md, int force, hipStream_t s) {
    hipLaunchKernelGGL(k_shuffle, dim3(md->R, 3), dim3(256), md->shuf_lds, s, md->reps, md->chain, md->kx, md->ky,
                       md->kz, force, md->shuf_kcap, md->shuf_bmax > 0 ? 1 : 0);
    if (md->shuf_bmax > 0) {  // slots that can hold more than SHUF_PAR_MIN active particles
        hipLaunchKernelGGL(k_shuf_part, dim3(md->shuf_bmax, 3, md->R), dim3(64), 0, s, md->reps, md->chain, force,
                           md->shuf_bmax, md->shuf_ent, md->shuf_ent_n);
        hipLaunchKernelGGL(k_shuf_apply, dim3(md->R, 3), dim3(2 * SHUF_M), md->shuf_lds, s, md->reps, md->chain,
                           md->kx, md->ky, md->kz, force, md->shuf_kcap, md->shuf_bmax, md->shuf_ent,
                           md->shuf_ent_n);
    }
}
// with_shuffle: the step's shuffle runs in k_first_half's first 3*R
// workgroups (handles whose shuffle LDS is small, see shuf_fused)
static void launch_first_half(ac4_md *md, bool with_shuffle) {
    const ShufArgs sa{md->kx, md->ky, md->kz, md->shuf_kcap, with_shuffle ? 3 * md->R : 0};
    hipLaunchKernelGGL(k_first_half, dim3(md->nblk + sa.nblk), dim3(BLK), with_shuffle ? md->shuf_lds : 0, md->stream,
                       md->reps, md->chain, md->blk_rep, md->blk_first, md->P.x, md->P.y, md->P.z, md->P.xu, md->P.yu,
                       md->P.zu, md->P.px, md->P.py, md->P.pz, md->fx, md->fy, md->fz, md->xo, md->yo, md->zo,
                       md->P.type, md->pos4, md->posq, md->prec == AC4_FP32 ? 1 : 0, md->chain, md->blk_idx, md->cnt,
                       md->wl_n, sa);
}

// One MD step, one chain of launches on the handle's stream (also when
// captured in a graph).  The shuffle, which only the force kernel needs, once
// ran on a second graph branch beside the half kick and the rebuild: the
// fork and join cost more than the overlap gave (about 9 us per step).  It
// now runs inside k_first_half where its LDS allows, else as k_shuffle first.
static int launch_step(ac4_md *md) {
    if (md->shuf_fused) {
        launch_first_half(md, true);
    } else {
        launch_shuffle(md, 0, md->stream);
        launch_first_half(md, false);
    }
    AC4_HIP(hipGetLastError());
    TRY(launch_rebuild(md));
    return launch_force<true>(md);
}

static int check_err(ac4_md *md) {
    std::vector<ChainS> c;
    TRY(get_chains(md, c));
    for (int r = 0; r < md->R; r++) {
        AC4_REQUIRE(!(c[r].err & 1), AC4_ECAP, "replica %d: neighbour list exceeded %d entries", r, md->cap);
        AC4_REQUIRE(!(c[r].err & 2), AC4_ESTATE, "replica %d: non-finite or runaway coordinate (run blew up)", r);
    }
    return AC4_OK;
}

static int force_rebuild_all(ac4_md *md) {
    std::vector<ChainS> c;
    TRY(get_chains(md, c));
    for (auto &e : c) e.force_rebuild = 1;
    TRY(put_chains(md, c));
    TRY(launch_rebuild(md, true));
    TRY(get_chains(md, c));
    for (auto &e : c) { e.force_rebuild = 0; e.maxbits = 0ull; }
    return put_chains(md, c);
}

extern "C" int ac4_md_update_cell_list(ac4_md *md, int r_out, int *cell_out, int *full_count_out) {
    AC4_REQUIRE(md, AC4_EARG, "NULL handle");
    TRY(force_rebuild_all(md));
    if (r_out >= 0 && r_out < md->R) {
        std::vector<int> id;
        TRY(get_ids(md, r_out, id));
        if (cell_out) TRY(fetch_unpermuted(md, r_out, (const int *)md->P.cell_ref, cell_out, id));
        if (full_count_out) TRY(fetch_unpermuted(md, r_out, (const int *)md->nn_ref, full_count_out, id));
    }
    return check_err(md);
}

extern "C" int ac4_md_calculate_forces(ac4_md *md, double *out2) {
    AC4_REQUIRE(md, AC4_EARG, "NULL handle");
    std::vector<ChainS> c;
    TRY(get_chains(md, c));
    bool stale = false;
    for (auto &e : c) stale |= e.force_rebuild != 0;
    if (stale) TRY(force_rebuild_all(md));
    TRY(launch_force<false>(md));
    TRY(get_chains(md, c));
    if (out2)
        for (int r = 0; r < md->R; r++) { out2[2 * r] = c[r].pe; out2[2 * r + 1] = c[r].vir; }
    md->have_forces = true;
    return check_err(md);
}

extern "C" int ac4_md_step(ac4_md *md, int nsteps) {
    AC4_REQUIRE(md && nsteps >= 0, AC4_EARG, "bad arguments");
    if (!md->have_forces) TRY(ac4_md_calculate_forces(md, nullptr));
    for (int k = 0; k < nsteps; k++) TRY(launch_step(md));
    return AC4_OK;
}

// Steps are replayed from captured graphs: GRAPH_CHUNK steps per launch (one
// launch per ten steps instead of one per step: 1.53e9 -> 1.60e9
// particle-steps/s on the 64-replica ensemble in three launch groups; 25 steps
// instead of ten: 2.093e9 -> 2.101e9), and a one-step graph for the remainder.  Every step's kernels and arguments are the same (the step
// state lives on the device), so a chunk replays exactly the steps launched
// one by one.
#define GRAPH_CHUNK 25
static int capture_steps(ac4_md *md, int k) {
    const int m = k ? GRAPH_CHUNK : 1;
    AC4_HIP(hipStreamBeginCapture(md->stream, hipStreamCaptureModeThreadLocal));
    int rc = AC4_OK;
    for (int q = 0; q < m && rc == AC4_OK; q++) rc = launch_step(md);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(md->stream, &g);
    if (rc != AC4_OK || e != hipSuccess) {  // nothing of a failed capture is kept
        if (g) (void)hipGraphDestroy(g);
        if (rc != AC4_OK) return rc;
        AC4_HIP(e);
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        AC4_HIP(e);
    }
    md->graph[k] = g;
    md->gexec[k] = x;
    return AC4_OK;
}

static int ready_graphs(ac4_md *md) {
    if (!md->have_forces) TRY(ac4_md_calculate_forces(md, nullptr));
    for (int k = 0; k < 2; k++)
        if (!md->gexec[k]) TRY(capture_steps(md, k));
    return AC4_OK;
}

extern "C" int ac4_md_step_graph(ac4_md *md, int nsteps) {
    AC4_REQUIRE(md && nsteps >= 0, AC4_EARG, "bad arguments");
    TRY(ready_graphs(md));
    for (int k = 0; k < nsteps / GRAPH_CHUNK; k++) AC4_HIP(hipGraphLaunch(md->gexec[1], md->stream));
    for (int k = 0; k < nsteps % GRAPH_CHUNK; k++) AC4_HIP(hipGraphLaunch(md->gexec[0], md->stream));
    return AC4_OK;
}

// Several handles stepped together: each on its own stream, the graph
// replays interleaved step by step so that the handles' kernels overlap on
// the device (one handle's force kernel beside another's list rebuild).
// Every handle's arithmetic is that of ac4_md_step_graph alone.
extern "C" int ac4_md_step_graph_multi(ac4_md *const *mds, int nh, int nsteps) {
    AC4_REQUIRE(mds && nh > 0 && nsteps >= 0, AC4_EARG, "step_graph_multi: bad arguments");
    for (int h = 0; h < nh; h++) {
        AC4_REQUIRE(mds[h], AC4_EARG, "step_graph_multi: handle %d is NULL", h);
        for (int q = 0; q < h; q++)
            AC4_REQUIRE(mds[q] != mds[h], AC4_EARG, "step_graph_multi: handle %d appears twice", h);
        TRY(ready_graphs(mds[h]));  // forces and the captured steps of every handle first
    }
    for (int k = 0; k < nsteps / GRAPH_CHUNK; k++)
        for (int h = 0; h < nh; h++) AC4_HIP(hipGraphLaunch(mds[h]->gexec[1], mds[h]->stream));
    for (int k = 0; k < nsteps % GRAPH_CHUNK; k++)
        for (int h = 0; h < nh; h++) AC4_HIP(hipGraphLaunch(mds[h]->gexec[0], mds[h]->stream));
    return AC4_OK;
}

extern "C" int ac4_md_relax(ac4_md *md, int nsteps, double f_start, double f_end) {
    AC4_REQUIRE(md && nsteps > 0 && f_start > 0 && f_end > 0, AC4_EARG, "relax: bad arguments");
    if (!md->have_forces) TRY(ac4_md_calculate_forces(md, nullptr));
    std::vector<ChainS> c;
    TRY(get_chains(md, c));
    for (auto &e : c) { e.mode = 1; e.ramp_n = nsteps; e.ramp_k = 0; e.ramp0 = f_start; e.ramp1 = f_end; }
    TRY(put_chains(md, c));
    for (int k = 0; k < nsteps; k++) TRY(launch_step(md));
    return check_err(md);
}

// Eager steps with HIP events (on the handle's stream) around each launch
// group; ms = average ms per step of {first_half, rebuild group, force,
// shuffle, force end}: the force entry is k_force alone, the last k_force_end;
// where the shuffle runs inside k_first_half its time is in the first entry
// and the fourth is the empty interval between two events.
extern "C" int ac4_md_kernel_times(ac4_md *md, int nsteps, double *ms) {
    AC4_REQUIRE(md && nsteps > 0 && ms, AC4_EARG, "kernel_times: bad arguments");
    if (!md->have_forces) TRY(ac4_md_calculate_forces(md, nullptr));
    hipStream_t s = md->stream;
    hipEvent_t ev[6];
    for (auto &e : ev) AC4_HIP(hipEventCreate(&e));
    double acc[5] = {0, 0, 0, 0, 0};
    for (int k = 0; k < nsteps; k++) {
        AC4_HIP(hipEventRecord(ev[0], s));
        if (!md->shuf_fused) launch_shuffle(md, 0, s);
        AC4_HIP(hipEventRecord(ev[1], s));
        launch_first_half(md, md->shuf_fused);
        AC4_HIP(hipEventRecord(ev[2], s));
        TRY(launch_rebuild(md));
        AC4_HIP(hipEventRecord(ev[3], s));
        TRY(launch_force_pass<true>(md));
        AC4_HIP(hipEventRecord(ev[4], s));
        TRY(launch_force_end<true>(md));
        AC4_HIP(hipEventRecord(ev[5], s));
        AC4_HIP(hipEventSynchronize(ev[5]));
        float t[5];
        for (int q = 0; q < 5; q++) AC4_HIP(hipEventElapsedTime(&t[q], ev[q], ev[q + 1]));
        acc[3] += t[0]; acc[0] += t[1]; acc[1] += t[2]; acc[2] += t[3]; acc[4] += t[4];
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    for (int q = 0; q < 5; q++) ms[q] = acc[q] / nsteps;
    return check_err(md);
}

extern "C" int ac4_md_shuffle(ac4_md *md) {
    AC4_REQUIRE(md, AC4_EARG, "NULL handle");
    launch_shuffle(md, 1, md->stream);
    AC4_HIP(hipGetLastError());
    std::vector<ChainS> c;
    TRY(get_chains(md, c));  // synchronises
    for (auto &e : c) e.shuffles += 1;
    return put_chains(md, c);
}

extern "C" int ac4_md_shuffle_uniforms(ac4_md *md, int r, double *u) {
    TRY(check_rep(md, r));
    AC4_REQUIRE(u, AC4_EARG, "NULL output");
    ChainS c;
    TRY(get_chain(md, r, &c));
    const RepC &R = md->reps_h[r];
    for (int comp = 0; comp < 3; comp++)
        for (int i = 0; i < R.nact; i++)
            u[comp * R.nact + i] = ac4_shuffle_uniform(R.key0, R.key1, (uint64_t)c.shuffles, comp, i);
    return AC4_OK;
}

extern "C" int ac4_md_stats(ac4_md *md, int r, double *out8) {
    TRY(check_rep(md, r));
    ChainS c;
    TRY(get_chain(md, r, &c));
    const RepC &R = md->reps_h[r];
    out8[0] = c.pe;
    out8[1] = c.vir;
    out8[2] = c.ke;
    out8[3] = c.inst_T;
    out8[4] = R.density * c.inst_T + c.vir / (R.L * R.L * R.L);  // print.c:83
    out8[5] = c.drift;
    out8[6] = c.heff;
    out8[7] = (double)c.step;
    return AC4_OK;
}

extern "C" int ac4_md_rebuilds(ac4_md *md, int r) {
    TRY(check_rep(md, r));
    ChainS c;
    TRY(get_chain(md, r, &c));
    return c.rebuilds;
}

extern "C" int ac4_md_sync(ac4_md *md) {
    AC4_REQUIRE(md, AC4_EARG, "NULL handle");
    AC4_HIP(hipStreamSynchronize(md->stream));
    return check_err(md);
}

extern "C" int ac4_philox4x32(const uint32_t *ctr4, const uint32_t *key2, uint32_t *out4) {
    AC4_REQUIRE(ctr4 && key2 && out4, AC4_EARG, "NULL argument");
    ac4_philox(ctr4, key2, out4);
    return AC4_OK;
}

// Reference-order x_u of every active slot into the export buffers and one
// copy to the host (xyz[3][ntot], slot offsets as ac4_md_slot_offset).
extern "C" int ac4_md_export_all(ac4_md *md, double *xu, double *yu, double *zu) {
    AC4_REQUIRE(md && xu && yu && zu, AC4_EARG, "export_all: bad arguments");
    if (!md->ex) {
        TRY(dalloc(&md->ex, md->ntot)); TRY(dalloc(&md->ey, md->ntot)); TRY(dalloc(&md->ez, md->ntot));
        TRY(dalloc(&md->et, md->ntot));
    }
    hipLaunchKernelGGL(k_export, dim3(md->nblk), dim3(BLK), 0, md->stream, md->reps, md->blk_rep, md->blk_first,
                       md->P.id, md->P.xu, md->P.yu, md->P.zu, md->P.type, md->ex, md->ey, md->ez, md->et, -1);
    AC4_HIP(hipGetLastError());
    TRY(d2h(xu, (const double *)md->ex, md->ntot, md->stream));
    TRY(d2h(yu, (const double *)md->ey, md->ntot, md->stream));
    TRY(d2h(zu, (const double *)md->ez, md->ntot, md->stream));
    AC4_HIP(hipStreamSynchronize(md->stream));
    return check_err(md);
}

extern "C" int ac4_md_overlap_enable(ac4_md *md, int sample_every, int origin_every, int n_origin, int n_lags,
                                     double a) {
    AC4_REQUIRE(md && sample_every > 0 && origin_every > 0 && origin_every % sample_every == 0 && n_origin > 0 &&
                    n_lags > 0 && a > 0,
                AC4_EARG, "overlap_enable: bad arguments (origin_every must be a multiple of sample_every)");
    AC4_REQUIRE(!md->ox, AC4_ESTATE, "overlap accumulators already enabled");
    md->ov = OvCfg{sample_every, origin_every, n_origin, n_lags, a * a};
    const size_t snap = (size_t)md->ntot * n_origin, acc = (size_t)md->R * (n_lags + 1);
    TRY(dalloc(&md->ox, snap)); TRY(dalloc(&md->oy, snap)); TRY(dalloc(&md->oz, snap));
    TRY(dalloc(&md->ov_q, acc)); TRY(dalloc(&md->ov_q2, acc)); TRY(dalloc(&md->ov_n, acc));
    TRY(dalloc(&md->ov_cnt, (size_t)md->R * n_origin));
    TRY(dalloc(&md->ov_step, (size_t)md->R * n_origin));
    TRY(dalloc(&md->ov_list, (size_t)md->R));
    TRY(dalloc(&md->ov_out, 3 * (size_t)(n_lags + 1)));
    std::vector<long long> none((size_t)md->R * n_origin, -1);
    AC4_HIP(hipMemcpy(md->ov_step, none.data(), sizeof(long long) * none.size(), hipMemcpyHostToDevice));
    return AC4_OK;
}

// Counts overlaps for the current step (every replica) and stores a new
// origin when due.  Call when the step counters are multiples of sample_every.
extern "C" int ac4_md_overlap_sample(ac4_md *md) {
    AC4_REQUIRE(md && md->ox, AC4_ESTATE, "overlap accumulators not enabled");
    hipStream_t s = md->stream;
    hipLaunchKernelGGL(k_ov_count, dim3(md->nblk, md->ov.n_origin), dim3(BLK), 0, s, md->reps, md->chain,
                       md->blk_rep, md->blk_first, md->ov, md->P.id, md->P.xu, md->P.yu, md->P.zu, md->ox, md->oy,
                       md->oz, md->ov_step, md->ov_cnt);
    const int nt = md->R * md->ov.n_origin;
    hipLaunchKernelGGL(k_ov_fold, dim3((nt + 255) / 256), dim3(256), 0, s, md->reps, md->chain, md->ov, md->R,
                       md->ov_step, md->ov_cnt, md->ov_q, md->ov_q2, md->ov_n);
    hipLaunchKernelGGL(k_ov_origin, dim3(md->nblk), dim3(BLK), 0, s, md->reps, md->chain, md->blk_rep,
                       md->blk_first, md->ov, md->P.id, md->P.xu, md->P.yu, md->P.zu, md->ox, md->oy, md->oz,
                       md->ov_step);
    AC4_HIP(hipGetLastError());
    return AC4_OK;
}

// Per-lag sums of replica r: sq[l] = sum_o Q_o(l*sample_every), sq2 likewise
// for Q^2, cnt[l] = number of origins (l = 0..n_lags).
extern "C" int ac4_md_overlap_get(ac4_md *md, int r, double *sq, double *sq2, int *cnt) {
    AC4_REQUIRE(md && md->ox && r >= 0 && r < md->R, AC4_EARG, "overlap_get: bad arguments");
    const size_t nl = md->ov.n_lags + 1, off = (size_t)r * nl;
    if (sq) TRY(d2h(sq, (const double *)md->ov_q + off, nl, md->stream));
    if (sq2) TRY(d2h(sq2, (const double *)md->ov_q2 + off, nl, md->stream));
    if (cnt) TRY(d2h(cnt, (const int *)md->ov_n + off, nl, md->stream));
    AC4_HIP(hipStreamSynchronize(md->stream));
    return AC4_OK;
}

// chi4(l) pooled over the listed replicas and their origins (k_ov_chi4):
// out3[3*l + {0,1,2}] = {Nbar*(<Q^2>-<Q>^2), <Q>, origins}, l = 0..n_lags.
extern "C" int ac4_md_overlap_chi4(ac4_md *md, int nlist, const int *list, double *out3) {
    AC4_REQUIRE(md && md->ox && nlist > 0 && list && out3, AC4_EARG, "overlap_chi4: bad arguments");
    for (int k = 0; k < nlist; k++)
        AC4_REQUIRE(list[k] >= 0 && list[k] < md->R && md->reps_h[list[k]].n > 0, AC4_EARG,
                    "overlap_chi4: entry %d names replica %d, which is not a running replica", k, list[k]);
    const int nl = md->ov.n_lags + 1;
    AC4_HIP(hipMemcpyAsync(md->ov_list, list, sizeof(int) * nlist, hipMemcpyHostToDevice, md->stream));
    hipLaunchKernelGGL(k_ov_chi4, dim3((nl + 255) / 256), dim3(256), 0, md->stream, md->reps, md->ov, md->ov_list,
                       nlist, md->ov_q, md->ov_q2, md->ov_n, md->ov_out);
    AC4_HIP(hipGetLastError());
    TRY(d2h(out3, (const double *)md->ov_out, 3 * (size_t)nl, md->stream));
    AC4_HIP(hipStreamSynchronize(md->stream));
    return AC4_OK;
}

// Starts replica r's accumulators afresh (no origins, zero sums).
extern "C" int ac4_md_overlap_reset(ac4_md *md, int r) {
    AC4_REQUIRE(md && md->ox && r >= 0 && r < md->R, AC4_EARG, "overlap_reset: bad arguments");
    const int nt = std::max(md->ov.n_origin, md->ov.n_lags + 1);
    hipLaunchKernelGGL(k_ov_reset, dim3((nt + 255) / 256), dim3(256), 0, md->stream, md->ov, r, md->ov_step,
                       md->ov_cnt, md->ov_q, md->ov_q2, md->ov_n);
    AC4_HIP(hipGetLastError());
    return AC4_OK;
}

extern "C" int ac4_md_slot_offset(const ac4_md *md, int slot) {
    if (!md || slot < 0 || slot >= md->R) return AC4_EARG;
    return md->reps_h[slot].off;
}

// Reference-order device copies of x_u and type of replica r for in-situ
// analysis (ac4_analysis.hip); the copies are complete when this returns.
extern "C" int ac4_md_device_xu(ac4_md *md, int r, const double **xu, const double **yu, const double **zu,
                                const int **type, int *n) {
    TRY(check_rep(md, r));
    if (!md->ex) {
        TRY(dalloc(&md->ex, md->ntot)); TRY(dalloc(&md->ey, md->ntot)); TRY(dalloc(&md->ez, md->ntot));
        TRY(dalloc(&md->et, md->ntot));
    }
    hipLaunchKernelGGL(k_export, dim3(md->nblk), dim3(BLK), 0, md->stream, md->reps, md->blk_rep, md->blk_first,
                       md->P.id, md->P.xu, md->P.yu, md->P.zu, md->P.type, md->ex, md->ey, md->ez, md->et, r);
    AC4_HIP(hipGetLastError());
    AC4_HIP(hipStreamSynchronize(md->stream));
    const RepC &R = md->reps_h[r];
    *xu = md->ex + R.off;
    *yu = md->ey + R.off;
    *zu = md->ez + R.off;
    *type = md->et + R.off;
    *n = R.n;
    return AC4_OK;
}

// ------------------------------------------------------ exact checkpoints
// A replica's complete device state in slot (cell) order: particle arrays,
// forces, the last-rebuild positions, the cell tables, the neighbour list and
// the chain.  Restoring copies it back without recomputing anything, so the
// continued run is bitwise the one that was interrupted.  Slot and cell
// indices are stored replica-local, so the target may be another slot, of a
// handle with the same shape (precision, list capacity and layout, team
// width, skin) configured with the same replica parameters.
struct CkHead {
    char magic[8];
    int32_t version, n, nact, ncell, cap, lc, tpp, prec;
    double L, skin, T, f0;
    int32_t tau, pad;
    uint64_t seed, chain_bytes, total;
    // in-situ overlap state (ac4_md_overlap_enable) when the handle has it:
    // the origin snapshots in reference order, their steps, the per-lag sums
    int32_t ov_sample, ov_origin_every, ov_n_origin, ov_n_lags;
    double ov_a2;
};
static const char k_ck_magic[8] = {'A', 'C', '4', 'C', 'K', 'P', 'T', '1'};

static size_t ck_ov_bytes(const ac4_md *md, const RepC &R) {
    if (!md->ox) return 0;
    const size_t no = md->ov.n_origin, nl = md->ov.n_lags + 1;
    return no * 3 * (size_t)R.n * sizeof(double) + no * sizeof(long long) + nl * (2 * sizeof(double) + sizeof(int));
}
static size_t ck_bytes(const ac4_md *md, const RepC &R) {
    const size_t n = R.n, nc = R.ncell;
    return sizeof(CkHead) + sizeof(ChainS) + n * (15 * sizeof(double) + sizeof(double4) + sizeof(QPos)) +
           n * 7 * sizeof(int) + 3 * (size_t)R.nact + nc * 2 * sizeof(int) + n * sizeof(int) +
           nc * NRANGE * sizeof(int2) + (size_t)md->cap * n * sizeof(int) + ck_ov_bytes(md, R);
}

extern "C" long long ac4_md_checkpoint_size(ac4_md *md, int r) {
    if (check_rep(md, r) != AC4_OK) return AC4_EARG;
    return (long long)ck_bytes(md, md->reps_h[r]);
}

template <class T>
static int ck_io(bool save, T *dev, size_t n, char *&p, hipStream_t s) {
    if (save)
        TRY(d2h((T *)p, (const T *)dev, n, s));
    else
        TRY(h2d(dev, (const T *)p, n, s));
    p += n * sizeof(T);
    return AC4_OK;
}

static int ck_arrays(ac4_md *md, int r, bool save, char *p) {
    const RepC &R = md->reps_h[r];
    hipStream_t s = md->stream;
    const size_t g = R.off, n = R.n, c0 = R.off_cell, nc = R.ncell;
    double *dv[15] = {md->P.x, md->P.y, md->P.z, md->P.xu, md->P.yu, md->P.zu, md->P.px, md->P.py, md->P.pz,
                      md->xo, md->yo, md->zo, md->fx, md->fy, md->fz};
    for (double *a : dv) TRY(ck_io(save, a + g, n, p, s));
    TRY(ck_io(save, md->pos4 + g, n, p, s));
    TRY(ck_io(save, md->posq + g, n, p, s));
    int *iv[7] = {md->P.type, md->P.act, md->P.id, md->P.cell_ref, md->P.cell_of, md->nn, md->nn_ref};
    for (int *a : iv) TRY(ck_io(save, a + g, n, p, s));
    TRY(ck_io(save, md->kx + R.off_act, (size_t)R.nact, p, s));
    TRY(ck_io(save, md->ky + R.off_act, (size_t)R.nact, p, s));
    TRY(ck_io(save, md->kz + R.off_act, (size_t)R.nact, p, s));
    TRY(ck_io(save, md->cnt + c0, nc, p, s));
    TRY(ck_io(save, md->start + c0, nc, p, s));
    TRY(ck_io(save, md->members + g, n, p, s));
    int2 *rg = (int2 *)p;
    TRY(ck_io(save, md->ranges + c0 * NRANGE, nc * NRANGE, p, s));
    int *nl = (int *)p;
    TRY(ck_io(save, md->nl + g * md->cap, (size_t)md->cap * n, p, s));
    if (md->ox) {  // origin snapshots of this replica, its origin steps, its per-lag sums
        const size_t no = md->ov.n_origin, nl1 = md->ov.n_lags + 1;
        for (size_t o = 0; o < no; o++) {
            const size_t base = g * no + o * (size_t)R.cap_n;
            TRY(ck_io(save, md->ox + base, n, p, s));
            TRY(ck_io(save, md->oy + base, n, p, s));
            TRY(ck_io(save, md->oz + base, n, p, s));
        }
        TRY(ck_io(save, md->ov_step + (size_t)r * no, no, p, s));
        TRY(ck_io(save, md->ov_q + (size_t)r * nl1, nl1, p, s));
        TRY(ck_io(save, md->ov_q2 + (size_t)r * nl1, nl1, p, s));
        TRY(ck_io(save, md->ov_n + (size_t)r * nl1, nl1, p, s));
    }
    AC4_HIP(hipStreamSynchronize(s));
    if (save) {  // replica-local slot indices
        for (size_t k = 0; k < nc * NRANGE; k++)
            if (rg[k].y > rg[k].x) { rg[k].x -= (int)g; rg[k].y -= (int)g; }
        for (size_t k = 0; k < (size_t)md->cap * n; k++) nl[k] -= (int)g;
    }
    return AC4_OK;
}

extern "C" int ac4_md_checkpoint(ac4_md *md, int r, void *buf, long long cap) {
    TRY(check_rep(md, r));
    const RepC &R = md->reps_h[r];
    AC4_REQUIRE(buf && cap >= (long long)ck_bytes(md, R), AC4_EARG, "checkpoint: buffer smaller than %zu bytes",
                ck_bytes(md, R));
    AC4_REQUIRE(R.n > 0, AC4_ESTATE, "checkpoint: slot %d is idle", r);
    AC4_HIP(hipStreamSynchronize(md->stream));
    CkHead h;
    memset(&h, 0, sizeof h);
    memcpy(h.magic, k_ck_magic, 8);
    h.version = 3; h.n = R.n; h.nact = R.nact; h.ncell = R.ncell; h.cap = md->cap; h.lc = R.lc; h.tpp = R.tpp;
    h.prec = md->prec; h.L = R.L; h.skin = R.skin; h.T = R.T; h.f0 = R.f0; h.tau = R.tau;
    h.seed = (uint64_t)R.key0 | ((uint64_t)R.key1 << 32);
    h.chain_bytes = sizeof(ChainS);
    h.total = ck_bytes(md, R);
    if (md->ox) {
        h.ov_sample = md->ov.sample_every; h.ov_origin_every = md->ov.origin_every;
        h.ov_n_origin = md->ov.n_origin; h.ov_n_lags = md->ov.n_lags; h.ov_a2 = md->ov.a2;
    }
    char *p = (char *)buf;
    memcpy(p, &h, sizeof h);
    p += sizeof h;
    ChainS c;
    TRY(get_chain(md, r, &c));
    memcpy(p, &c, sizeof c);
    p += sizeof c;
    return ck_arrays(md, r, true, p);
}

extern "C" int ac4_md_restore(ac4_md *md, int r, const void *buf, long long len) {
    TRY(check_rep(md, r));
    const RepC &R = md->reps_h[r];
    AC4_REQUIRE(buf && len >= (long long)sizeof(CkHead), AC4_EARG, "restore: no checkpoint");
    CkHead h;
    memcpy(&h, buf, sizeof h);
    AC4_REQUIRE(!memcmp(h.magic, k_ck_magic, 8) && h.version == 3 && h.chain_bytes == sizeof(ChainS), AC4_EARG,
                "restore: not a checkpoint of this library version");
    AC4_REQUIRE((long long)h.total == len && h.total == ck_bytes(md, R), AC4_EARG,
                "restore: checkpoint of %lld bytes does not fit slot %d", len, r);
    AC4_REQUIRE(h.n == R.n && h.nact == R.nact && h.ncell == R.ncell && h.cap == md->cap && h.lc == R.lc &&
                    h.tpp == R.tpp && h.prec == md->prec && h.L == R.L && h.skin == R.skin,
                AC4_EARG, "restore: the slot's replica or handle shape differs from the checkpoint's");
    const bool ov_same = md->ox ? (h.ov_sample == md->ov.sample_every && h.ov_origin_every == md->ov.origin_every &&
                                   h.ov_n_origin == md->ov.n_origin && h.ov_n_lags == md->ov.n_lags &&
                                   h.ov_a2 == md->ov.a2)
                                : h.ov_n_origin == 0;
    AC4_REQUIRE(ov_same, AC4_EARG, "restore: the overlap accumulators of the checkpoint and the handle differ");
    const uint64_t seed = (uint64_t)R.key0 | ((uint64_t)R.key1 << 32);
    AC4_REQUIRE(h.seed == seed && h.T == R.T && h.f0 == R.f0 && h.tau == R.tau, AC4_EARG,
                "restore: slot %d runs seed %llu, T %g, f0 %g, tau %d; the checkpoint seed %llu, T %g, f0 %g, tau %d",
                r, (unsigned long long)seed, R.T, R.f0, R.tau, (unsigned long long)h.seed, h.T, h.f0, h.tau);
    std::vector<char> tmp((const char *)buf, (const char *)buf + len);  // local copy: indices are rebased
    char *p = tmp.data() + sizeof h;
    ChainS c;
    memcpy(&c, p, sizeof c);
    p += sizeof c;
    const size_t g = R.off, n = R.n, nc = R.ncell;
    // rebase the replica-local indices onto this slot before copying
    char *q = p + n * (15 * sizeof(double) + sizeof(double4) + sizeof(QPos)) + n * 7 * sizeof(int) +
              3 * (size_t)R.nact + nc * 2 * sizeof(int) + n * sizeof(int);
    int2 *rg = (int2 *)q;
    for (size_t k = 0; k < nc * NRANGE; k++)
        if (rg[k].y > rg[k].x) { rg[k].x += (int)g; rg[k].y += (int)g; }
    int *nl = (int *)(q + nc * NRANGE * sizeof(int2));
    for (size_t k = 0; k < (size_t)md->cap * n; k++) nl[k] += (int)g;
    TRY(ck_arrays(md, r, false, p));
    c.done_blocks = 0u;
    c.bin_done = 0u;
    return put_chain(md, r, &c);
}
