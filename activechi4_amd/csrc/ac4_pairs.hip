// ac4_pairs.hip -- pair observables of a device frame store:
//
//  ac4_frames_overlap_cage  Analysis/overlap_cage_rel.c:4-269 (paths relative
//      to the reference checkout): the overlap of each particle's
//      displacement relative to the mean displacement of its neighbours
//      within 1.3 sigma at the origin (its cage), relaxMPICR1_a_<a>.dat.
//  ac4_frames_guu           Analysis/guu.c:3-300: the displacement-
//      displacement correlation g_uu(r) of the active particles at the lag
//      nearest tau, guuTauAlphaAct.dat.
//
// Cage overlap, per origin: particles are binned on the cell grid of
// GetList (cells of edge 1/(int)(L/cell_cut) in box units); k_cg_list
// gives every particle its neighbours in the order GetList's traversal
// appends them (for particle i: the j < i whose own scan finds i, ascending;
// then i's own scan, own cell then the 13 half-stencil cells, each ascending;
// then the j > i whose scan finds i, ascending), so the cage sums are formed
// in the reference's order; the per-lag count is exact.  Needs at least 3
// cells per edge, where the 27 cells around a cell are distinct (the
// reference's small-box double counting is not reproduced).
//
// g_uu, per origin: displacements of the active particles (k_gu_disp), the
// running MSD and mean-displacement sums in particle order on one thread
// (k_gu_sums), then all active pairs i < j on a triangle of 256 x 256 tiles
// (k_gu_hist) with the histogram of a bin range in LDS.  Bin indices
// (int)(sqrt(r2)/bin) use correctly rounded sqrt and division, so the pair
// counts are exact; the two weighted histograms are FP64 sums whose order
// follows the atomics (agreement with the reference to rounding).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <vector>
#include "ac4_internal.h"

#define PR_THREADS 256
#define CG_CAP 64            // neighbours per particle within 1.3 sigma (about 13 at rho = 1.2)
#define GU_TILE 256
#define GU_LDS_BINS 6144     // bins per pass: 6144 * 20 B = 120 KB of LDS
#define PR_TIME_STEP 0.005   // Analysis/global.h:19
#define PR_DIVIDE_TO 2.0     // Analysis/global.h:21
#define PR_PI 3.14159265358979
#define PR_RHO 1.2

__device__ __forceinline__ double pr_wrap(double v) {  // the while loops of read_all_data.c LoadOrigin
    return (v >= 1.0 || v < 0.0) ? v - floor(v) : v;
}
__device__ __forceinline__ double pr_mi(double d) {  // overlap_cage_rel.c:85-96, guu.c:157-168
    if (d >= 0.5)
        d -= 1.0;
    else if (d < -0.5)
        d += 1.0;
    return d;
}

// ------------------------------------------------------------ cage overlap

// wrapped origin positions and their GetList cell (1-based, :210-216)
__global__ void __launch_bounds__(PR_THREADS) k_cg_bin(const double *__restrict__ x0, const double *__restrict__ y0,
                                                       const double *__restrict__ z0, int n, int nc, double cs,
                                                       double *wx, double *wy, double *wz, int *cell,
                                                       unsigned int *cnt) {
    const int i = blockIdx.x * PR_THREADS + threadIdx.x;
    if (i >= n) return;
    const double x = pr_wrap(x0[i]), y = pr_wrap(y0[i]), z = pr_wrap(z0[i]);
    wx[i] = x;
    wy[i] = y;
    wz[i] = z;
    int c = (int)(1 + floor(x / cs) + floor(y / cs) * nc + floor(z / cs) * nc * nc);
    const int ncell = nc * nc * nc;
    c = c < 1 ? 1 : (c > ncell ? ncell : c);
    cell[i] = c;
    atomicAdd(&cnt[c], 1u);
}

// exclusive scan of cnt[0..m) into start[0..m], one block
__global__ void __launch_bounds__(1024) k_pr_scan(const unsigned int *__restrict__ cnt, int m, int *start) {
    __shared__ int s_part[1024];
    const int t = threadIdx.x, per = (m + 1023) / 1024;
    const int b0 = t * per, b1 = min(m, b0 + per);
    int sum = 0;
    for (int b = b0; b < b1; b++) sum += (int)cnt[b];
    s_part[t] = sum;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const int v = t >= d ? s_part[t - d] : 0;
        __syncthreads();
        s_part[t] += v;
        __syncthreads();
    }
    int run = s_part[t] - sum;
    for (int b = b0; b < b1; b++) {
        start[b] = run;
        run += (int)cnt[b];
    }
    if (t == 1023) start[m] = s_part[1023];
}

__global__ void __launch_bounds__(PR_THREADS) k_cg_scatter(const int *__restrict__ cell, int n,
                                                           const int *__restrict__ start, unsigned int *fill,
                                                           int *mem) {
    const int i = blockIdx.x * PR_THREADS + threadIdx.x;
    if (i >= n) return;
    const int c = cell[i];
    mem[start[c] + (int)atomicAdd(&fill[c], 1u)] = i;
}

// len_check of overlap_cage_rel.c:253-269
__device__ __forceinline__ bool cg_near(double ax, double ay, double az, double bx, double by, double bz, double L) {
    double dx = ax - bx, dy = ay - by, dz = az - bz;
    if (dx < -0.5) dx += 1.0;
    if (dx > 0.5) dx -= 1.0;
    if (dy < -0.5) dy += 1.0;
    if (dy > 0.5) dy -= 1.0;
    if (dz < -0.5) dz += 1.0;
    if (dz > 0.5) dz -= 1.0;
    const double dr2 = (dx * dx + dy * dy + dz * dz) * L * L;
    return dr2 < 1.3 * 1.3;
}

// the 13 half-stencil offsets of maps() (overlap_cage_rel.c:297-309)
__constant__ int c_hs[13][3] = {{1, 0, 0},  {1, 1, 0},  {0, 1, 0},  {-1, 1, 0}, {1, 0, -1}, {1, 1, -1}, {0, 1, -1},
                                {-1, 1, -1}, {1, 0, 1}, {1, 1, 1},  {0, 1, 1},  {-1, 1, 1}, {0, 0, 1}};

__device__ __forceinline__ int cg_cell_at(int c, int dx, int dy, int dz, int nc) {  // getcell_I (:273-277)
    const int q = c - 1;
    const int ix = q % nc, iy = (q / nc) % nc, iz = q / (nc * nc);
    return 1 + (ix + dx + nc) % nc + ((iy + dy + nc) % nc) * nc + ((iz + dz + nc) % nc) * nc * nc;
}

// Neighbour list of particle i in GetList's append order.  keys: group 0 =
// found by j < i, 1 = i's own scan (sub = stencil slot + 1, own cell 0),
// 2 = found by j > i; sorted by (group, sub, index).
__global__ void __launch_bounds__(PR_THREADS) k_cg_list(const double *__restrict__ wx, const double *__restrict__ wy,
                                                        const double *__restrict__ wz, const int *__restrict__ cell,
                                                        int n, int nc, double L, const int *__restrict__ start,
                                                        const int *__restrict__ mem, int *nn, int *nb,
                                                        unsigned long long *key, int *overflow) {
    const int i = blockIdx.x * PR_THREADS + threadIdx.x;
    if (i >= n) return;
    const double xi = wx[i], yi = wy[i], zi = wz[i];
    const int c = cell[i];
    int k = 0;
    int *lst = nb + (size_t)i * CG_CAP;
    unsigned long long *ky = key + (size_t)i * CG_CAP;
    auto put = [&](unsigned long long kk, int j) {
        if (k < CG_CAP) {
            ky[k] = kk;
            lst[k] = j;
        }
        k++;
    };
    for (int q = start[c]; q < start[c + 1]; q++) {  // own cell (:221-233)
        const int j = mem[q];
        if (j == i || !cg_near(xi, yi, zi, wx[j], wy[j], wz[j], L)) continue;
        put(j < i ? (unsigned long long)j : (1ull << 60) | (unsigned long long)j, j);
    }
    for (int s = 0; s < 13; s++) {
        const int c2 = cg_cell_at(c, c_hs[s][0], c_hs[s][1], c_hs[s][2], nc);  // i's scan (:235-248)
        for (int q = start[c2]; q < start[c2 + 1]; q++) {
            const int j = mem[q];
            if (cg_near(xi, yi, zi, wx[j], wy[j], wz[j], L))
                put((1ull << 60) | ((unsigned long long)(s + 1) << 32) | (unsigned long long)j, j);
        }
        const int c3 = cg_cell_at(c, -c_hs[s][0], -c_hs[s][1], -c_hs[s][2], nc);  // scans that reach i
        for (int q = start[c3]; q < start[c3 + 1]; q++) {
            const int j = mem[q];
            if (cg_near(wx[j], wy[j], wz[j], xi, yi, zi, L))
                put(j < i ? (unsigned long long)j : (2ull << 60) | (unsigned long long)j, j);
        }
    }
    if (k > CG_CAP) {
        atomicMax(overflow, k);
        k = CG_CAP;
    }
    for (int a = 1; a < k; a++) {  // insertion sort by key
        const unsigned long long kk = ky[a];
        const int jj = lst[a];
        int b = a - 1;
        while (b >= 0 && ky[b] > kk) {
            ky[b + 1] = ky[b];
            lst[b + 1] = lst[b];
            b--;
        }
        ky[b + 1] = kk;
        lst[b + 1] = jj;
    }
    nn[i] = k;
}

// DX = L * (one-correction minimum image of x_t - x0), unwrapped (:80-100);
// grid (particle blocks, lags of the origin)
__global__ void __launch_bounds__(PR_THREADS) k_cg_disp(const double *__restrict__ X, const double *__restrict__ Y,
                                                        const double *__restrict__ Z, int n, size_t f0,
                                                        const int *__restrict__ lagf, double L, double *DX,
                                                        double *DY, double *DZ) {
    const int i = blockIdx.x * PR_THREADS + threadIdx.x;
    if (i >= n) return;
    const size_t ft = (size_t)lagf[blockIdx.y] * n, o = (size_t)blockIdx.y * n + i;
    DX[o] = L * pr_mi(X[ft + i] - X[f0 + i]);
    DY[o] = L * pr_mi(Y[ft + i] - Y[f0 + i]);
    DZ[o] = L * pr_mi(Z[ft + i] - Z[f0 + i]);
}

// cage-relative overlap count of one lag (:109-131); grid (particle blocks, lags)
__global__ void __launch_bounds__(PR_THREADS) k_cg_count(const double *__restrict__ DX, const double *__restrict__ DY,
                                                         const double *__restrict__ DZ, int n,
                                                         const int *__restrict__ nn, const int *__restrict__ nb,
                                                         double param, unsigned int *count) {
    const int i = blockIdx.x * PR_THREADS + threadIdx.x;
    const size_t o = (size_t)blockIdx.y * n;
    int hit = 0;
    if (i < n) {
        const int m = nn[i];
        const int *lst = nb + (size_t)i * CG_CAP;
        double dx = 0.0, dy = 0.0, dz = 0.0;
        for (int k = 0; k < m; k++) {
            const size_t p = o + lst[k];
            dx += DX[p];
            dy += DY[p];
            dz += DZ[p];
        }
        dx = dx / (double)m;
        dy = dy / (double)m;
        dz = dz / (double)m;
        const double Xi = DX[o + i], Yi = DY[o + i], Zi = DZ[o + i];
        double r2 = (Xi * Xi + Yi * Yi + Zi * Zi);
        r2 += (dx * dx + dy * dy + dz * dz);
        r2 -= 2.00 * (Xi * dx + Yi * dy + Zi * dz);
        hit = r2 < param ? 1 : 0;
    }
    const unsigned long long b = __ballot(hit);
    __shared__ unsigned int s_c;
    if (threadIdx.x == 0) s_c = 0;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) atomicAdd(&s_c, (unsigned int)__popcll(b));
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(&count[blockIdx.y], s_c);
}

// ------------------------------------------------------------------- g_uu

// active particles m (ascending index act[m]): displacement to the lag frame
// (guu.c:151-180) and wrapped origin position (LoadOrigin)
__global__ void __launch_bounds__(PR_THREADS) k_gu_disp(const double *__restrict__ X, const double *__restrict__ Y,
                                                        const double *__restrict__ Z, int n, size_t f0, size_t ft,
                                                        const int *__restrict__ act, int na, double L, double *g) {
    const int m = blockIdx.x * PR_THREADS + threadIdx.x;
    if (m >= na) return;
    const int i = act[m];
    const double dx = pr_mi(X[ft + i] - X[f0 + i]), dy = pr_mi(Y[ft + i] - Y[f0 + i]),
                 dz = pr_mi(Z[ft + i] - Z[f0 + i]);
    const double r2 = (dx * dx + dy * dy + dz * dz);
    g[0 * (size_t)na + m] = pr_wrap(X[f0 + i]);
    g[1 * (size_t)na + m] = pr_wrap(Y[f0 + i]);
    g[2 * (size_t)na + m] = pr_wrap(Z[f0 + i]);
    g[3 * (size_t)na + m] = L * sqrt(r2);  // drt
    g[4 * (size_t)na + m] = L * dx;
    g[5 * (size_t)na + m] = L * dy;
    g[6 * (size_t)na + m] = L * dz;
    g[7 * (size_t)na + m] = r2;
}

// MSDtPP += L*L*r2, MDtPP += r in particle order, carried over origins (guu.c:172-173)
__global__ void k_gu_sums(const double *__restrict__ g, int na, double L, double *acc) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double msd = acc[0], md = acc[1];
    for (int m = 0; m < na; m++) {
        msd += L * L * g[7 * (size_t)na + m];
        md += g[3 * (size_t)na + m];
    }
    acc[0] = msd;
    acc[1] = md;
}

// pairs (row m1 < column m2) of one tile of the upper triangle; bins [b0, b0+nb)
__global__ void __launch_bounds__(GU_TILE) k_gu_hist(const double *__restrict__ g, int na, double L, double bin,
                                                     double maxbin, int b0, int nbins, int ntile,
                                                     unsigned long long *hc, double *h1, double *h2) {
    extern __shared__ unsigned char s_raw[];
    unsigned int *s_c = (unsigned int *)s_raw;
    double *s_1 = (double *)(s_raw + ((sizeof(unsigned int) * nbins + 15) & ~(size_t)15));
    double *s_2 = s_1 + nbins;
    __shared__ double s_col[7][GU_TILE];
    // tile (tr, tc), tr <= tc, from the linear block index
    int t = blockIdx.x, tr = 0;
    while (t >= ntile - tr) {
        t -= ntile - tr;
        tr++;
    }
    const int tc = tr + t;
    for (int b = threadIdx.x; b < nbins; b += GU_TILE) {
        s_c[b] = 0;
        s_1[b] = 0.0;
        s_2[b] = 0.0;
    }
    const int cj = tc * GU_TILE + threadIdx.x;
    if (cj < na) {
        s_col[0][threadIdx.x] = g[0 * (size_t)na + cj];
        s_col[1][threadIdx.x] = g[1 * (size_t)na + cj];
        s_col[2][threadIdx.x] = g[2 * (size_t)na + cj];
        s_col[3][threadIdx.x] = g[3 * (size_t)na + cj];
        s_col[4][threadIdx.x] = g[4 * (size_t)na + cj];
        s_col[5][threadIdx.x] = g[5 * (size_t)na + cj];
        s_col[6][threadIdx.x] = g[6 * (size_t)na + cj];
    }
    __syncthreads();
    const int mi = tr * GU_TILE + threadIdx.x;
    if (mi < na) {
        const double xi = g[mi], yi = g[(size_t)na + mi], zi = g[2 * (size_t)na + mi];
        const double ri = g[3 * (size_t)na + mi], ui = g[4 * (size_t)na + mi], vi = g[5 * (size_t)na + mi],
                     wi = g[6 * (size_t)na + mi];
        const int jn = min(GU_TILE, na - tc * GU_TILE);
        for (int jj = 0; jj < jn; jj++) {
            const int mj = tc * GU_TILE + jj;
            if (mj <= mi) continue;
            // guu.c:198-213: x0_j - x0_i, one correction
            const double dx = pr_mi(s_col[0][jj] - xi), dy = pr_mi(s_col[1][jj] - yi), dz = pr_mi(s_col[2][jj] - zi);
            const double r2 = L * L * (dx * dx + dy * dy + dz * dz);
            const int index = (int)(sqrt(r2) / bin);
            if (index < maxbin && index >= b0 && index < b0 + nbins) {
                const int b = index - b0;
                atomicAdd(&s_c[b], 1u);
                atomicAdd(&s_1[b], 2.0 * (ri * s_col[3][jj]));
                atomicAdd(&s_2[b], 2.0 * (ui * s_col[4][jj] + vi * s_col[5][jj] + wi * s_col[6][jj]));
            }
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nbins; b += GU_TILE)
        if (s_c[b]) {
            atomicAdd(&hc[b0 + b], (unsigned long long)s_c[b]);
            atomicAdd(&h1[b0 + b], s_1[b]);
            atomicAdd(&h2[b0 + b], s_2[b]);
        }
}

#define PTRY(e) do { int rc_ = (e); if (rc_ != AC4_OK) return rc_; } while (0)

extern "C" int ac4_frames_overlap_cage(ac4_frames *f, int gap, double factor, double a, double cell_cut,
                                       const char *path, int *pair_counts, int cap_pairs, int *npairs) {
    ac4_frames_view v;
    PTRY(ac4_frames_view_get(f, &v));
    AC4_REQUIRE(gap > 0 && factor > 1.0 && a > 0.0 && cell_cut > 0.0 && v.count >= 2, AC4_EARG,
                "ac4_frames_overlap_cage: bad arguments");
    const int n = v.n, nf = v.count;
    const double L = v.L;
    std::vector<int> fidx;
    PTRY(ac4_frames_steps(v, fidx));
    const int nc = (int)(L / cell_cut);  // maps(), :283
    AC4_REQUIRE(nc >= 3, AC4_EARG,
                "ac4_frames_overlap_cage: %d cells per edge (L/cell_cut); at least 3 are needed", nc);
    const int ncell = nc * nc * nc;
    const double cs = 1.0 / (double)nc;
    ac4_walk_t w;
    memset(&w, 0, sizeof w);
    PTRY(ac4_walk(fidx.data(), nf, gap, factor, 1, &w));  // origins counted on k*gap+1 (:23-30)
    const int maxint = w.maximal_interval;
    std::vector<int> cur(w.cur, w.cur + w.np), oth(w.oth, w.oth + w.np), slot(w.slot, w.slot + w.np);
    ac4_walk_free(&w);
    AC4_REQUIRE(maxint > 0, AC4_EARG, "ac4_frames_overlap_cage: no lag");
    const int noi = 10 + (int)(log((double)maxint) / log(factor));
    const int np = (int)cur.size();
    int maxlag = 0;
    for (int p = 0, q; p < np; p = q) {
        for (q = p; q < np && cur[q] == cur[p]; q++) {}
        maxlag = std::max(maxlag, q - p);
    }
    AcDevBuf B;
    double *d_w, *d_D;
    int *d_cell, *d_start, *d_mem, *d_nn, *d_nb, *d_lag, *d_over;
    unsigned int *d_cnt, *d_count;
    unsigned long long *d_key;
    PTRY(B.get(&d_w, (size_t)3 * n));
    PTRY(B.get(&d_D, (size_t)3 * n * std::max(maxlag, 1)));
    PTRY(B.get(&d_cell, n));
    PTRY(B.get(&d_cnt, (size_t)2 * (ncell + 1)));
    PTRY(B.get(&d_start, ncell + 2));
    PTRY(B.get(&d_mem, n));
    PTRY(B.get(&d_nn, n));
    PTRY(B.get(&d_nb, (size_t)n * CG_CAP));
    PTRY(B.get(&d_key, (size_t)n * CG_CAP));
    PTRY(B.get(&d_lag, std::max(maxlag, 1)));
    PTRY(B.get(&d_over, 1));
    PTRY(B.get(&d_count, std::max(np, 1)));
    hipStream_t s = v.s;
    AC4_HIP(hipMemsetAsync(d_count, 0, sizeof(unsigned int) * std::max(np, 1), s));
    AC4_HIP(hipMemsetAsync(d_over, 0, sizeof(int), s));
    const unsigned gb = (unsigned)((n + PR_THREADS - 1) / PR_THREADS);
    // origins visited by the do-while of :59-154; a pair's origin is its cur
    for (int p = 0, q; p < np; p = q) {
        for (q = p; q < np && cur[q] == cur[p]; q++) {}
        const size_t f0 = (size_t)cur[p] * n;
        AC4_HIP(hipMemsetAsync(d_cnt, 0, sizeof(unsigned int) * 2 * (ncell + 1), s));
        hipLaunchKernelGGL(k_cg_bin, dim3(gb), dim3(PR_THREADS), 0, s, v.x + f0, v.y + f0, v.z + f0, n, nc, cs, d_w,
                           d_w + n, d_w + 2 * (size_t)n, d_cell, d_cnt);
        hipLaunchKernelGGL(k_pr_scan, dim3(1), dim3(1024), 0, s, d_cnt, ncell + 1, d_start);
        hipLaunchKernelGGL(k_cg_scatter, dim3(gb), dim3(PR_THREADS), 0, s, d_cell, n, d_start, d_cnt + ncell + 1,
                           d_mem);
        hipLaunchKernelGGL(k_cg_list, dim3(gb), dim3(PR_THREADS), 0, s, d_w, d_w + n, d_w + 2 * (size_t)n, d_cell, n,
                           nc, L, d_start, d_mem, d_nn, d_nb, d_key, d_over);
        std::vector<int> lags(oth.begin() + p, oth.begin() + q);
        AC4_HIP(hipMemcpyAsync(d_lag, lags.data(), sizeof(int) * lags.size(), hipMemcpyHostToDevice, s));
        const size_t nl = lags.size();
        hipLaunchKernelGGL(k_cg_disp, dim3(gb, (unsigned)nl), dim3(PR_THREADS), 0, s, v.x, v.y, v.z, n, f0, d_lag, L,
                           d_D, d_D + nl * n, d_D + 2 * nl * n);
        hipLaunchKernelGGL(k_cg_count, dim3(gb, (unsigned)nl), dim3(PR_THREADS), 0, s, d_D, d_D + nl * n,
                           d_D + 2 * nl * n, n, d_nn, d_nb, a * a, d_count + p);
        AC4_HIP(hipGetLastError());
        AC4_HIP(hipStreamSynchronize(s));  // lags and scratch are reused by the next origin
    }
    std::vector<unsigned int> cnt(std::max(np, 1));
    int over = 0;
    AC4_HIP(hipMemcpyAsync(cnt.data(), d_count, sizeof(unsigned int) * std::max(np, 1), hipMemcpyDeviceToHost, s));
    AC4_HIP(hipMemcpyAsync(&over, d_over, sizeof(int), hipMemcpyDeviceToHost, s));
    AC4_HIP(hipStreamSynchronize(s));
    AC4_REQUIRE(over == 0, AC4_ECAP, "ac4_frames_overlap_cage: %d neighbours within 1.3 (at most %d)", over, CG_CAP);
    // :132-167: SUM = count/N per pair, accumulated per slot in walk order
    std::vector<double> c(noi, 0.0), x4(noi, 0.0), dt(noi, 0.0);
    std::vector<int> counters(noi, 0);
    for (int p = 0; p < np; p++) {
        const int t = slot[p];
        AC4_REQUIRE(t < noi, AC4_EARG, "ac4_frames_overlap_cage: slot %d beyond %d", t, noi);
        const double SUM = (double)cnt[p] / (double)n;
        c[t] += SUM;
        x4[t] += SUM * SUM;
        counters[t]++;
        dt[t] = PR_TIME_STEP * (double)(fidx[oth[p]] - fidx[cur[p]]);
        if (pair_counts && p < cap_pairs) pair_counts[p] = (int)cnt[p];
    }
    if (npairs) *npairs = np;
    if (path) {
        FILE *fo = fopen(path, "wb");
        AC4_REQUIRE(fo, AC4_EIO, "ac4_frames_overlap_cage: cannot write %s", path);
        fprintf(fo, "0.0\t1.0\t0.0\n");
        for (int k = 1; k < noi; k++)
            if (counters[k])
                fprintf(fo, "%.14lf\t%.14lf\t%.14lf\n", dt[k], c[k] / (double)counters[k],
                        x4[k] / (double)counters[k] - c[k] * c[k] / ((double)(counters[k] * counters[k])));
        AC4_REQUIRE(fclose(fo) == 0, AC4_EIO, "ac4_frames_overlap_cage: write error on %s", path);
    }
    return AC4_OK;
}

extern "C" int ac4_frames_guu(ac4_frames *f, const int *act_tag, int gap, double factor, double tau,
                              const char *path, double *hist3, int *maxbin_out, int *frame_num_out,
                              int *n_origins_out) {
    ac4_frames_view v;
    PTRY(ac4_frames_view_get(f, &v));
    AC4_REQUIRE(act_tag && gap > 0 && factor > 1.0 && v.count >= 2, AC4_EARG, "ac4_frames_guu: bad arguments");
    const int n = v.n, nf = v.count;
    const double L = v.L;
    std::vector<int> fidx;
    PTRY(ac4_frames_steps(v, fidx));
    const double bin = 0.005;
    const double MaxBin = (int)(0.5 * L / bin);  // guu.c:17
    const int maxbin = (int)MaxBin;
    const int maxint = 200 * gap / PR_DIVIDE_TO;  // guu.c:30-31
    // guu.c:67-115: lag ladder of the first frame, FrameNum nearest tau
    int frame_num = 0;
    {
        std::vector<double> dt(1, 0.0);
        const int cf = fidx[0];
        int other = 0, offset = 1;
        while (other < nf && offset < maxint) {
            const int of = cf + offset;
            while (other < nf && fidx[other] != of) other++;
            if (other < nf) {
                dt.push_back(PR_TIME_STEP * (double)(of - cf));
                if ((int)(offset * factor) == offset)
                    offset++;
                else
                    offset = (int)(((double)offset) * factor);
            }
        }
        double inter = 1000.0;
        for (size_t i = 1; i < dt.size(); i++)
            if (fabs(dt[i] - tau) < inter) {
                frame_num = (int)(dt[i] * 200.0);
                inter = fabs(dt[i] - tau);
            }
    }
    std::vector<int> act;
    for (int i = 0; i < n; i++)
        if (act_tag[i] == 1) act.push_back(i);
    const int na = (int)act.size();
    AcDevBuf B;
    int *d_act;
    double *d_g, *d_acc, *d_h1, *d_h2;
    unsigned long long *d_hc;
    PTRY(B.get(&d_act, std::max(na, 1)));
    PTRY(B.get(&d_g, (size_t)8 * std::max(na, 1)));
    PTRY(B.get(&d_acc, 2));
    PTRY(B.get(&d_hc, std::max(maxbin, 1)));
    PTRY(B.get(&d_h1, std::max(maxbin, 1)));
    PTRY(B.get(&d_h2, std::max(maxbin, 1)));
    hipStream_t s = v.s;
    if (na) AC4_HIP(hipMemcpyAsync(d_act, act.data(), sizeof(int) * na, hipMemcpyHostToDevice, s));
    AC4_HIP(hipMemsetAsync(d_acc, 0, sizeof(double) * 2, s));
    AC4_HIP(hipMemsetAsync(d_hc, 0, sizeof(unsigned long long) * std::max(maxbin, 1), s));
    AC4_HIP(hipMemsetAsync(d_h1, 0, sizeof(double) * std::max(maxbin, 1), s));
    AC4_HIP(hipMemsetAsync(d_h2, 0, sizeof(double) * std::max(maxbin, 1), s));
    const int ntile = (na + GU_TILE - 1) / GU_TILE;
    const long long ntri = (long long)ntile * (ntile + 1) / 2;
    const int pass_bins = std::min(maxbin, GU_LDS_BINS);
    const size_t lds = ((sizeof(unsigned int) * pass_bins + 15) & ~(size_t)15) + sizeof(double) * 2 * pass_bins;
    AC4_HIP(hipFuncSetAttribute((const void *)k_gu_hist, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int current = 0, on = 0;
    do {  // guu.c:138-231
        const int cf = fidx[current];
        int other = current;
        while (other < nf && fidx[other] != cf + frame_num) other++;
        if (other >= nf) break;
        if (na) {
            hipLaunchKernelGGL(k_gu_disp, dim3((na + PR_THREADS - 1) / PR_THREADS), dim3(PR_THREADS), 0, s, v.x, v.y,
                               v.z, n, (size_t)current * n, (size_t)other * n, d_act, na, L, d_g);
            hipLaunchKernelGGL(k_gu_sums, dim3(1), dim3(64), 0, s, d_g, na, L, d_acc);
            for (int b0 = 0; b0 < maxbin && ntri > 0; b0 += pass_bins)
                hipLaunchKernelGGL(k_gu_hist, dim3((unsigned)ntri), dim3(GU_TILE), lds, s, d_g, na, L, bin, MaxBin,
                                   b0, std::min(pass_bins, maxbin - b0), ntile, d_hc, d_h1, d_h2);
            AC4_HIP(hipGetLastError());
        }
        on++;
        while (current < nf && fidx[current] != on * gap) current++;
    } while (current < nf);
    std::vector<unsigned long long> hc(std::max(maxbin, 1));
    std::vector<double> h1(std::max(maxbin, 1)), h2(std::max(maxbin, 1));
    double acc[2];
    AC4_HIP(hipMemcpyAsync(hc.data(), d_hc, sizeof(unsigned long long) * maxbin, hipMemcpyDeviceToHost, s));
    AC4_HIP(hipMemcpyAsync(h1.data(), d_h1, sizeof(double) * maxbin, hipMemcpyDeviceToHost, s));
    AC4_HIP(hipMemcpyAsync(h2.data(), d_h2, sizeof(double) * maxbin, hipMemcpyDeviceToHost, s));
    AC4_HIP(hipMemcpyAsync(acc, d_acc, sizeof acc, hipMemcpyDeviceToHost, s));
    AC4_HIP(hipStreamSynchronize(s));
    // guu.c:250-276
    const double MSDt = acc[0] / ((double)n * (double)on);
    const double MDt = acc[1] / ((double)n * (double)on);
    const double norm = (4.0 * PR_PI * (PR_RHO)*bin * (double)n * (double)on);
    const double norm1 = (4.0 * PR_PI * PR_RHO * bin * (double)n * (double)on * (double)(MDt * MDt));
    const double norm2 = (4.0 * PR_PI * PR_RHO * bin * (double)n * (double)on * (double)(MSDt));
    if (hist3)
        for (int i = 0; i < maxbin; i++) {
            hist3[i] = 2.0 * (double)hc[i];
            hist3[maxbin + i] = h1[i];
            hist3[2 * (size_t)maxbin + i] = h2[i];
        }
    if (maxbin_out) *maxbin_out = maxbin;
    if (frame_num_out) *frame_num_out = frame_num;
    if (n_origins_out) *n_origins_out = on;
    if (path) {
        FILE *fo = fopen(path, "wb");
        AC4_REQUIRE(fo, AC4_EIO, "ac4_frames_guu: cannot write %s", path);
        for (int i = 1; i < MaxBin; i++) {
            const double r = (double)(i + 1) * bin;
            const double hist = 2.0 * (double)hc[i];
            fprintf(fo, "%0.10g\t%.10g\t%.10g\t%.10g\t%.10g\n", r, (hist / (norm * r * r)), (h1[i] / (norm1 * r * r)),
                    h2[i] / (norm1 * r * r), h2[i] / (norm2 * r * r));
        }
        AC4_REQUIRE(fclose(fo) == 0, AC4_EIO, "ac4_frames_guu: write error on %s", path);
    }
    return AC4_OK;
}
