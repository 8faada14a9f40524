// ac4_block.hip -- block (sub-box) analysis of a device frame store.
//
// Replaces Analysis_Block/msd_pos.c, overlap.c and vanHove.c (paths relative
// to the reference checkout) for one block edge.  The box is cut into ncx^3
// blocks, ncx = (int)(L / cell_cut), and each particle is tagged with the
// block of its wrapped origin position.  Per (origin, lag) pair and block:
//   - the overlap count #{i : L*L*(dx*dx+dy*dy+dz*dz) < a*a} on wrapped
//     positions with one minimum-image correction (overlap.c:110-143);
//   - the sum of those r^2 (msd_pos.c:131-167);
//   - the block centre-of-mass displacement sum(x + correction) - sum(x0)
//     (msd_pos.c:104-150, vanHove.c:97-143).
// Outputs, written by the host in the reference's formats:
//   BlockAnalysis/MSD/MSDLb<Lb>.dat, BlockAnalysis/Relax/RelaxLb<Lb>.dat,
//   BlockAnalysis/Relax/RelaxTotLb<Lb>.dat, BlockAnalysis/vanHove<Lb>.dat.
//
// Kernels:
//   k_blk_count, k_blk_offsets, k_blk_place  block ids and a stable counting
//               sort per origin (16 waves each), so each block's members are
//               listed in ascending particle index;
//   k_blk_pair  one thread per (pair, block): walks the block's members in
//               that order, so every sum is formed in the reference's order
//               and is bit-identical to it;
//   k_blk_slot  one thread per (slot, block): accumulates the pairs of a lag
//               slot in origin order (overlap.c:144-148, msd_pos.c:172);
//   k_blk_tot   one wave per slot and sum: the block-summed accumulators
//               (overlap.c:149-151, msd_pos.c:173-179), whose addends
//               k_blk_pair has formed, added one after another in
//               pair-then-block order (64 addends per coalesced load, taken
//               lane by lane with readlane: the chain stays sequential).
// The final averages over blocks run on the host in the reference's order.
//
// Empty blocks: the reference divides by a zero block count and carries the
// resulting NaN into its output (printed "-nan"); the host gives every NaN
// the bit pattern an x86-64 division 0.0/0.0 returns, so the text matches.
//
// vanHove.c reads uninitialised locals: its lag ladder (:31-44) is taken from
// the first frame with offset 1 and slot 0, only the slots it fills are
// searched (:46-51), and FrameNum is 0 when no lag lies within 1000 time
// units of Time (each origin is then paired with itself), which is what the
// reference binary built from those sources does.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <sys/stat.h>
#include <vector>
#include "ac4_internal.h"

#define BLK_THREADS 256
#define BLK_SORT_MAX 16384     // blocks whose cursors fit the sort's 64 KB of LDS
#define BLK_SORT_WAVES 16      // chunks of consecutive particles per origin
#define BLK_TIME_STEP 0.005    // Analysis_Block/global.h:14
#define BLK_DIVIDE_TO 2.0      // Analysis_Block/global.h:16

// The while-loop wrap of msd_pos.c:86-91.  Each x-1 step is exact for x >= 1
// and each x+1 step is exact until the last, which rounds once, so the loops
// equal x - floor(x) rounded once; values in [0,1) (and -0) stay as they are.
__device__ __forceinline__ double blk_wrap(double v) { return (v >= 1.0 || v < 0.0) ? v - floor(v) : v; }

// msd_pos.c:95-100: block of a wrapped position.  A wrapped 1.0 (an input
// within 1e-16 below zero) would index past the grid in the reference; it is
// clamped into the last block here.
__device__ __forceinline__ int blk_cell(double x, double y, double z, int ncx, int ncell) {
    const double inv = (double)ncx;
    const double c = floor(x * inv) + floor(y * inv) * ncx + floor(z * inv) * (ncx * ncx);
    const int ci = (int)c;
    return ci < 0 ? 0 : (ci >= ncell ? ncell - 1 : ci);
}

__device__ __forceinline__ double blk_mi(double d) {  // overlap.c:127-138
    if (d >= 0.5)
        d -= 1.0;
    else if (d < -0.5)
        d += 1.0;
    return d;
}

__device__ __forceinline__ void blk_corr(double &acc, double d) {  // msd_pos.c:139-150
    if (d >= 0.5)
        acc -= 1.0;
    else if (d < -0.5)
        acc += 1.0;
}

// Stable counting sort by block, per origin, in three passes over
// BLK_SORT_WAVES chunks of consecutive particles: k_blk_count counts each
// chunk's particles per block, k_blk_offsets turns the counts into the
// first slot of every (chunk, block) and start[o][b], k_blk_place writes the
// members.  Chunks are placed in chunk order and, inside a chunk, lanes of
// the same block in lane order, so members[o][start[o][b] ...] lists block
// b's particles in ascending index (start[o][ncell] = n).
__global__ void __launch_bounds__(64) k_blk_count(const double *__restrict__ X, const double *__restrict__ Y,
                                                  const double *__restrict__ Z, int n, const int *__restrict__ ofr,
                                                  int ncx, int ncell, int chunk, int *__restrict__ cnt) {
    extern __shared__ int s_cur[];
    const int o = blockIdx.x, w = blockIdx.y, lane = threadIdx.x;
    const size_t fb = (size_t)ofr[o] * n;
    for (int b = lane; b < ncell; b += 64) s_cur[b] = 0;
    __syncthreads();
    const int i1 = min(n, (w + 1) * chunk);
    for (int i = w * chunk + lane; i < i1; i += 64)
        atomicAdd(&s_cur[blk_cell(blk_wrap(X[fb + i]), blk_wrap(Y[fb + i]), blk_wrap(Z[fb + i]), ncx, ncell)], 1);
    __syncthreads();
    int *c = cnt + ((size_t)o * gridDim.y + w) * ncell;
    for (int b = lane; b < ncell; b += 64) c[b] = s_cur[b];
}

__global__ void __launch_bounds__(256) k_blk_offsets(int ncell, int nw, int n, int *__restrict__ cnt,
                                                     int *__restrict__ start) {
    __shared__ int s_x[256];
    const int o = blockIdx.x, t = threadIdx.x;
    int *c = cnt + (size_t)o * nw * ncell;
    int *st = start + (size_t)o * (ncell + 1);
    int run = 0;
    for (int b0 = 0; b0 < ncell; b0 += 256) {
        const int b = b0 + t;
        int tot = 0;
        if (b < ncell)
            for (int w = 0; w < nw; w++) tot += c[(size_t)w * ncell + b];
        s_x[t] = tot;
        __syncthreads();
        for (int d = 1; d < 256; d <<= 1) {
            const int u = t >= d ? s_x[t - d] : 0;
            __syncthreads();
            s_x[t] += u;
            __syncthreads();
        }
        if (b < ncell) {
            int acc = run + s_x[t] - tot;
            st[b] = acc;
            for (int w = 0; w < nw; w++) {
                const int k = c[(size_t)w * ncell + b];
                c[(size_t)w * ncell + b] = acc;
                acc += k;
            }
        }
        run += s_x[255];
        __syncthreads();
    }
    if (t == 0) st[ncell] = n;
}

__global__ void __launch_bounds__(64) k_blk_place(const double *__restrict__ X, const double *__restrict__ Y,
                                                  const double *__restrict__ Z, int n, const int *__restrict__ ofr,
                                                  int ncx, int ncell, int chunk, const int *__restrict__ cur,
                                                  int *__restrict__ members) {
    extern __shared__ int s_cur[];
    const int o = blockIdx.x, w = blockIdx.y, lane = threadIdx.x;
    const size_t fb = (size_t)ofr[o] * n;
    const int *c = cur + ((size_t)o * gridDim.y + w) * ncell;
    for (int b = lane; b < ncell; b += 64) s_cur[b] = c[b];
    __syncthreads();
    int *mem = members + (size_t)o * n;
    const unsigned long long below = (1ull << lane) - 1ull;
    const int i1 = min(n, (w + 1) * chunk);
    for (int i0 = w * chunk; i0 < i1; i0 += 64) {
        const int i = i0 + lane;
        const int b = i < i1 ? blk_cell(blk_wrap(X[fb + i]), blk_wrap(Y[fb + i]), blk_wrap(Z[fb + i]), ncx, ncell) : -1;
        unsigned long long todo = __ballot(i < i1);
        while (todo) {  // one block at a time; lanes of that block keep lane order
            const int leader = __ffsll((long long)todo) - 1;
            const int bl = __shfl(b, leader);
            const unsigned long long m = __ballot(b == bl);
            int base = lane == leader ? s_cur[bl] : 0;
            base = __shfl(base, leader);
            if (b == bl) mem[base + __popcll(m & below)] = i;
            if (lane == leader) s_cur[bl] = base + __popcll(m);
            todo &= ~m;
        }
    }
}

// One thread per (pair, block); pairs[p] = (origin, lag frame).
__global__ void __launch_bounds__(BLK_THREADS) k_blk_pair(
    const double *__restrict__ X, const double *__restrict__ Y, const double *__restrict__ Z, int n,
    const int2 *__restrict__ pairs, const int *__restrict__ ofr, long long total, int ncell,
    const int *__restrict__ start, const int *__restrict__ members, double L, double param, int *__restrict__ ov,
    double *__restrict__ s2, double *__restrict__ ex, double *__restrict__ ey, double *__restrict__ ez,
    double *__restrict__ add) {
    const long long t = (long long)blockIdx.x * BLK_THREADS + threadIdx.x;
    if (t >= total) return;
    const int p = (int)(t / ncell), b = (int)(t - (long long)p * ncell);
    const int2 pr = pairs[p];
    const size_t f0 = (size_t)ofr[pr.x] * n, f1 = (size_t)pr.y * n;
    const int *st = start + (size_t)pr.x * (ncell + 1);
    const int *mem = members + (size_t)pr.x * n;
    const int k1 = st[b + 1];
    int cnt = 0;
    double s = 0.0, cx = 0.0, cy = 0.0, cz = 0.0, ax = 0.0, ay = 0.0, az = 0.0;
#pragma unroll 4
    for (int k = st[b]; k < k1; k++) {
        const int i = mem[k];
        const double x0 = blk_wrap(X[f0 + i]), y0 = blk_wrap(Y[f0 + i]), z0 = blk_wrap(Z[f0 + i]);
        const double x = blk_wrap(X[f1 + i]), y = blk_wrap(Y[f1 + i]), z = blk_wrap(Z[f1 + i]);
        double dx = x - x0, dy = y - y0, dz = z - z0;
        cx += x;
        cy += y;
        cz += z;
        blk_corr(cz, dz);
        blk_corr(cy, dy);
        blk_corr(cx, dx);
        ax += x0;
        ay += y0;
        az += z0;
        dz = blk_mi(dz);
        dy = blk_mi(dy);
        dx = blk_mi(dx);
        const double r2 = L * L * (dx * dx + dy * dy + dz * dz);
        s += r2;
        cnt += r2 < param ? 1 : 0;
    }
    ov[t] = cnt;
    s2[t] = s;
    ex[t] = cx - ax;
    ey[t] = cy - ay;
    ez[t] = cz - az;
    // the pair's addends to the block-summed accumulators: overlap.c:149,
    // msd_pos.c:173-177 (NumInCell = the member count; 0 gives NaN, as there)
    const double nb = (double)(k1 - st[b]);
    const double dx = (cx - ax) / nb, dy = (cy - ay) / nb, dz = (cz - az) / nb;
    add[t] = (double)cnt / nb;
    add[total + t] = L * L * (dx * dx + dy * dy + dz * dz);
    add[2 * total + t] = L * L * L * L * (dx * dx * dx * dx + dy * dy * dy * dy + dz * dz * dz * dz);
}

// One thread per (slot, block): the pairs of the slot in walk (origin) order.
__global__ void __launch_bounds__(BLK_THREADS) k_blk_slot(int noi, int ncell, const int *__restrict__ sptr,
                                                          const int *__restrict__ spairs,
                                                          const int2 *__restrict__ pairs,
                                                          const int *__restrict__ start, const int *__restrict__ ov,
                                                          const double *__restrict__ s2, double *c, double *x4,
                                                          int *cn, double *msd) {
    const long long t = (long long)blockIdx.x * BLK_THREADS + threadIdx.x;
    if (t >= (long long)noi * ncell) return;
    const int slot = (int)(t / ncell), b = (int)(t - (long long)slot * ncell);
    double a = 0.0, a2 = 0.0, m = 0.0;
    int k = 0;
    for (int j = sptr[slot]; j < sptr[slot + 1]; j++) {
        const int p = spairs[j];
        const int *st = start + (size_t)pairs[p].x * (ncell + 1);
        const double nb = (double)(st[b + 1] - st[b]);
        const size_t q = (size_t)p * ncell + b;
        const double v = (double)ov[q] / nb;
        a += v;
        a2 += v * v;
        k++;
        m += s2[q] / nb;
    }
    const size_t o = (size_t)b * noi + slot;
    c[o] = a;
    x4[o] = a2;
    cn[o] = k;
    msd[o] = m;
}

// Wave (slot, q): out[q][slot] = sum over the slot's pairs (walk order) and
// blocks (ascending) of add[q][pair][block], one addition at a time.
__global__ void __launch_bounds__(64) k_blk_tot(int noi, int ncell, long long total, const int *__restrict__ sptr,
                                                const int *__restrict__ spairs, const double *__restrict__ add,
                                                double *out) {
    const int slot = blockIdx.x / 3, q = blockIdx.x - 3 * (blockIdx.x / 3), lane = threadIdx.x;
    const double *src = add + (size_t)q * total;
    double acc = 0.0;
    for (int j = sptr[slot]; j < sptr[slot + 1]; j++) {
        const double *row = src + (size_t)spairs[j] * ncell;
        for (int b0 = 0; b0 < ncell; b0 += 64) {
            const double v = b0 + lane < ncell ? row[b0 + lane] : 0.0;
            const int lo = __double2loint(v), hi = __double2hiint(v);
            const int m = min(64, ncell - b0);
            for (int k = 0; k < m; k++)
                acc += __hiloint2double(__builtin_amdgcn_readlane(hi, k), __builtin_amdgcn_readlane(lo, k));
        }
    }
    if (lane == 0) out[q * noi + slot] = acc;
}

namespace {

// The NaN an x86-64 SSE division 0.0/0.0 returns (sign bit set; glibc prints
// it as "-nan"): the only NaN the reference's arithmetic produces.
double x86_nan(double v) {
    if (v == v) return v;
    const uint64_t bits = 0xFFF8000000000000ull;
    memcpy(&v, &bits, sizeof v);
    return v;
}

// mkdir -p (Analysis_Block/main.c:50-55 runs "mkdir -p <sub_dir>/BlockAnalysis/...").
void mkdir_p(const char *path) {
    char buf[1024];
    snprintf(buf, sizeof buf, "%s", path);
    for (char *q = buf + 1; *q; q++)
        if (*q == '/') {
            *q = 0;
            mkdir(buf, 0755);
            *q = '/';
        }
    mkdir(buf, 0755);
}

}  // namespace

#define BTRY(e) do { int rc_ = (e); if (rc_ != AC4_OK) return rc_; } while (0)

extern "C" int ac4_frames_block(ac4_frames *f, int gap, double factor, double a, double cell_cut, double vh_time,
                                const char *sub_dir) {
    ac4_frames_view v;
    BTRY(ac4_frames_view_get(f, &v));
    AC4_REQUIRE(sub_dir && gap > 0 && factor > 1.0 && a > 0.0 && cell_cut > 0.0 && v.count >= 2, AC4_EARG,
                "ac4_frames_block: bad arguments (gap %d, %d frames)", gap, v.count);
    const int nf = v.count, n = v.n;
    const double L = v.L;
    std::vector<int> fidx;
    BTRY(ac4_frames_steps(v, fidx));
    const int ncx = (int)(L / cell_cut);  // msd_pos.c:33
    AC4_REQUIRE(ncx >= 1 && (long long)ncx * ncx * ncx <= BLK_SORT_MAX, AC4_EARG,
                "ac4_frames_block: %d blocks per edge (1..%d supported)", ncx, 25);
    const int ncell = ncx * ncx * ncx;

    ac4_walk_t w;
    memset(&w, 0, sizeof w);
    BTRY(ac4_walk(fidx.data(), nf, gap, factor, 1, &w));  // overlap.c:16-29, 74-167
    const int maxint = w.maximal_interval;
    if (maxint <= 0) {
        ac4_walk_free(&w);
        ac4_set_error("ac4_frames_block: no lag (maximal interval %d)", maxint);
        return AC4_EARG;
    }
    const int noi = 10 + (int)(log((double)maxint) / log(factor));
    std::vector<int> cur(w.cur, w.cur + w.np), oth(w.oth, w.oth + w.np), slot(w.slot, w.slot + w.np);
    ac4_walk_free(&w);
    const int np = (int)cur.size();
    for (int p = 0; p < np; p++)
        AC4_REQUIRE(slot[p] < noi, AC4_EARG, "ac4_frames_block: slot %d beyond %d", slot[p], noi);

    // origins visited by the do-while of msd_pos.c:70-195 (frame positions)
    std::vector<int> orig, seq_of(nf, -1);
    {
        int current = 0, k = 0;
        do {
            seq_of[current] = (int)orig.size();
            orig.push_back(current);
            k++;
            while (current < nf && fidx[current] != k * gap) current++;
        } while (current < nf);
    }
    const int nor = (int)orig.size();

    // vanHove.c:30-51: the lag nearest vh_time on the first frame's ladder
    int frame_num = 0;
    {
        std::vector<double> dl;
        int other = 0, offset = 1;
        const int cf = fidx[0];
        while (other < nf && offset < maxint) {
            const int of = cf + offset;
            while (other < nf && fidx[other] != of) other++;
            if (other < nf) {
                dl.push_back(BLK_TIME_STEP * (double)(of - cf));
                if ((int)(offset * factor) == offset)
                    offset++;
                else
                    offset = (int)(((double)offset) * factor);
            }
        }
        double inter = 1000.0;
        for (int i = 1; i < noi && i < (int)dl.size(); i++)
            if (fabs(dl[i] - vh_time) < inter) {
                frame_num = (int)(dl[i] * 200.0);
                inter = fabs(dl[i] - vh_time);
            }
    }

    // pairs: the walk's, then one per origin for the van Hove lag (vanHove.c:64-172)
    std::vector<int2> pairs;
    pairs.reserve(np + nor);
    for (int p = 0; p < np; p++) pairs.push_back(make_int2(seq_of[cur[p]], oth[p]));
    for (int o = 0; o < nor; o++) {
        const int want = fidx[orig[o]] + frame_num;
        int other = orig[o];
        while (other < nf && fidx[other] != want) other++;
        if (other < nf) pairs.push_back(make_int2(o, other));
    }
    const int nvh = (int)pairs.size() - np;
    const int npt = (int)pairs.size();
    std::vector<int> sptr(noi + 1, 0), spairs(np);
    std::vector<double> dt(noi, 0.0);
    for (int p = 0; p < np; p++) {
        sptr[slot[p] + 1]++;
        dt[slot[p]] = BLK_TIME_STEP * (double)(fidx[oth[p]] - fidx[cur[p]]);
    }
    for (int t = 0; t < noi; t++) sptr[t + 1] += sptr[t];
    {
        std::vector<int> fill(sptr.begin(), sptr.end() - 1);
        for (int p = 0; p < np; p++) spairs[fill[slot[p]]++] = p;  // walk order within a slot
    }

    const size_t res = (size_t)npt * ncell, bs = (size_t)ncell * noi;
    AcDevBuf B;
    int *d_ofr, *d_start, *d_mem, *d_ov, *d_sptr, *d_spairs, *d_cn, *d_wcnt;
    int2 *d_pairs;
    double *d_s2, *d_ex, *d_ey, *d_ez, *d_add, *d_c, *d_x4, *d_msd, *d_tot;
    BTRY(B.get(&d_ofr, nor));
    BTRY(B.get(&d_start, (size_t)nor * (ncell + 1)));
    BTRY(B.get(&d_mem, (size_t)nor * n));
    BTRY(B.get(&d_wcnt, (size_t)nor * BLK_SORT_WAVES * ncell));
    BTRY(B.get(&d_pairs, npt));
    BTRY(B.get(&d_ov, res));
    BTRY(B.get(&d_s2, res));
    BTRY(B.get(&d_ex, res));
    BTRY(B.get(&d_ey, res));
    BTRY(B.get(&d_ez, res));
    BTRY(B.get(&d_sptr, noi + 1));
    BTRY(B.get(&d_spairs, np));
    BTRY(B.get(&d_c, bs));
    BTRY(B.get(&d_x4, bs));
    BTRY(B.get(&d_cn, bs));
    BTRY(B.get(&d_msd, bs));
    BTRY(B.get(&d_add, 3 * res));
    BTRY(B.get(&d_tot, (size_t)3 * noi));
    hipStream_t s = v.s;
    AC4_HIP(hipMemcpyAsync(d_ofr, orig.data(), sizeof(int) * nor, hipMemcpyHostToDevice, s));
    AC4_HIP(hipMemcpyAsync(d_pairs, pairs.data(), sizeof(int2) * npt, hipMemcpyHostToDevice, s));
    AC4_HIP(hipMemcpyAsync(d_sptr, sptr.data(), sizeof(int) * (noi + 1), hipMemcpyHostToDevice, s));
    if (np) AC4_HIP(hipMemcpyAsync(d_spairs, spairs.data(), sizeof(int) * np, hipMemcpyHostToDevice, s));

    const int chunk = (n + BLK_SORT_WAVES - 1) / BLK_SORT_WAVES;
    hipLaunchKernelGGL(k_blk_count, dim3(nor, BLK_SORT_WAVES), dim3(64), sizeof(int) * ncell, s, v.x, v.y, v.z, n,
                       d_ofr, ncx, ncell, chunk, d_wcnt);
    hipLaunchKernelGGL(k_blk_offsets, dim3(nor), dim3(256), 0, s, ncell, BLK_SORT_WAVES, n, d_wcnt, d_start);
    hipLaunchKernelGGL(k_blk_place, dim3(nor, BLK_SORT_WAVES), dim3(64), sizeof(int) * ncell, s, v.x, v.y, v.z, n,
                       d_ofr, ncx, ncell, chunk, d_wcnt, d_mem);
    AC4_HIP(hipGetLastError());
    if (res) {
        hipLaunchKernelGGL(k_blk_pair, dim3((unsigned)((res + BLK_THREADS - 1) / BLK_THREADS)), dim3(BLK_THREADS), 0,
                           s, v.x, v.y, v.z, n, d_pairs, d_ofr, (long long)res, ncell, d_start, d_mem, L, a * a, d_ov,
                           d_s2, d_ex, d_ey, d_ez, d_add);
        AC4_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_blk_slot, dim3((unsigned)((bs + BLK_THREADS - 1) / BLK_THREADS)), dim3(BLK_THREADS), 0, s,
                       noi, ncell, d_sptr, d_spairs, d_pairs, d_start, d_ov, d_s2, d_c, d_x4, d_cn, d_msd);
    AC4_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_blk_tot, dim3(3 * noi), dim3(64), 0, s, noi, ncell, (long long)res, d_sptr, d_spairs, d_add,
                       d_tot);
    AC4_HIP(hipGetLastError());

    std::vector<double> c(bs), x4(bs), msd(bs), cT(noi), mc(noi), mc2(noi);
    std::vector<int> cn(bs), cTn(noi), binder(noi), last(ncell + 1);
    std::vector<double> vex((size_t)nvh * ncell), vey((size_t)nvh * ncell), vez((size_t)nvh * ncell);
    std::vector<int> vst((size_t)nvh * (ncell + 1));
    AC4_HIP(hipMemcpyAsync(c.data(), d_c, sizeof(double) * bs, hipMemcpyDeviceToHost, s));
    AC4_HIP(hipMemcpyAsync(x4.data(), d_x4, sizeof(double) * bs, hipMemcpyDeviceToHost, s));
    AC4_HIP(hipMemcpyAsync(msd.data(), d_msd, sizeof(double) * bs, hipMemcpyDeviceToHost, s));
    AC4_HIP(hipMemcpyAsync(cn.data(), d_cn, sizeof(int) * bs, hipMemcpyDeviceToHost, s));
    AC4_HIP(hipMemcpyAsync(cT.data(), d_tot, sizeof(double) * noi, hipMemcpyDeviceToHost, s));
    AC4_HIP(hipMemcpyAsync(mc.data(), d_tot + noi, sizeof(double) * noi, hipMemcpyDeviceToHost, s));
    AC4_HIP(hipMemcpyAsync(mc2.data(), d_tot + 2 * noi, sizeof(double) * noi, hipMemcpyDeviceToHost, s));
    for (int t = 0; t < noi; t++) cTn[t] = binder[t] = (sptr[t + 1] - sptr[t]) * ncell;  // countersT, CounterBinder
    AC4_HIP(hipMemcpyAsync(last.data(), d_start + (size_t)(nor - 1) * (ncell + 1), sizeof(int) * (ncell + 1),
                           hipMemcpyDeviceToHost, s));
    if (nvh) {
        const size_t o = (size_t)np * ncell, cnt = (size_t)nvh * ncell;
        AC4_HIP(hipMemcpyAsync(vex.data(), d_ex + o, sizeof(double) * cnt, hipMemcpyDeviceToHost, s));
        AC4_HIP(hipMemcpyAsync(vey.data(), d_ey + o, sizeof(double) * cnt, hipMemcpyDeviceToHost, s));
        AC4_HIP(hipMemcpyAsync(vez.data(), d_ez + o, sizeof(double) * cnt, hipMemcpyDeviceToHost, s));
        for (int k = 0; k < nvh; k++)
            AC4_HIP(hipMemcpyAsync(vst.data() + (size_t)k * (ncell + 1),
                                   d_start + (size_t)pairs[np + k].x * (ncell + 1), sizeof(int) * (ncell + 1),
                                   hipMemcpyDeviceToHost, s));
    }
    AC4_HIP(hipStreamSynchronize(s));
    for (auto *vec : {&c, &x4, &msd, &cT, &mc, &mc2})
        for (double &e : *vec) e = x86_nan(e);

    // host: averages over blocks and the output files (msd_pos.c:197-223,
    // overlap.c:171-219, vanHove.c:174-179); NumInCell is the last origin's
    char path[1024];
    snprintf(path, sizeof path, "%s/BlockAnalysis/MSD", sub_dir);
    mkdir_p(path);
    snprintf(path, sizeof path, "%s/BlockAnalysis/Relax", sub_dir);
    mkdir_p(path);
    const double lb = L / (double)ncx;
    FILE *fm, *fr, *ft, *fv;
    snprintf(path, sizeof path, "%s/BlockAnalysis/MSD/MSDLb%.3lf.dat", sub_dir, lb);
    fm = fopen(path, "wb");
    snprintf(path, sizeof path, "%s/BlockAnalysis/Relax/RelaxLb%.3lf.dat", sub_dir, lb);
    fr = fopen(path, "wb");
    snprintf(path, sizeof path, "%s/BlockAnalysis/Relax/RelaxTotLb%.3lf.dat", sub_dir, lb);
    ft = fopen(path, "wb");
    snprintf(path, sizeof path, "%s/BlockAnalysis/vanHove%.3lf.dat", sub_dir, lb);
    fv = fopen(path, "wb");
    if (!fm || !fr || !ft || !fv) {
        for (FILE *q : {fm, fr, ft, fv})
            if (q) fclose(q);
        ac4_set_error("ac4_frames_block: cannot write under %s/BlockAnalysis", sub_dir);
        return AC4_EIO;
    }
    fprintf(fm, "0.0\t0.0\t0.0\t0.0\n");
    double avg_nin = 0.0;
    for (int i = 0; i < ncell; i++) avg_nin += (double)(last[i + 1] - last[i]);
    avg_nin /= (double)ncell;
    for (int j = 1; j < noi; j++) {
        double bm = 0.0, bc = 0.0, bx = 0.0, x4t = 0.0;
        int bn = 0;
        for (int i = 0; i < ncell; i++) {
            const size_t k = (size_t)i * noi + j;
            if (cn[k]) {
                const int nin = last[i + 1] - last[i];
                bm += msd[k] / (double)cn[k];
                bc += c[k] / (double)cn[k];
                bx += (x4[k] / (double)cn[k] - c[k] * c[k] / ((double)(cn[k] * cn[k]))) * nin;
                bn++;
                x4t += x4[k];
            }
        }
        if (bn) {
            fprintf(fm, "%.14lf\t%.14lf\t%.14lf\t%.14lf\n", dt[j], bm / (double)bn, mc[j] / (double)binder[j],
                    mc2[j] / (double)binder[j]);
            fprintf(fr, "%.14lf\t%.14lf\t%.14lf\n", dt[j], bc / (double)bn, bx / (double)bn);
        }
        if (cTn[j])
            fprintf(ft, "%.14lf\t%.14lf\t%.14lf\n", dt[j], cT[j] / (double)cTn[j],
                    avg_nin * (x4t / (double)cTn[j] - cT[j] * cT[j] / ((double)cTn[j] * cTn[j])));
    }
    for (int k = 0; k < nvh; k++) {
        const int *st = vst.data() + (size_t)k * (ncell + 1);
        for (int i = 0; i < ncell; i++) {
            const size_t q = (size_t)k * ncell + i;
            const int nin = st[i + 1] - st[i];
            fprintf(fv, "%.14lf\n", (vex[q] / nin) * L);
            fprintf(fv, "%.14lf\n", (vey[q] / nin) * L);
            fprintf(fv, "%.14lf\n", (vez[q] / nin) * L);
        }
    }
    int bad = 0;
    for (FILE *q : {fm, fr, ft, fv}) bad |= fclose(q) != 0;
    AC4_REQUIRE(!bad, AC4_EIO, "ac4_frames_block: write error under %s/BlockAnalysis", sub_dir);
    return AC4_OK;
}
