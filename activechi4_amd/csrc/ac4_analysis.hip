// ac4_analysis.hip -- streaming dynamic-heterogeneity analysis on MI355X.
//
// Replaces Analysis/msd_pos.c, overlap.c, S4Qt.c and Fsqt.c (paths relative
// to the reference checkout).  The reference loads every frame into host RAM
// and walks (origin, lag) pairs with one MPI_Reduce per pair; here frames are
// pushed once in file order, origin frames stay resident in HBM, and each
// pushed frame is compared with all origins whose walk it completes in one
// launch (grid = particle blocks x pairs, x q groups for Fs); S4's phase sums
// are formed once per origin over all of its lags on the FP64 matrix cores.
// Per-pair results are reduced in a fixed order and accumulated into lag
// slots in the reference's origin order on the device, so integer
// observables (overlap counts) reproduce the reference's output bit for
// bit.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>
#include <vector>
#include <map>
#include "ac4_internal.h"

#define ABLK 256
#define QG 18                 // q vectors per thread group
#define NQG (AC4_NQ / QG)     // 9 groups
#define A_PI 3.14159265358979 // Analysis/global.h:14

__constant__ int c_vx[AC4_NQ], c_vy[AC4_NQ], c_vz[AC4_NQ], c_dg[AC4_NQ];

struct Job {  // one (origin, lag) pair evaluated against the pushed frame
    const double *x0, *y0, *z0;
    int pair, slot;
};

__device__ __forceinline__ double minimg_ana(double d) {  // overlap.c:67-78
    if (d >= 0.5)
        d -= 1.0;
    else if (d < -0.5)
        d += 1.0;
    return d;
}
__device__ __forceinline__ double wrap1(double v) {  // S4Qt.c:102-113
    if (v >= 1.0) v -= 1.0;
    if (v < 0.0) v += 1.0;
    return v;
}
// x0 after the per-lag single-if wrap has run `times` times (S4Qt.c:101-114
// wraps the resident origin coordinates again at every lag).
__device__ __forceinline__ double wrapn(double v, int times) {
    for (int k = 0; k < times; k++) {
        double w = wrap1(v);
        if (w == v) break;
        v = w;
    }
    return v;
}

template <int NV>
__device__ __forceinline__ void block_sum_store(double (&v)[NV], double *out) {
    __shared__ double s_red[NV][ABLK / 64];
    for (int k = 0; k < NV; k++) {
        double w = ac4_wave_sum(v[k]);
        if ((threadIdx.x & 63) == 0) s_red[k][threadIdx.x >> 6] = w;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.0;
        for (int w = 0; w < ABLK / 64; w++) s += s_red[threadIdx.x][w];
        out[threadIdx.x] = s;
    }
}

// MSD sums (msd_pos.c:64-76): grid (blocks, jobs); partial[(job*nb+blk)*3+v]
__global__ void __launch_bounds__(ABLK) k_msd(const Job *__restrict__ jobs, int n, double L,
                                              const double *__restrict__ x, const double *__restrict__ y,
                                              const double *__restrict__ z, const int *__restrict__ ty, int flag,
                                              double *partial) {
    const Job J = jobs[blockIdx.y];
    const int i = blockIdx.x * ABLK + threadIdx.x;
    double v[3] = {0.0, 0.0, 0.0};
    if (i < n && (ty[i] != 1 || flag == 0)) {
        double dx = x[i] - J.x0[i];
        double dy = y[i] - J.y0[i];
        double dz = z[i] - J.z0[i];
        double r2 = L * L * (dx * dx + dy * dy + dz * dz);
        double r3 = r2 * r2;
        double r4 = L * L * L * L * (dx * dx * dx * dx + dy * dy * dy * dy + dz * dz * dz * dz);
        v[0] = r2;  // sum
        v[1] = r4;  // sumSq
        v[2] = r3;  // sumCube
    }
    block_sum_store<3>(v, partial + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 3);
}

// overlap count (overlap.c:61-84): partial[(job*nb+blk)]
__global__ void __launch_bounds__(ABLK) k_overlap(const Job *__restrict__ jobs, int n, double L, double param,
                                                  const double *__restrict__ x, const double *__restrict__ y,
                                                  const double *__restrict__ z, const int *__restrict__ ty, int flag,
                                                  double *partial) {
    const Job J = jobs[blockIdx.y];
    const int i = blockIdx.x * ABLK + threadIdx.x;
    double v[1] = {0.0};
    if (i < n && (ty[i] != 1 || flag == 0)) {
        double dx = x[i] - J.x0[i];
        double dy = y[i] - J.y0[i];
        double dz = z[i] - J.z0[i];
        dz = minimg_ana(dz);
        dy = minimg_ana(dy);
        dx = minimg_ana(dx);
        double r2 = L * L * (dx * dx + dy * dy + dz * dz);
        if (r2 < param) v[0] = 1.0;
    }
    block_sum_store<1>(v, partial + ((size_t)blockIdx.y * gridDim.x + blockIdx.x));
}

// Permutation s of a degeneracy-d vector in the order of S4Qt.c:133-214.
__host__ __device__ __forceinline__ void perm_q(int d, int s, int a, int b, int c, int &o0, int &o1, int &o2) {
    if (d == 1 || s == 0) { o0 = a; o1 = b; o2 = c; return; }
    if (d == 3) {
        if (s == 1) { o0 = b; o1 = c; o2 = a; } else { o0 = c; o1 = a; o2 = b; }
        return;
    }
    switch (s) {
        case 1: o0 = a; o1 = c; o2 = b; break;
        case 2: o0 = b; o1 = a; o2 = c; break;
        case 3: o0 = b; o1 = c; o2 = a; break;
        case 4: o0 = c; o1 = a; o2 = b; break;
        default: o0 = c; o1 = b; o2 = a; break;
    }
}

// Fs (Fsqt.c:103-222): grid (blocks, NQG, jobs).  The phase is that of each
// particle's displacement, so every (origin, lag) pair takes its own 573
// sincos per particle.
// partial[((job*nb+blk)*AC4_NQ*2) + n] = cos sum, + AC4_NQ + n = sin sum.
__global__ void __launch_bounds__(ABLK) k_phase(const Job *__restrict__ jobs, int n, const double *__restrict__ x,
                                                const double *__restrict__ y, const double *__restrict__ z,
                                                const int *__restrict__ ty, int flag, double *partial) {
    const Job J = jobs[blockIdx.z];
    const int i = blockIdx.x * ABLK + threadIdx.x;
    const int qg = blockIdx.y;
    double v[2 * QG];
    for (int k = 0; k < 2 * QG; k++) v[k] = 0.0;
    if (i < n && (ty[i] != 1 || flag == 0)) {
        const double x0 = wrapn(J.x0[i], J.slot), y0 = wrapn(J.y0[i], J.slot), z0 = wrapn(J.z0[i], J.slot);
        const double xs = wrap1(x[i]), ys = wrap1(y[i]), zs = wrap1(z[i]);
        // phases of the displacement (Fsqt.c:145-146)
        const double p0 = minimg_ana(xs - x0), p1 = minimg_ana(ys - y0), p2 = minimg_ana(zs - z0);
        for (int k = 0; k < QG; k++) {
            const int nq = qg * QG + k;
            const int d = c_dg[nq];
            double cs = 0.0, sn = 0.0;
            for (int s = 0; s < d; s++) {
                int o0, o1, o2;
                perm_q(d, s, c_vx[nq], c_vy[nq], c_vz[nq], o0, o1, o2);
                // Fsqt.c:147-213 forms arg = q.L.p with q = 2*A_PI/L*(o0,o1,o2):
                // arg = 2 pi f with f = o.p turns, evaluated in turns.  The
                // reference's A_PI and its invL*L differ from pi and 1 in the
                // 15th digit: below 1e-12 in the phase for |f| <= 63.
                double sa, ca;
                ac4_sincos_turns((double)o0 * p0 + (double)o1 * p1 + (double)o2 * p2, &sa, &ca);
                cs += ca;
                sn += sa;
            }
            if (d == 3) { cs = cs / 3.0; sn = sn / 3.0; }
            else if (d == 6) { cs = cs / 6.0; sn = sn / 6.0; }
            v[k] = cs;
            v[QG + k] = sn;
        }
    }
    double *out = partial + ((size_t)blockIdx.z * gridDim.x + blockIdx.x) * (AC4_NQ * 2);
    __shared__ double s_red[2 * QG][ABLK / 64];
    for (int k = 0; k < 2 * QG; k++) {
        double w = ac4_wave_sum(v[k]);
        if ((threadIdx.x & 63) == 0) s_red[k][threadIdx.x >> 6] = w;
    }
    __syncthreads();
    if (threadIdx.x < 2 * QG) {
        double s = 0.0;
        for (int w = 0; w < ABLK / 64; w++) s += s_red[threadIdx.x][w];
        const int k = threadIdx.x;
        if (k < QG) out[qg * QG + k] = s;
        else out[AC4_NQ + qg * QG + (k - QG)] = s;
    }
}

// ---------------------------------------------------------------- S4
// S4Qt.c:101-216 sums, for each (origin, lag) pair and GetQ vector q, the
// phase exp(i q.r0) of every particle that still overlaps its origin
// position at the lag (|r(t) - r0| < a), averaged over q's permutations.
// The phase depends on the origin frame alone, so it is evaluated once per
// origin and shared by all of that origin's lags:
//   k_s4_mask   at each lag frame: the overlap mask of the pair, one bit per
//               particle (and the overlap count of the pair);
//   k_s4_origin when the origin's last lag has been masked: the phase matrix
//               P[particle][573 permuted vectors] times the mask matrix
//               M[particle][lags] on FP64 matrix cores, P formed in registers
//               and never stored;
//   k_s4_reduce sums the chunk partials in a fixed order and averages each
//               vector's permutations.
// This evaluates 573 sincos per particle per origin instead of per pair.
#define S4_NCOL 576           // the 573 permuted vectors, padded to whole 16-column tiles
#define S4_NCT (S4_NCOL / 16) // column tiles
#define S4_RT_MAX 4           // lag rows of one pass: up to 64 lags
#define S4_CHUNKS 64          // particle chunks a pass is split into (at most)

__constant__ int c_pv[3 * S4_NCOL];  // permuted vector (o0, o1, o2) per column; zero in the padding
__constant__ int c_pcol[AC4_NQ];     // first column of each GetQ vector

typedef double ac4_d4 __attribute__((ext_vector_type(4)));

// Overlap mask of S4Qt.c:125-133 for each pair: grid (blocks, jobs).  Bit i
// of mask[pair*nwords + i/32] is set when particle i overlaps; the count per
// block goes to mpartial[job*nb+blk].  The origin coordinates take the
// reference's per-lag single-if wraps (wrapn) exactly as it applies them.
__global__ void __launch_bounds__(ABLK) k_s4_mask(const Job *__restrict__ jobs, int n, double L, double param,
                                                  const double *__restrict__ x, const double *__restrict__ y,
                                                  const double *__restrict__ z, int nwords, unsigned *mask,
                                                  double *mpartial) {
    const Job J = jobs[blockIdx.y];
    const int i = blockIdx.x * ABLK + threadIdx.x;
    bool use = false;
    if (i < n) {
        const double x0 = wrapn(J.x0[i], J.slot), y0 = wrapn(J.y0[i], J.slot), z0 = wrapn(J.z0[i], J.slot);
        const double xs = wrap1(x[i]), ys = wrap1(y[i]), zs = wrap1(z[i]);
        double dx = (xs - x0), dy = (ys - y0), dz = (zs - z0);
        dz = minimg_ana(dz);
        dy = minimg_ana(dy);
        dx = minimg_ana(dx);
        const double r2 = L * L * (dx * dx + dy * dy + dz * dz);  // S4Qt.c:131-132
        use = r2 < param;
    }
    const unsigned long long b = __ballot(use);
    const int lane = threadIdx.x & 63;
    const int w = (blockIdx.x * ABLK + (threadIdx.x & ~63)) >> 5;  // first of the wave's two words
    unsigned *row = mask + (size_t)J.pair * nwords;
    if (lane == 0 && w < nwords) row[w] = (unsigned)b;
    if (lane == 1 && w + 1 < nwords) row[w + 1] = (unsigned)(b >> 32);
    double m[1] = {use ? 1.0 : 0.0};
    block_sum_store<1>(m, mpartial + ((size_t)blockIdx.y * gridDim.x + blockIdx.x));
}

// One origin, one pass of up to RT*16 of its lags: grid (chunks, S4_NCT),
// 256 threads.  Wave w of block (c, ct) takes the mask words
// [(4c+w)*wpw, (4c+w+1)*wpw) -- 32 particles each, 4 per MFMA step -- and
// the 16 permuted vectors of column tile ct.  v_mfma_f64_16x16x4f64 operand
// layout: lane l = i + 16k holds A[i][k] (mask bit of lag i, particle k) and
// B[k][j] with j = i (phase of vector j for particle k); the accumulator's
// element e of lane l is row l/16 + 4*e, column l%16.  The four waves'
// sums are added in wave order through LDS, and the block writes
// part[((c*rows + row)*2 + {cos,sin})*S4_NCOL + column].
template <int RT>
__global__ void __launch_bounds__(256) k_s4_origin(const double *__restrict__ x0, const double *__restrict__ y0,
                                                   const double *__restrict__ z0, int n, int nwords,
                                                   const unsigned *__restrict__ mask, const int *__restrict__ lags,
                                                   int nl, int wpw, double *__restrict__ part) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int ct = blockIdx.y;
    const int i16 = lane & 15, k = lane >> 4;
    const int col = ct * 16 + i16;
    const double o0 = (double)c_pv[3 * col], o1 = (double)c_pv[3 * col + 1], o2 = (double)c_pv[3 * col + 2];
    const unsigned *mrow[RT];
    bool live[RT];
    for (int r = 0; r < RT; r++) {
        const int l = 16 * r + i16;
        live[r] = l < nl;
        mrow[r] = mask + (size_t)lags[live[r] ? l : 0] * nwords;
    }
    ac4_d4 ac[RT], as[RT];
    for (int r = 0; r < RT; r++) {
        ac[r] = ac4_d4{0.0, 0.0, 0.0, 0.0};
        as[r] = ac4_d4{0.0, 0.0, 0.0, 0.0};
    }
    const int w0 = (blockIdx.x * 4 + wv) * wpw;
    const int w1 = min(w0 + wpw, nwords);
    for (int w = w0; w < w1; w++) {
        unsigned mw[RT];
        for (int r = 0; r < RT; r++) mw[r] = live[r] ? mrow[r][w] : 0u;
#pragma unroll
        for (int kk = 0; kk < 8; kk++) {
            // A particle past n has no mask bit: its (clamped) phase is
            // multiplied by zero.  exp(i q.r0) is periodic in r0 with the
            // box, so the stored origin coordinate is used as it is, whatever
            // single-if wraps the reference applies to it at a given lag.
            const int i = min(32 * w + 4 * kk + k, n - 1);
            double s, c;
            ac4_sincos_turns(fma(o2, z0[i], fma(o1, y0[i], o0 * x0[i])), &s, &c);
            const int bit = 4 * kk + k;
            for (int r = 0; r < RT; r++) {
                const double a = ((mw[r] >> bit) & 1u) ? 1.0 : 0.0;
                ac[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, c, ac[r], 0, 0, 0);
                as[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, s, as[r], 0, 0, 0);
            }
        }
    }
    __shared__ double s_acc[RT * 16 * 32];
    for (int step = 0; step < 4; step++) {
        if (wv == step)
            for (int r = 0; r < RT; r++)
                for (int e = 0; e < 4; e++) {
                    const int row = 16 * r + k + 4 * e;
                    double *pc = &s_acc[row * 32 + i16], *ps = &s_acc[row * 32 + 16 + i16];
                    if (step == 0) { *pc = ac[r][e]; *ps = as[r][e]; }
                    else { *pc += ac[r][e]; *ps += as[r][e]; }
                }
        __syncthreads();
    }
    const int rows = RT * 16;
    for (int t = threadIdx.x; t < rows * 32; t += 256) {
        const int row = t >> 5, cc = t & 31, sn = cc >> 4;
        part[((size_t)(blockIdx.x * rows + row) * 2 + sn) * S4_NCOL + ct * 16 + (cc & 15)] = s_acc[t];
    }
}

// Per lag l of the pass and value v (v < AC4_NQ: cos sum of vector v, else
// sin sum of vector v - AC4_NQ): grid (nl, ceil(2*AC4_NQ/64)), 256 threads.
// Wave w sums chunks w, w+4, ... (each over the vector's permutations in
// order); the four wave sums are added in wave order and divided by the
// number of permutations.  s4[pair*2*AC4_NQ + v].
__global__ void __launch_bounds__(256) k_s4_reduce(const double *__restrict__ part, int nchunk, int rows,
                                                   const int *__restrict__ lags, double *__restrict__ s4) {
    const int l = blockIdx.x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int v = blockIdx.y * 64 + lane;
    const bool ok = v < 2 * AC4_NQ;
    const int sn = ok && v >= AC4_NQ ? 1 : 0;
    const int q = ok ? v - sn * AC4_NQ : 0;
    const int d = c_dg[q], c0 = c_pcol[q];
    double acc = 0.0;
    if (ok)
        for (int c = wv; c < nchunk; c += 4) {
            const double *src = part + ((size_t)(c * rows + l) * 2 + sn) * S4_NCOL + c0;
            double t = 0.0;
            for (int s = 0; s < d; s++) t += src[s];
            acc += t;
        }
    __shared__ double s_w[4][64];
    s_w[wv][lane] = acc;
    __syncthreads();
    if (wv == 0 && ok) {
        const double tot = ((s_w[0][lane] + s_w[1][lane]) + s_w[2][lane]) + s_w[3][lane];
        s4[(size_t)lags[l] * 2 * AC4_NQ + v] = tot / (double)d;
    }
}

// Fixed-order sum over blocks: out[pair*nv + v] = sum_b partial[(job*nb+b)*nv + v].
// Grid (jobs, ceil(nv/64)), 256 threads: lane l of wave w takes value
// v = 64*blockIdx.y + l (consecutive values, coalesced loads) and the blocks
// b = w, w+4, w+8, ... with 8 loads in flight; the four wave sums are added
// in wave order.  The order depends on nb and nv only.
__global__ void __launch_bounds__(256) k_reduce_jobs(const Job *__restrict__ jobs, int nb, int nv,
                                                     const double *__restrict__ partial, double *out) {
    __shared__ double s_w[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, v = blockIdx.y * 64 + lane;
    double acc = 0.0;
    if (v < nv) {
        const double *p = partial + (size_t)blockIdx.x * nb * nv + v;
        int b = w;
        for (; b + 28 < nb; b += 32) {
            double t[8];
#pragma unroll
            for (int u = 0; u < 8; u++) t[u] = p[(size_t)(b + 4 * u) * nv];
#pragma unroll
            for (int u = 0; u < 8; u++) acc += t[u];
        }
        for (; b < nb; b += 4) acc += p[(size_t)b * nv];
    }
    s_w[w][lane] = acc;
    __syncthreads();
    if (w == 0 && v < nv) {
        const Job J = jobs[blockIdx.x];
        out[(size_t)J.pair * nv + v] = ((s_w[0][lane] + s_w[1][lane]) + s_w[2][lane]) + s_w[3][lane];
    }
}

// Slot accumulation in walk (origin) order, msd_pos.c:87-93 / overlap.c:89-94 /
// S4Qt.c:231-236.  mode 0: sum += val/div; mode 1: s = val/div (div>0) or val;
// sum += s; sq += s*s.
__global__ void k_slots(const double *__restrict__ vals, int nv, const int *__restrict__ ptr,
                        const int *__restrict__ pairs, int nslots, int mode, double div, double *sum, double *sq) {
    const int t = blockIdx.x;
    for (int v = threadIdx.x; v < nv; v += blockDim.x) {
        double a = 0.0, b = 0.0;
        for (int k = ptr[t]; k < ptr[t + 1]; k++) {
            double s = vals[(size_t)pairs[k] * nv + v];
            if (div > 0.0) s = s / div;
            a += s;
            if (mode == 1) b += s * s;
        }
        sum[(size_t)t * nv + v] = a;
        if (mode == 1) sq[(size_t)t * nv + v] = b;
    }
}

// ------------------------------------------------------------------ host
struct Frame {
    double *x = nullptr, *y = nullptr, *z = nullptr;
    int *t = nullptr;
};

struct WalkInfo {
    ac4_walk_t w;
    std::vector<std::vector<int>> by_oth;  // pair ids completed by frame pos
    std::vector<int> last_use;             // per origin pos: last frame pos needing it
    std::vector<int> slot_ptr, slot_pairs; // CSR of pairs per slot (walk order)
    int nslots = 0;
};

struct ac4_ana {
    int n = 0, nframes = 0, gap = 0, obs = 0, flag = 0;
    double L = 0, invL = 0, factor = 1.4, a = 0.3, param = 0.09;
    std::vector<int> fidx;
    WalkInfo W[2];
    std::vector<Frame> store;   // resident origin frames by pos
    Frame cur;                  // staging for the pushed frame
    int last_pos = -1;
    double *d_msd = nullptr, *d_ov = nullptr, *d_s4 = nullptr, *d_s4m = nullptr, *d_fs = nullptr;
    double *d_part = nullptr, *d_mpart = nullptr;
    Job *d_jobs = nullptr;
    // S4 by origin: pair ids per origin (CSR, walk order), origins whose
    // last lag is frame pos, the pairs' overlap masks and the chunk partials
    std::vector<int> s4_optr, s4_opairs;
    std::vector<std::vector<int>> s4_end;
    std::vector<char> s4_done;
    int *d_s4_opairs = nullptr;
    unsigned *d_s4mask = nullptr;
    double *d_s4part = nullptr;
    int nwords = 0, s4_wpw = 1, s4_nchunk = 1;
    int max_jobs = 0, nb = 0;
    std::vector<int> vx, vy, vz, qsq;
    std::vector<double> qval;
    hipStream_t stream = nullptr;
    bool finalized = false;
    // slot results
    std::vector<double> m0, m1, m2, c, x4, s4c, s4s, s4c2, s4s2, fsc, fss, fsc2, fss2;
    std::vector<int> cnt0, cnt1;
    std::vector<double> dt0, dt1;
    int noi_msd = 0, noi_ov = 0;
};

static int frame_alloc(Frame &f, int n) {
    AC4_HIP(hipMalloc(&f.x, sizeof(double) * n));
    AC4_HIP(hipMalloc(&f.y, sizeof(double) * n));
    AC4_HIP(hipMalloc(&f.z, sizeof(double) * n));
    AC4_HIP(hipMalloc(&f.t, sizeof(int) * n));
    return AC4_OK;
}
static void frame_free(Frame &f) {
    if (f.x) (void)hipFree(f.x);
    if (f.y) (void)hipFree(f.y);
    if (f.z) (void)hipFree(f.z);
    if (f.t) (void)hipFree(f.t);
    f = Frame();
}

#define TRY(e) do { int rc_ = (e); if (rc_ != AC4_OK) return rc_; } while (0)

static void build_walk_info(ac4_ana *an, WalkInfo &W, int rule) {
    ac4_walk(an->fidx.data(), an->nframes, an->gap, an->factor, rule, &W.w);
    W.by_oth.assign(an->nframes, {});
    W.last_use.assign(an->nframes, -1);
    for (int p = 0; p < W.w.np; p++) {
        W.by_oth[W.w.oth[p]].push_back(p);
        W.last_use[W.w.cur[p]] = std::max(W.last_use[W.w.cur[p]], W.w.oth[p]);
    }
}

static void build_slots(WalkInfo &W, int nslots) {
    W.nslots = nslots;
    W.slot_ptr.assign(nslots + 1, 0);
    for (int p = 0; p < W.w.np; p++) W.slot_ptr[W.w.slot[p] + 1]++;
    for (int t = 0; t < nslots; t++) W.slot_ptr[t + 1] += W.slot_ptr[t];
    W.slot_pairs.assign(W.w.np, 0);
    std::vector<int> fill(W.slot_ptr.begin(), W.slot_ptr.end() - 1);
    for (int p = 0; p < W.w.np; p++) W.slot_pairs[fill[W.w.slot[p]]++] = p;  // walk order kept
}

// S4 by origin: the permuted-vector columns, the per-origin pair lists and
// the device buffers.  Masks and sums start at zero, so a pair whose lag
// frame never arrives contributes nothing.
static int s4_setup(ac4_ana *an, const std::vector<int> &dg) {
    std::vector<int> pv(3 * S4_NCOL, 0), pcol(AC4_NQ, 0);
    int col = 0;
    for (int q = 0; q < AC4_NQ; q++) {
        pcol[q] = col;
        for (int s = 0; s < dg[q]; s++, col++)
            perm_q(dg[q], s, an->vx[q], an->vy[q], an->vz[q], pv[3 * col], pv[3 * col + 1], pv[3 * col + 2]);
    }
    AC4_REQUIRE(col <= S4_NCOL, AC4_EARG, "S4: %d permuted vectors exceed %d columns", col, S4_NCOL);
    AC4_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_pv), pv.data(), sizeof(int) * 3 * S4_NCOL));
    AC4_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_pcol), pcol.data(), sizeof(int) * AC4_NQ));
    const ac4_walk_t &w = an->W[1].w;
    an->s4_optr.assign(an->nframes + 1, 0);
    for (int p = 0; p < w.np; p++) an->s4_optr[w.cur[p] + 1]++;
    for (int o = 0; o < an->nframes; o++) an->s4_optr[o + 1] += an->s4_optr[o];
    an->s4_opairs.assign(std::max(w.np, 1), 0);
    std::vector<int> fill(an->s4_optr.begin(), an->s4_optr.end() - 1);
    for (int p = 0; p < w.np; p++) an->s4_opairs[fill[w.cur[p]]++] = p;
    an->s4_end.assign(an->nframes, {});
    for (int o = 0; o < an->nframes; o++)
        if (an->W[1].last_use[o] >= 0) an->s4_end[an->W[1].last_use[o]].push_back(o);
    an->s4_done.assign(an->nframes, 0);
    an->nwords = (an->n + 31) / 32;
    an->s4_wpw = std::max(1, (an->nwords + 4 * S4_CHUNKS - 1) / (4 * S4_CHUNKS));
    an->s4_nchunk = (an->nwords + 4 * an->s4_wpw - 1) / (4 * an->s4_wpw);
    const size_t np1 = std::max(w.np, 1);
    AC4_HIP(hipMalloc(&an->d_s4_opairs, sizeof(int) * np1));
    AC4_HIP(hipMemcpy(an->d_s4_opairs, an->s4_opairs.data(), sizeof(int) * np1, hipMemcpyHostToDevice));
    AC4_HIP(hipMalloc(&an->d_s4mask, sizeof(unsigned) * np1 * an->nwords));
    AC4_HIP(hipMalloc(&an->d_s4part, sizeof(double) * (size_t)an->s4_nchunk * 16 * S4_RT_MAX * 2 * S4_NCOL));
    AC4_HIP(hipMemset(an->d_s4mask, 0, sizeof(unsigned) * np1 * an->nwords));
    AC4_HIP(hipMemset(an->d_s4, 0, sizeof(double) * 2 * AC4_NQ * np1));
    AC4_HIP(hipMemset(an->d_s4m, 0, sizeof(double) * np1));
    return AC4_OK;
}

// The S4 sums of every pair of origin o, from its resident frame and the
// masks its lags left; passes of up to 16*S4_RT_MAX lags.
static int s4_origin(ac4_ana *an, int o) {
    const Frame &F = an->store[o];
    AC4_REQUIRE(F.x, AC4_ESTATE, "S4: origin frame %d is not resident", o);
    hipStream_t s = an->stream;
    const int b = an->s4_optr[o], e = an->s4_optr[o + 1];
    for (int p0 = b; p0 < e; p0 += 16 * S4_RT_MAX) {
        const int nl = std::min(16 * S4_RT_MAX, e - p0), rt = (nl + 15) / 16;
        const int *lags = an->d_s4_opairs + p0;
        const dim3 g(an->s4_nchunk, S4_NCT);
        switch (rt) {
            case 1: hipLaunchKernelGGL(k_s4_origin<1>, g, dim3(256), 0, s, F.x, F.y, F.z, an->n, an->nwords,
                                       an->d_s4mask, lags, nl, an->s4_wpw, an->d_s4part); break;
            case 2: hipLaunchKernelGGL(k_s4_origin<2>, g, dim3(256), 0, s, F.x, F.y, F.z, an->n, an->nwords,
                                       an->d_s4mask, lags, nl, an->s4_wpw, an->d_s4part); break;
            case 3: hipLaunchKernelGGL(k_s4_origin<3>, g, dim3(256), 0, s, F.x, F.y, F.z, an->n, an->nwords,
                                       an->d_s4mask, lags, nl, an->s4_wpw, an->d_s4part); break;
            default: hipLaunchKernelGGL(k_s4_origin<4>, g, dim3(256), 0, s, F.x, F.y, F.z, an->n, an->nwords,
                                        an->d_s4mask, lags, nl, an->s4_wpw, an->d_s4part); break;
        }
        hipLaunchKernelGGL(k_s4_reduce, dim3(nl, (2 * AC4_NQ + 63) / 64), dim3(256), 0, s, an->d_s4part,
                           an->s4_nchunk, rt * 16, lags, an->d_s4);
        AC4_HIP(hipGetLastError());
    }
    an->s4_done[o] = 1;
    return AC4_OK;
}

extern "C" int ac4_ana_create(int n, double L, int nframes, const int *fidx, int gap, double factor, double a,
                              int obs, int type_flag, ac4_ana **out) {
    AC4_REQUIRE(n > 0 && L > 0 && nframes >= 2 && fidx && out && factor > 1.0 && a > 0, AC4_EARG,
                "ac4_ana_create: bad arguments");
    int ndev = 0;
    AC4_REQUIRE(hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0, AC4_EHIP,
                "ac4_ana_create: no HIP device visible (libac4 has no CPU path)");
    for (int i = 1; i < nframes; i++)
        AC4_REQUIRE(fidx[i] > fidx[i - 1], AC4_EARG, "ac4_ana_create: frame index list not increasing at %d", i);
    ac4_ana *an = new ac4_ana();
    an->n = n; an->L = L; an->invL = 1.0 / L; an->nframes = nframes; an->factor = factor; an->a = a;
    an->param = a * a;  // overlap.c:5
    an->obs = obs; an->flag = type_flag;
    an->fidx.assign(fidx, fidx + nframes);
    an->gap = gap > 0 ? gap : ac4_origin_gap(fidx, nframes);
    if (an->gap <= 0) { delete an; return AC4_EARG; }
    build_walk_info(an, an->W[0], 0);
    build_walk_info(an, an->W[1], 1);
    // slot counts: msd_pos.c:29 / overlap.c:29 (+ bounds), S4/Fs use num_of_frames
    const double lf = log(an->factor);
    an->noi_msd = 10 + (int)(log((double)an->W[0].w.maximal_interval) / lf) + 10;
    an->noi_ov = 10 + (int)(log((double)an->W[1].w.maximal_interval) / lf);
    build_slots(an->W[0], std::max(an->noi_msd, an->W[0].w.max_slot + 1));
    build_slots(an->W[1], std::max(std::max(an->noi_ov, nframes), an->W[1].w.max_slot + 1));
    an->store.assign(nframes, Frame());
    an->vx.resize(AC4_NQ); an->vy.resize(AC4_NQ); an->vz.resize(AC4_NQ); an->qsq.resize(AC4_NQ + 8);
    ac4_getq(an->vx.data(), an->vy.data(), an->vz.data(), an->qsq.data());
    std::vector<int> dg(AC4_NQ);
    an->qval.assign(AC4_NQ, 0.0);
    for (int q = 0; q < AC4_NQ; q++) {
        dg[q] = ac4_degen(an->vx[q], an->vy[q], an->vz[q]);
        // Q[n] as left by S4Qt.c:135-208: first permutation for d<=3, last (z,y,x) for d=6
        int o0 = an->vx[q], o1 = an->vy[q], o2 = an->vz[q];
        if (dg[q] == 6) { o0 = an->vz[q]; o1 = an->vy[q]; o2 = an->vx[q]; }
        double qx = 2.0 * A_PI * an->invL * (double)o0, qy = 2.0 * A_PI * an->invL * (double)o1,
               qz = 2.0 * A_PI * an->invL * (double)o2;
        an->qval[q] = qx * qx + qy * qy + qz * qz;
    }
    int rc = AC4_OK;
    auto fail = [&](int e) { ac4_ana_destroy(an); return e; };
    if (hipStreamCreate(&an->stream) != hipSuccess) return fail(AC4_EHIP);
    AC4_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_vx), an->vx.data(), sizeof(int) * AC4_NQ));
    AC4_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_vy), an->vy.data(), sizeof(int) * AC4_NQ));
    AC4_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_vz), an->vz.data(), sizeof(int) * AC4_NQ));
    AC4_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_dg), dg.data(), sizeof(int) * AC4_NQ));
    an->nb = (n + ABLK - 1) / ABLK;
    for (int w = 0; w < 2; w++)
        for (int f = 0; f < nframes; f++) an->max_jobs = std::max(an->max_jobs, (int)an->W[w].by_oth[f].size());
    an->max_jobs = std::max(an->max_jobs, 1);
    const int np0 = std::max(an->W[0].w.np, 1), np1 = std::max(an->W[1].w.np, 1);
    if (hipMalloc(&an->d_jobs, sizeof(Job) * an->max_jobs) != hipSuccess) return fail(AC4_EHIP);
    if (hipMalloc(&an->d_part, sizeof(double) * (size_t)an->max_jobs * an->nb * AC4_NQ * 2) != hipSuccess) return fail(AC4_EHIP);
    if (hipMalloc(&an->d_mpart, sizeof(double) * (size_t)an->max_jobs * an->nb) != hipSuccess) return fail(AC4_EHIP);
    if (obs & AC4_OBS_MSD) if (hipMalloc(&an->d_msd, sizeof(double) * 3 * (size_t)np0) != hipSuccess) return fail(AC4_EHIP);
    if (obs & AC4_OBS_OVERLAP) if (hipMalloc(&an->d_ov, sizeof(double) * (size_t)np1) != hipSuccess) return fail(AC4_EHIP);
    if (obs & AC4_OBS_S4) {
        if (hipMalloc(&an->d_s4, sizeof(double) * 2 * AC4_NQ * (size_t)np1) != hipSuccess) return fail(AC4_EHIP);
        if (hipMalloc(&an->d_s4m, sizeof(double) * (size_t)np1) != hipSuccess) return fail(AC4_EHIP);
        if ((rc = s4_setup(an, dg)) != AC4_OK) return fail(rc);
    }
    if (obs & AC4_OBS_FS) if (hipMalloc(&an->d_fs, sizeof(double) * 2 * AC4_NQ * (size_t)np1) != hipSuccess) return fail(AC4_EHIP);
    if ((rc = frame_alloc(an->cur, n)) != AC4_OK) return fail(rc);
    *out = an;
    return AC4_OK;
}

extern "C" int ac4_ana_destroy(ac4_ana *an) {
    if (!an) return AC4_OK;
    if (an->stream) (void)hipStreamSynchronize(an->stream);
    for (auto &f : an->store) frame_free(f);
    frame_free(an->cur);
    void *ptrs[] = {an->d_msd, an->d_ov, an->d_s4, an->d_s4m, an->d_fs, an->d_part, an->d_mpart, an->d_jobs,
                    an->d_s4_opairs, an->d_s4mask, an->d_s4part};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (auto &W : an->W) ac4_walk_free(&W.w);
    if (an->stream) (void)hipStreamDestroy(an->stream);
    delete an;
    return AC4_OK;
}

extern "C" int ac4_ana_gap(const ac4_ana *an) { return an ? an->gap : AC4_EARG; }

static bool is_origin(const ac4_ana *an, int pos) {
    return an->W[0].last_use[pos] >= 0 || an->W[1].last_use[pos] >= 0;
}

// Runs every pair completed by the frame now in `f` (device arrays).
static int process_frame(ac4_ana *an, int pos, const double *x, const double *y, const double *z, const int *t) {
    hipStream_t s = an->stream;
    const int nb = an->nb;
    for (int w = 0; w < 2; w++) {
        const std::vector<int> &ps = an->W[w].by_oth[pos];
        if (ps.empty()) continue;
        const bool want = w == 0 ? (an->obs & AC4_OBS_MSD) != 0
                                 : (an->obs & (AC4_OBS_OVERLAP | AC4_OBS_S4 | AC4_OBS_FS)) != 0;
        if (!want) continue;
        std::vector<Job> jobs(ps.size());
        for (size_t k = 0; k < ps.size(); k++) {
            const int p = ps[k], o = an->W[w].w.cur[p];
            const Frame &F = an->store[o];
            AC4_REQUIRE(F.x, AC4_ESTATE, "origin frame %d of pair %d was not pushed", o, p);
            jobs[k] = Job{F.x, F.y, F.z, p, an->W[w].w.slot[p]};
        }
        const int nj = (int)jobs.size();
        AC4_HIP(hipMemcpyAsync(an->d_jobs, jobs.data(), sizeof(Job) * nj, hipMemcpyHostToDevice, s));
        if (w == 0) {
            hipLaunchKernelGGL(k_msd, dim3(nb, nj), dim3(ABLK), 0, s, an->d_jobs, an->n, an->L, x, y, z, t, an->flag,
                               an->d_part);
            hipLaunchKernelGGL(k_reduce_jobs, dim3(nj, 1), dim3(256), 0, s, an->d_jobs, nb, 3, an->d_part, an->d_msd);
        } else {
            if (an->obs & AC4_OBS_OVERLAP) {
                hipLaunchKernelGGL(k_overlap, dim3(nb, nj), dim3(ABLK), 0, s, an->d_jobs, an->n, an->L, an->param, x,
                                   y, z, t, an->flag, an->d_part);
                hipLaunchKernelGGL(k_reduce_jobs, dim3(nj, 1), dim3(256), 0, s, an->d_jobs, nb, 1, an->d_part, an->d_ov);
            }
            if (an->obs & AC4_OBS_S4) {
                hipLaunchKernelGGL(k_s4_mask, dim3(nb, nj), dim3(ABLK), 0, s, an->d_jobs, an->n, an->L, an->param, x,
                                   y, z, an->nwords, an->d_s4mask, an->d_mpart);
                hipLaunchKernelGGL(k_reduce_jobs, dim3(nj, 1), dim3(256), 0, s, an->d_jobs, nb, 1, an->d_mpart, an->d_s4m);
            }
            if (an->obs & AC4_OBS_FS) {
                hipLaunchKernelGGL(k_phase, dim3(nb, NQG, nj), dim3(ABLK), 0, s, an->d_jobs, an->n, x, y, z, t,
                                   an->flag, an->d_part);
                hipLaunchKernelGGL(k_reduce_jobs, dim3(nj, (2 * AC4_NQ + 63) / 64), dim3(256), 0, s, an->d_jobs, nb, 2 * AC4_NQ, an->d_part,
                                   an->d_fs);
            }
        }
        AC4_HIP(hipGetLastError());
        // jobs[] is a pageable host vector: the copy must finish before it dies
        AC4_HIP(hipStreamSynchronize(s));
    }
    // origins whose last lag this frame was: their S4 sums, before the
    // origin frames are released
    if ((an->obs & AC4_OBS_S4) && !an->s4_end[pos].empty()) {
        for (int o : an->s4_end[pos]) TRY(s4_origin(an, o));
        AC4_HIP(hipStreamSynchronize(s));
    }
    return AC4_OK;
}

static int push_common(ac4_ana *an, int pos, const double *x, const double *y, const double *z, const int *t,
                       hipMemcpyKind kind) {
    AC4_REQUIRE(an && pos >= 0 && pos < an->nframes && x && y && z && t, AC4_EARG, "push_frame: bad arguments");
    AC4_REQUIRE(pos > an->last_pos, AC4_ESTATE, "push_frame: frame %d after %d (frames must increase)", pos,
                an->last_pos);
    AC4_REQUIRE(!an->finalized, AC4_ESTATE, "push_frame after finalize");
    hipStream_t s = an->stream;
    const size_t n = an->n;
    Frame dst = an->cur;
    if (is_origin(an, pos)) {
        TRY(frame_alloc(an->store[pos], an->n));
        dst = an->store[pos];
    }
    AC4_HIP(hipMemcpyAsync(dst.x, x, sizeof(double) * n, kind, s));
    AC4_HIP(hipMemcpyAsync(dst.y, y, sizeof(double) * n, kind, s));
    AC4_HIP(hipMemcpyAsync(dst.z, z, sizeof(double) * n, kind, s));
    AC4_HIP(hipMemcpyAsync(dst.t, t, sizeof(int) * n, kind, s));
    AC4_HIP(hipStreamSynchronize(s));
    TRY(process_frame(an, pos, dst.x, dst.y, dst.z, dst.t));
    an->last_pos = pos;
    // release origins no pair needs any more
    for (int o = 0; o <= pos; o++)
        if (an->store[o].x && an->W[0].last_use[o] <= pos && an->W[1].last_use[o] <= pos) frame_free(an->store[o]);
    return AC4_OK;
}

extern "C" int ac4_ana_push_frame(ac4_ana *an, int pos, const double *x, const double *y, const double *z,
                                  const int *type) {
    return push_common(an, pos, x, y, z, type, hipMemcpyHostToDevice);
}

extern "C" int ac4_ana_push_frame_device(ac4_ana *an, int pos, const double *x, const double *y, const double *z,
                                         const int *type) {
    return push_common(an, pos, x, y, z, type, hipMemcpyDeviceToDevice);
}

extern "C" int ac4_md_device_xu(ac4_md *md, int r, const double **xu, const double **yu, const double **zu,
                                const int **type, int *n);
extern "C" int ac4_ana_push_md(ac4_ana *an, int pos, ac4_md *md, int r) {
    const double *x, *y, *z;
    const int *t;
    int n = 0;
    TRY(ac4_md_device_xu(md, r, &x, &y, &z, &t, &n));
    AC4_REQUIRE(an && n == an->n, AC4_EARG, "push_md: replica has %d particles, analysis expects %d", n,
                an ? an->n : -1);
    return push_common(an, pos, x, y, z, t, hipMemcpyDeviceToDevice);
}

static int slots_run(ac4_ana *an, int w, const double *dvals, int nv, int mode, double div, std::vector<double> &sum,
                     std::vector<double> *sq) {
    WalkInfo &W = an->W[w];
    const int ns = W.nslots;
    int *dptr = nullptr, *dpairs = nullptr;
    double *dsum = nullptr, *dsq = nullptr;
    AC4_HIP(hipMalloc(&dptr, sizeof(int) * (ns + 1)));
    AC4_HIP(hipMalloc(&dpairs, sizeof(int) * std::max(W.w.np, 1)));
    AC4_HIP(hipMalloc(&dsum, sizeof(double) * (size_t)ns * nv));
    AC4_HIP(hipMalloc(&dsq, sizeof(double) * (size_t)ns * nv));
    AC4_HIP(hipMemcpy(dptr, W.slot_ptr.data(), sizeof(int) * (ns + 1), hipMemcpyHostToDevice));
    if (W.w.np) AC4_HIP(hipMemcpy(dpairs, W.slot_pairs.data(), sizeof(int) * W.w.np, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_slots, dim3(ns), dim3(256), 0, an->stream, dvals, nv, dptr, dpairs, ns, mode, div, dsum, dsq);
    AC4_HIP(hipGetLastError());
    sum.assign((size_t)ns * nv, 0.0);
    AC4_HIP(hipMemcpyAsync(sum.data(), dsum, sizeof(double) * ns * nv, hipMemcpyDeviceToHost, an->stream));
    if (sq) {
        sq->assign((size_t)ns * nv, 0.0);
        AC4_HIP(hipMemcpyAsync(sq->data(), dsq, sizeof(double) * ns * nv, hipMemcpyDeviceToHost, an->stream));
    }
    AC4_HIP(hipStreamSynchronize(an->stream));
    (void)hipFree(dptr); (void)hipFree(dpairs); (void)hipFree(dsum); (void)hipFree(dsq);
    return AC4_OK;
}

static void split_phase(const std::vector<double> &in, int ns, std::vector<double> &c, std::vector<double> &s) {
    c.assign((size_t)ns * AC4_NQ, 0.0);
    s.assign((size_t)ns * AC4_NQ, 0.0);
    for (int t = 0; t < ns; t++)
        for (int q = 0; q < AC4_NQ; q++) {
            c[(size_t)t * AC4_NQ + q] = in[(size_t)t * 2 * AC4_NQ + q];
            s[(size_t)t * AC4_NQ + q] = in[(size_t)t * 2 * AC4_NQ + AC4_NQ + q];
        }
}

extern "C" int ac4_ana_finalize(ac4_ana *an) {
    AC4_REQUIRE(an, AC4_EARG, "NULL handle");
    if (an->finalized) return AC4_OK;
    if (an->obs & AC4_OBS_S4)  // origins whose last lag frame was never pushed
        for (int o = 0; o < an->nframes; o++)
            if (!an->s4_done[o] && an->store[o].x && an->s4_optr[o + 1] > an->s4_optr[o]) TRY(s4_origin(an, o));
    AC4_HIP(hipStreamSynchronize(an->stream));
    const double tot = an->flag == 1 ? 0.8 * an->n : (double)an->n;  // msd_pos.c:26-27
    for (int w = 0; w < 2; w++) {  // counters and delta_t (msd_pos.c:90-92)
        WalkInfo &W = an->W[w];
        std::vector<int> &cn = w == 0 ? an->cnt0 : an->cnt1;
        std::vector<double> &dt = w == 0 ? an->dt0 : an->dt1;
        cn.assign(W.nslots, 0);
        dt.assign(W.nslots, 0.0);
        for (int p = 0; p < W.w.np; p++) {
            cn[W.w.slot[p]]++;
            dt[W.w.slot[p]] = 0.005 * (double)(an->fidx[W.w.oth[p]] - an->fidx[W.w.cur[p]]);
        }
    }
    if (an->obs & AC4_OBS_MSD) {
        std::vector<double> s;
        TRY(slots_run(an, 0, an->d_msd, 3, 0, tot, s, nullptr));
        const int ns = an->W[0].nslots;
        an->m0.assign(ns, 0); an->m1.assign(ns, 0); an->m2.assign(ns, 0);
        for (int t = 0; t < ns; t++) { an->m0[t] = s[3 * t]; an->m1[t] = s[3 * t + 1]; an->m2[t] = s[3 * t + 2]; }
    }
    if (an->obs & AC4_OBS_OVERLAP) TRY(slots_run(an, 1, an->d_ov, 1, 1, tot, an->c, &an->x4));
    if (an->obs & AC4_OBS_S4) {
        std::vector<double> s, s2;
        TRY(slots_run(an, 1, an->d_s4, 2 * AC4_NQ, 1, 0.0, s, &s2));
        split_phase(s, an->W[1].nslots, an->s4c, an->s4s);
        split_phase(s2, an->W[1].nslots, an->s4c2, an->s4s2);
    }
    if (an->obs & AC4_OBS_FS) {
        std::vector<double> s, s2;
        TRY(slots_run(an, 1, an->d_fs, 2 * AC4_NQ, 1, 0.0, s, &s2));
        split_phase(s, an->W[1].nslots, an->fsc, an->fss);
        split_phase(s2, an->W[1].nslots, an->fsc2, an->fss2);
    }
    an->finalized = true;
    return AC4_OK;
}

extern "C" int ac4_ana_nslots(const ac4_ana *an, int obs) {
    if (!an) return AC4_EARG;
    return obs == AC4_OBS_MSD ? an->W[0].nslots : an->W[1].nslots;
}

extern "C" int ac4_ana_get_msd(ac4_ana *an, double *m0, double *m1, double *m2, int *counters, double *delta_t,
                               int *num_intervals) {
    AC4_REQUIRE(an && an->finalized && (an->obs & AC4_OBS_MSD), AC4_ESTATE, "get_msd: not finalized / not requested");
    const int ns = an->W[0].nslots;
    for (int t = 0; t < ns; t++) {
        if (m0) m0[t] = an->m0[t];
        if (m1) m1[t] = an->m1[t];
        if (m2) m2[t] = an->m2[t];
        if (counters) counters[t] = an->cnt0[t];
        if (delta_t) delta_t[t] = an->dt0[t];
    }
    if (num_intervals) *num_intervals = an->noi_msd;
    return AC4_OK;
}

extern "C" int ac4_ana_get_overlap(ac4_ana *an, double *c, double *x4, int *counters, double *delta_t,
                                   int *num_intervals) {
    AC4_REQUIRE(an && an->finalized && (an->obs & AC4_OBS_OVERLAP), AC4_ESTATE,
                "get_overlap: not finalized / not requested");
    const int ns = an->W[1].nslots;
    for (int t = 0; t < ns; t++) {
        if (c) c[t] = an->c[t];
        if (x4) x4[t] = an->x4[t];
        if (counters) counters[t] = an->cnt1[t];
        if (delta_t) delta_t[t] = an->dt1[t];
    }
    if (num_intervals) *num_intervals = an->noi_ov;
    return AC4_OK;
}

extern "C" int ac4_ana_get_phase(ac4_ana *an, int obs, double *cs, double *ss, double *cs2, double *ss2,
                                 int *counters, double *delta_t, double *qval) {
    AC4_REQUIRE(an && an->finalized && (obs == AC4_OBS_S4 || obs == AC4_OBS_FS) && (an->obs & obs), AC4_ESTATE,
                "get_phase: not finalized / not requested");
    const bool s4 = obs == AC4_OBS_S4;
    const size_t ns = an->W[1].nslots, m = ns * AC4_NQ;
    const std::vector<double> &C = s4 ? an->s4c : an->fsc, &S = s4 ? an->s4s : an->fss, &C2 = s4 ? an->s4c2 : an->fsc2,
                              &S2 = s4 ? an->s4s2 : an->fss2;
    if (cs) memcpy(cs, C.data(), sizeof(double) * m);
    if (ss) memcpy(ss, S.data(), sizeof(double) * m);
    if (cs2) memcpy(cs2, C2.data(), sizeof(double) * m);
    if (ss2) memcpy(ss2, S2.data(), sizeof(double) * m);
    if (counters) memcpy(counters, an->cnt1.data(), sizeof(int) * ns);
    if (delta_t) memcpy(delta_t, an->dt1.data(), sizeof(double) * ns);
    if (qval) memcpy(qval, an->qval.data(), sizeof(double) * AC4_NQ);
    return AC4_OK;
}

extern "C" int ac4_ana_npairs(const ac4_ana *an, int obs) {
    if (!an) return AC4_EARG;
    return obs == AC4_OBS_MSD ? an->W[0].w.np : an->W[1].w.np;
}

extern "C" int ac4_ana_get_pairs(ac4_ana *an, int obs, double *vals, int *ivals) {
    AC4_REQUIRE(an && (an->obs & obs), AC4_ESTATE, "get_pairs: observable not requested");
    AC4_HIP(hipStreamSynchronize(an->stream));
    const size_t np = obs == AC4_OBS_MSD ? an->W[0].w.np : an->W[1].w.np;
    if (np == 0) return AC4_OK;
    if (obs == AC4_OBS_MSD && vals) AC4_HIP(hipMemcpy(vals, an->d_msd, sizeof(double) * 3 * np, hipMemcpyDeviceToHost));
    if (obs == AC4_OBS_OVERLAP && ivals) {
        std::vector<double> v(np);
        AC4_HIP(hipMemcpy(v.data(), an->d_ov, sizeof(double) * np, hipMemcpyDeviceToHost));
        for (size_t p = 0; p < np; p++) ivals[p] = (int)v[p];
    }
    if (obs == AC4_OBS_S4) {
        if (vals) AC4_HIP(hipMemcpy(vals, an->d_s4, sizeof(double) * 2 * AC4_NQ * np, hipMemcpyDeviceToHost));
        if (ivals) {
            std::vector<double> v(np);
            AC4_HIP(hipMemcpy(v.data(), an->d_s4m, sizeof(double) * np, hipMemcpyDeviceToHost));
            for (size_t p = 0; p < np; p++) ivals[p] = (int)v[p];
        }
    }
    if (obs == AC4_OBS_FS && vals) AC4_HIP(hipMemcpy(vals, an->d_fs, sizeof(double) * 2 * AC4_NQ * np, hipMemcpyDeviceToHost));
    return AC4_OK;
}

extern "C" int ac4_ana_write(ac4_ana *an, const char *sub_dir) {
    AC4_REQUIRE(an && sub_dir, AC4_EARG, "write: bad arguments");
    if (!an->finalized) TRY(ac4_ana_finalize(an));
    char path[2048];
    const double tot = an->flag == 1 ? 0.8 * an->n : (double)an->n;
    if (an->obs & AC4_OBS_MSD) {
        snprintf(path, sizeof path, "%s/%s", sub_dir, an->flag == 1 ? "msdA.dat" : "msd.dat");
        TRY(ac4_write_msd(path, an->noi_msd, an->m0.data(), an->m1.data(), an->m2.data(), an->cnt0.data(),
                          an->dt0.data()));
    }
    if (an->obs & AC4_OBS_OVERLAP) {
        snprintf(path, sizeof path, an->flag == 1 ? "%s/relaxAA_a_%.2lf.dat" : "%s/relax_a_%.2lf.dat", sub_dir, an->a);
        TRY(ac4_write_overlap(path, an->noi_ov, an->c.data(), an->x4.data(), an->cnt1.data(), an->dt1.data()));
    }
    // S4Qt.c/Fsqt.c loop over num_of_frames slots
    const int ns = std::min(an->W[1].nslots, an->nframes);
    if (an->obs & AC4_OBS_S4) {
        snprintf(path, sizeof path, "%s/S4qt_a_%.2lf.dat", sub_dir, an->a);
        TRY(ac4_write_phase(path, 1, an->n, tot, ns, an->vx.data(), an->vy.data(), an->vz.data(), an->s4c.data(),
                            an->s4s.data(), an->s4c2.data(), an->s4s2.data(), an->cnt1.data(), an->dt1.data(),
                            an->qval.data()));
    }
    if (an->obs & AC4_OBS_FS) {
        snprintf(path, sizeof path, an->flag == 1 ? "%s/FsqtAA_a_%.2lf.dat" : "%s/Fsqt_a_%.2lf.dat", sub_dir, an->a);
        TRY(ac4_write_phase(path, 0, an->n, tot, ns, an->vx.data(), an->vy.data(), an->vz.data(), an->fsc.data(),
                            an->fss.data(), an->fsc2.data(), an->fss2.data(), an->cnt1.data(), an->dt1.data(),
                            an->qval.data()));
    }
    return AC4_OK;
}

extern "C" int ac4_sincos_turns_host(const double *f, int n, double *s, double *c) {
    AC4_REQUIRE(f && s && c && n >= 0, AC4_EARG, "sincos_turns: bad arguments");
    for (int i = 0; i < n; i++) ac4_sincos_turns(f[i], &s[i], &c[i]);
    return AC4_OK;
}

extern "C" int ac4_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}
