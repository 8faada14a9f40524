// ac4_frames.hip -- a device-resident trajectory of one replica and the
// four-point structure factor S4(q, t) on |q| shells with seeded random
// directions, weighted by the per-particle overlap.
//
// No reference counterpart for the store: the reference writes every frame
// to data_coor.out and rereads it (Analysis/read_all_data.c:61-91).  Here the
// frames stay in HBM (x_u of 10^6 particles is 24 MB a frame, 500 frames
// 12 GB), so an analysis can pair any origin with any lag after the run.
//
// S4 on shells generalises Analysis/S4Qt.c (which uses the 162 lattice
// vectors of GetQ): for |q_k| = q0 + k*dq, k < nshell, and unit directions
// n_d, d < ndir, drawn from a seed,
//   C_{p,k,d} = sum_i w_i cos(|q_k| L n_d.s_i),  S likewise with sin,
// over particles i, with s_i the origin position of pair p wrapped into
// [0,1) and w_i = 1 when particle i overlaps between the origin and the lag
// frame of the pair (Analysis/overlap.c:61-84 arithmetic: unwrapped
// difference, one minimum-image correction, L*L*((dx*dx+dy*dy)+dz*dz) < a*a).
// From the C and S of all pairs (origins at one lag):
//   s4_ref[k] = ndir * [Var_p(Cbar) + Var_p(Sbar)] / N, Cbar = mean_d C
//               (the structure of S4Qt.c:133-267: the directions play the
//               role of the permutations, ndir the degeneracy),
//   s4_dir[k] = mean_d [Var_p(C) + Var_p(S)] / N  (the direction average of
//               (<|W(q)|^2> - |<W(q)>|^2)/N),
// with population variances over the pairs.
//
// Kernel: a block stages a tile of particles (weight and wrapped origin
// position) in LDS; each thread owns one direction and a run of SHELL_RUN
// consecutive shells and walks the tile: two sincos per particle, then the
// phase advances from shell to shell by one complex multiplication,
// exp(i q_{k+1} x) = exp(i q_k x) exp(i dq x), all in FP64.  Per-block partial
// sums are reduced over blocks in a fixed order, so results are deterministic.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <algorithm>
#include <vector>
#include "ac4_internal.h"

extern "C" int ac4_md_device_xu(ac4_md *md, int r, const double **xu, const double **yu, const double **zu,
                                const int **type, int *n);

#define TRY(e) do { int rc_ = (e); if (rc_ != AC4_OK) return rc_; } while (0)
#define FR_THREADS 256
#define FR_TILE 512
#define SHELL_RUN 27
#define FR_BLOCKS 240

struct ac4_frames {
    int n = 0, cap = 0, count = 0;
    double L = 0.0;
    double *x = nullptr, *y = nullptr, *z = nullptr;  // [cap][n], box-scaled x_u, reference order
    std::vector<long long> step;
    hipStream_t s = nullptr;
    double *part = nullptr;  // [FR_BLOCKS][ndir][nshell][2]
    size_t part_n = 0;
    double *cs = nullptr;    // [ndir][nshell][2] of one pair
    size_t cs_n = 0;
    double *dirs = nullptr;
    int dirs_n = 0;
    unsigned int *cnt = nullptr;
};

// Analysis/overlap.c:67-78: one minimum-image correction of an unwrapped difference
__device__ __forceinline__ double fr_minimg(double d) {
    if (d >= 0.5)
        d -= 1.0;
    else if (d < -0.5)
        d += 1.0;
    return d;
}

__global__ void __launch_bounds__(FR_THREADS) k_fr_overlap(const double *__restrict__ x, const double *__restrict__ y,
                                                           const double *__restrict__ z, int n, size_t fo, size_t ft,
                                                           double L, double a2, unsigned int *cnt) {
    const int i = blockIdx.x * FR_THREADS + threadIdx.x;
    int hit = 0;
    if (i < n) {
        double dx = x[ft + i] - x[fo + i], dy = y[ft + i] - y[fo + i], dz = z[ft + i] - z[fo + i];
        dz = fr_minimg(dz);
        dy = fr_minimg(dy);
        dx = fr_minimg(dx);
        hit = L * L * (dx * dx + dy * dy + dz * dz) < a2;
    }
    const int w = ac4_wave_sum_i(hit);
    if ((threadIdx.x & 63) == 0 && w) atomicAdd(cnt, (unsigned int)w);
}

__global__ void __launch_bounds__(FR_THREADS) k_fr_phase(const double *__restrict__ x, const double *__restrict__ y,
                                                         const double *__restrict__ z, int n, size_t fo, size_t ft,
                                                         double L, double a2, double q0, double dq, int nshell,
                                                         int ndir, int ngroup, const double *__restrict__ dirs,
                                                         double *part) {
    __shared__ double s_x[FR_TILE], s_y[FR_TILE], s_z[FR_TILE];
    __shared__ int s_w[FR_TILE];
    const int t = threadIdx.x, d = t / ngroup, g = t - d * ngroup;
    const bool active = d < ndir;
    const int k0 = g * SHELL_RUN, kn = active ? std::min(SHELL_RUN, nshell - k0) : 0;
    double nx = 0.0, ny = 0.0, nz = 0.0;
    if (active) { nx = dirs[3 * d]; ny = dirs[3 * d + 1]; nz = dirs[3 * d + 2]; }
    const double qa = (q0 + (double)k0 * dq) * L, qd = dq * L;  // phases per unit of n.s
    double ac[SHELL_RUN], as[SHELL_RUN];
#pragma unroll
    for (int k = 0; k < SHELL_RUN; k++) { ac[k] = 0.0; as[k] = 0.0; }
    const int per = (n + gridDim.x - 1) / gridDim.x;
    const int i0 = blockIdx.x * per, i1 = std::min(n, i0 + per);
    for (int tb = i0; tb < i1; tb += FR_TILE) {
        const int m = std::min(FR_TILE, i1 - tb);
        __syncthreads();
        for (int j = t; j < m; j += FR_THREADS) {
            const int i = tb + j;
            const double ox = x[fo + i], oy = y[fo + i], oz = z[fo + i];
            double dx = x[ft + i] - ox, dy = y[ft + i] - oy, dz = z[ft + i] - oz;
            dz = fr_minimg(dz);
            dy = fr_minimg(dy);
            dx = fr_minimg(dx);
            s_w[j] = L * L * (dx * dx + dy * dy + dz * dz) < a2;
            s_x[j] = ox - floor(ox);
            s_y[j] = oy - floor(oy);
            s_z[j] = oz - floor(oz);
        }
        __syncthreads();
        if (kn <= 0) continue;
        for (int j = 0; j < m; j++) {
            if (!s_w[j]) continue;  // uniform over the block
            const double p = nx * s_x[j] + ny * s_y[j] + nz * s_z[j];
            double c, s, cd, sd;
            sincos(qa * p, &s, &c);
            sincos(qd * p, &sd, &cd);
#pragma unroll
            for (int k = 0; k < SHELL_RUN; k++) {
                ac[k] += c;
                as[k] += s;
                const double cn = c * cd - s * sd;
                s = s * cd + c * sd;
                c = cn;
            }
        }
    }
    if (!active) return;
    double *o = part + (size_t)blockIdx.x * ndir * nshell * 2 + ((size_t)d * nshell + k0) * 2;
    for (int k = 0; k < kn; k++) { o[2 * k] = ac[k]; o[2 * k + 1] = as[k]; }
}

// cs[e] = sum over blocks b = 0, 1, ... of part[b][e] (fixed order)
__global__ void k_fr_reduce(const double *__restrict__ part, int nblk, int ne, double *cs) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    double a = 0.0;
    for (int b = 0; b < nblk; b++) a += part[(size_t)b * ne + e];
    cs[e] = a;
}

extern "C" int ac4_frames_create(int n, double L, int cap, ac4_frames **out) {
    AC4_REQUIRE(n > 0 && L > 0 && cap > 0 && out, AC4_EARG, "ac4_frames_create: bad arguments");
    int ndev = 0;
    AC4_REQUIRE(hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0, AC4_EHIP,
                "ac4_frames_create: no HIP device visible (libac4 has no CPU path)");
    ac4_frames *f = new ac4_frames();
    f->n = n;
    f->L = L;
    f->cap = cap;
    const size_t bytes = sizeof(double) * (size_t)n * cap;
    if (hipMalloc(&f->x, bytes) != hipSuccess || hipMalloc(&f->y, bytes) != hipSuccess ||
        hipMalloc(&f->z, bytes) != hipSuccess || hipMalloc(&f->cnt, sizeof(unsigned int)) != hipSuccess ||
        hipStreamCreate(&f->s) != hipSuccess) {
        ac4_frames_destroy(f);
        ac4_set_error("ac4_frames_create: cannot allocate %zu bytes of frames", 3 * bytes);
        return AC4_EHIP;
    }
    *out = f;
    return AC4_OK;
}

extern "C" int ac4_frames_destroy(ac4_frames *f) {
    if (!f) return AC4_OK;
    if (f->s) (void)hipStreamSynchronize(f->s);
    void *p[] = {f->x, f->y, f->z, f->part, f->cs, f->dirs, f->cnt};
    for (void *q : p)
        if (q) (void)hipFree(q);
    if (f->s) (void)hipStreamDestroy(f->s);
    delete f;
    return AC4_OK;
}

extern "C" int ac4_frames_count(const ac4_frames *f) { return f ? f->count : 0; }

int ac4_frames_steps(const ac4_frames_view &v, std::vector<int> &fidx) {
    fidx.resize(v.count);
    for (int k = 0; k < v.count; k++) {
        AC4_REQUIRE(v.step[k] >= 0 && v.step[k] < 2147483647LL, AC4_EARG,
                    "frame step %lld does not fit the reference's int frame index", v.step[k]);
        fidx[k] = (int)v.step[k];
    }
    return AC4_OK;
}

int ac4_frames_view_get(const ac4_frames *f, ac4_frames_view *v) {
    AC4_REQUIRE(f && v, AC4_EARG, "frames view: bad arguments");
    v->n = f->n;
    v->count = f->count;
    v->L = f->L;
    v->x = f->x;
    v->y = f->y;
    v->z = f->z;
    v->step = f->step.data();
    v->s = f->s;
    return AC4_OK;
}

extern "C" int ac4_frames_find(const ac4_frames *f, long long step) {
    if (!f) return -1;
    for (int k = 0; k < f->count; k++)
        if (f->step[k] == step) return k;
    return -1;
}

extern "C" int ac4_frames_push(ac4_frames *f, const double *xu, const double *yu, const double *zu, long long step) {
    AC4_REQUIRE(f && xu && yu && zu, AC4_EARG, "frames_push: bad arguments");
    AC4_REQUIRE(f->count < f->cap, AC4_ECAP, "frames_push: store full (%d frames)", f->cap);
    const size_t o = (size_t)f->count * f->n, b = sizeof(double) * f->n;
    AC4_HIP(hipMemcpyAsync(f->x + o, xu, b, hipMemcpyHostToDevice, f->s));
    AC4_HIP(hipMemcpyAsync(f->y + o, yu, b, hipMemcpyHostToDevice, f->s));
    AC4_HIP(hipMemcpyAsync(f->z + o, zu, b, hipMemcpyHostToDevice, f->s));
    AC4_HIP(hipStreamSynchronize(f->s));
    f->step.push_back(step);
    f->count++;
    return AC4_OK;
}

extern "C" int ac4_frames_push_device(ac4_frames *f, const double *xu, const double *yu, const double *zu,
                                      long long step) {
    AC4_REQUIRE(f && xu && yu && zu, AC4_EARG, "frames_push_device: bad arguments");
    AC4_REQUIRE(f->count < f->cap, AC4_ECAP, "frames_push_device: store full (%d frames)", f->cap);
    const size_t o = (size_t)f->count * f->n, b = sizeof(double) * f->n;
    AC4_HIP(hipMemcpyAsync(f->x + o, xu, b, hipMemcpyDeviceToDevice, f->s));
    AC4_HIP(hipMemcpyAsync(f->y + o, yu, b, hipMemcpyDeviceToDevice, f->s));
    AC4_HIP(hipMemcpyAsync(f->z + o, zu, b, hipMemcpyDeviceToDevice, f->s));
    AC4_HIP(hipStreamSynchronize(f->s));
    f->step.push_back(step);
    f->count++;
    return AC4_OK;
}

extern "C" int ac4_frames_push_md(ac4_frames *f, ac4_md *md, int r, long long step) {
    AC4_REQUIRE(f && md, AC4_EARG, "frames_push_md: bad arguments");
    AC4_REQUIRE(f->count < f->cap, AC4_ECAP, "frames_push_md: store full (%d frames)", f->cap);
    const double *xu, *yu, *zu;
    const int *tp;
    int n = 0;
    TRY(ac4_md_device_xu(md, r, &xu, &yu, &zu, &tp, &n));  // synchronises the MD stream
    AC4_REQUIRE(n == f->n, AC4_EARG, "frames_push_md: replica has %d particles, store %d", n, f->n);
    const size_t o = (size_t)f->count * f->n, b = sizeof(double) * f->n;
    AC4_HIP(hipMemcpyAsync(f->x + o, xu, b, hipMemcpyDeviceToDevice, f->s));
    AC4_HIP(hipMemcpyAsync(f->y + o, yu, b, hipMemcpyDeviceToDevice, f->s));
    AC4_HIP(hipMemcpyAsync(f->z + o, zu, b, hipMemcpyDeviceToDevice, f->s));
    AC4_HIP(hipStreamSynchronize(f->s));
    f->step.push_back(step);
    f->count++;
    return AC4_OK;
}

extern "C" int ac4_frames_overlap(ac4_frames *f, int i0, int i1, double a, int *count) {
    AC4_REQUIRE(f && count && a > 0 && i0 >= 0 && i1 >= 0 && i0 < f->count && i1 < f->count, AC4_EARG,
                "frames_overlap: bad arguments");
    AC4_HIP(hipMemsetAsync(f->cnt, 0, sizeof(unsigned int), f->s));
    hipLaunchKernelGGL(k_fr_overlap, dim3((f->n + FR_THREADS - 1) / FR_THREADS), dim3(FR_THREADS), 0, f->s, f->x,
                       f->y, f->z, f->n, (size_t)i0 * f->n, (size_t)i1 * f->n, f->L, a * a, f->cnt);
    AC4_HIP(hipGetLastError());
    unsigned int c = 0;
    AC4_HIP(hipMemcpyAsync(&c, f->cnt, sizeof c, hipMemcpyDeviceToHost, f->s));
    AC4_HIP(hipStreamSynchronize(f->s));
    *count = (int)c;
    return AC4_OK;
}

// Unit directions from a seed: Philox-4x32-10 with counter (d, 0x5334, 0, 0),
// z = 1 - 2u, phi = 2 pi u' (uniform on the sphere, Archimedes).
extern "C" int ac4_shell_directions(int ndir, uint64_t seed, double *n3) {
    AC4_REQUIRE(ndir > 0 && n3, AC4_EARG, "shell_directions: bad arguments");
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int d = 0; d < ndir; d++) {
        uint32_t c[4] = {(uint32_t)d, 0x5334u, 0u, 0u}, o[4];
        ac4_philox(c, key, o);
        const double u1 = ((double)o[0] + 0.5) * (1.0 / 4294967296.0), u2 = ((double)o[1] + 0.5) * (1.0 / 4294967296.0);
        const double zz = 1.0 - 2.0 * u1, rr = sqrt(std::max(0.0, 1.0 - zz * zz)), ph = 6.283185307179586 * u2;
        n3[3 * d] = rr * cos(ph);
        n3[3 * d + 1] = rr * sin(ph);
        n3[3 * d + 2] = zz;
    }
    return AC4_OK;
}

// Variance reductions of S4 on shells from the per-pair phase sums
// ph[p][d][k][2] (host memory), as described at the top of this file.
static void s4_reduce_host(const double *ph, int npairs, int ndir, int nshell, double N, double *s4_ref,
                           double *s4_dir) {
    const size_t ne = (size_t)ndir * nshell * 2;
    const double inv_p = 1.0 / npairs;
    for (int k = 0; k < nshell; k++) {
        double mc = 0, mc2 = 0, ms = 0, ms2 = 0, dir_sum = 0;
        for (int p = 0; p < npairs; p++) {
            double c = 0, s = 0;
            for (int d = 0; d < ndir; d++) {
                const double *e = ph + (size_t)p * ne + ((size_t)d * nshell + k) * 2;
                c += e[0];
                s += e[1];
            }
            c /= ndir;
            s /= ndir;
            mc += c; mc2 += c * c; ms += s; ms2 += s * s;
        }
        mc *= inv_p; mc2 *= inv_p; ms *= inv_p; ms2 *= inv_p;
        for (int d = 0; d < ndir; d++) {
            double c1 = 0, c2 = 0, s1 = 0, s2 = 0;
            for (int p = 0; p < npairs; p++) {
                const double *e = ph + (size_t)p * ne + ((size_t)d * nshell + k) * 2;
                c1 += e[0]; c2 += e[0] * e[0]; s1 += e[1]; s2 += e[1] * e[1];
            }
            c1 *= inv_p; c2 *= inv_p; s1 *= inv_p; s2 *= inv_p;
            dir_sum += (c2 - c1 * c1) + (s2 - s1 * s1);
        }
        if (s4_ref) s4_ref[k] = ndir * ((mc2 - mc * mc) + (ms2 - ms * ms)) / N;
        if (s4_dir) s4_dir[k] = dir_sum / ndir / N;
    }
}

static int s4_check_args(ac4_frames *f, int npairs, const int *orig, const int *lagf, double a, int nshell, int ndir,
                         double dq) {
    AC4_REQUIRE(f && npairs > 0 && orig && lagf && a > 0 && nshell > 0 && ndir > 0 && dq >= 0, AC4_EARG,
                "frames_s4_shells: bad arguments");
    const int ngroup = (nshell + SHELL_RUN - 1) / SHELL_RUN;
    AC4_REQUIRE(ndir * ngroup <= FR_THREADS, AC4_EARG,
                "frames_s4_shells: ndir * ceil(nshell/%d) must be <= %d", SHELL_RUN, FR_THREADS);
    for (int p = 0; p < npairs; p++)
        AC4_REQUIRE(orig[p] >= 0 && orig[p] < f->count && lagf[p] >= 0 && lagf[p] < f->count, AC4_EARG,
                    "frames_s4_shells: pair %d refers to a frame not in the store", p);
    return AC4_OK;
}

extern "C" int ac4_frames_s4_shells(ac4_frames *f, int npairs, const int *orig, const int *lagf, double a, double q0,
                                    double dq, int nshell, int ndir, uint64_t seed, double *s4_ref, double *s4_dir,
                                    double *phases) {
    TRY(s4_check_args(f, npairs, orig, lagf, a, nshell, ndir, dq));
    const int ngroup = (nshell + SHELL_RUN - 1) / SHELL_RUN;
    const size_t ne = (size_t)ndir * nshell * 2;
    if (f->part_n < FR_BLOCKS * ne) {
        if (f->part) (void)hipFree(f->part);
        f->part = nullptr;
        AC4_HIP(hipMalloc(&f->part, sizeof(double) * FR_BLOCKS * ne));
        f->part_n = FR_BLOCKS * ne;
    }
    if (f->cs_n < ne) {
        if (f->cs) (void)hipFree(f->cs);
        f->cs = nullptr;
        AC4_HIP(hipMalloc(&f->cs, sizeof(double) * ne));
        f->cs_n = ne;
    }
    if (f->dirs_n < ndir) {
        if (f->dirs) (void)hipFree(f->dirs);
        f->dirs = nullptr;
        AC4_HIP(hipMalloc(&f->dirs, sizeof(double) * 3 * ndir));
        f->dirs_n = ndir;
    }
    std::vector<double> dir(3 * (size_t)ndir), ph((size_t)npairs * ne);
    TRY(ac4_shell_directions(ndir, seed, dir.data()));
    AC4_HIP(hipMemcpyAsync(f->dirs, dir.data(), sizeof(double) * dir.size(), hipMemcpyHostToDevice, f->s));
    for (int p = 0; p < npairs; p++) {
        hipLaunchKernelGGL(k_fr_phase, dim3(FR_BLOCKS), dim3(FR_THREADS), 0, f->s, f->x, f->y, f->z, f->n,
                           (size_t)orig[p] * f->n, (size_t)lagf[p] * f->n, f->L, a * a, q0, dq, nshell, ndir, ngroup,
                           f->dirs, f->part);
        hipLaunchKernelGGL(k_fr_reduce, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, f->s, f->part, FR_BLOCKS,
                           (int)ne, f->cs);
        AC4_HIP(hipGetLastError());
        AC4_HIP(hipMemcpyAsync(ph.data() + (size_t)p * ne, f->cs, sizeof(double) * ne, hipMemcpyDeviceToHost, f->s));
    }
    AC4_HIP(hipStreamSynchronize(f->s));
    if (phases) memcpy(phases, ph.data(), sizeof(double) * ph.size());
    s4_reduce_host(ph.data(), npairs, ndir, nshell, (double)f->n, s4_ref, s4_dir);
    return AC4_OK;
}

// ---------------------------------------------- S4 on shells, allocating nothing
// The per-block partial sums, the per-pair sums and the directions live in
// one file-scope device array instead of buffers of the frame store, so a
// call makes no device or host allocation.  Each GPU holds its own copy of
// the array until the process ends; its address on the current device is
// looked up once (hipGetSymbolAddress) and kept per device.  Calls that
// overlap on one GPU share the array and race.
// ndir * ceil(nshell / SHELL_RUN) <= FR_THREADS bounds ndir * nshell by
// FR_THREADS * SHELL_RUN = 6912, so ne = 2 * ndir * nshell <= 13824.
#define S4_NOALLOC_NE_MAX (2 * FR_THREADS * SHELL_RUN)
#define S4_NOALLOC_DIR_MAX FR_THREADS
#define S4_SCRATCH_DOUBLES ((size_t)FR_BLOCKS * S4_NOALLOC_NE_MAX + S4_NOALLOC_NE_MAX + 3 * S4_NOALLOC_DIR_MAX)
static __device__ __align__(256) double scratch[S4_SCRATCH_DOUBLES];  // 26.7 MB per GPU
#define S4_MAX_DEV 64
static double *g_scratch_ptr[S4_MAX_DEV];

static int scratch_ptr(double **out) {
    int dev = 0;
    AC4_HIP(hipGetDevice(&dev));
    AC4_REQUIRE(dev >= 0 && dev < S4_MAX_DEV, AC4_EARG, "s4_shells_noalloc: device %d beyond %d", dev, S4_MAX_DEV);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        AC4_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        AC4_REQUIRE(p != nullptr, AC4_EHIP, "s4_shells_noalloc: no device address for the scratch array");
        g_scratch_ptr[dev] = (double *)p;
    }
    *out = g_scratch_ptr[dev];
    return AC4_OK;
}

extern "C" int ac4_frames_s4_shells_noalloc(ac4_frames *f, int npairs, const int *orig, const int *lagf, double a,
                                            double q0, double dq, int nshell, int ndir, uint64_t seed,
                                            double *s4_ref, double *s4_dir, double *phases) {
    TRY(s4_check_args(f, npairs, orig, lagf, a, nshell, ndir, dq));
    AC4_REQUIRE(phases, AC4_EARG, "s4_shells_noalloc: phases (npairs*ndir*nshell*2 doubles) is required");
    const size_t ne = (size_t)ndir * nshell * 2;
    AC4_REQUIRE(ne <= S4_NOALLOC_NE_MAX, AC4_ECAP,
                "s4_shells_noalloc: ndir * nshell = %d exceeds the scratch limit of %d", ndir * nshell,
                S4_NOALLOC_NE_MAX / 2);
    const int ngroup = (nshell + SHELL_RUN - 1) / SHELL_RUN;
    double *sp = nullptr;
    TRY(scratch_ptr(&sp));
    double *part = sp, *cs = sp + (size_t)FR_BLOCKS * S4_NOALLOC_NE_MAX, *dirs = cs + S4_NOALLOC_NE_MAX;
    double dir[3 * S4_NOALLOC_DIR_MAX];  // ndir <= FR_THREADS (checked above)
    TRY(ac4_shell_directions(ndir, seed, dir));
    AC4_HIP(hipMemcpyAsync(dirs, dir, sizeof(double) * 3 * ndir, hipMemcpyHostToDevice, f->s));
    for (int p = 0; p < npairs; p++) {
        hipLaunchKernelGGL(k_fr_phase, dim3(FR_BLOCKS), dim3(FR_THREADS), 0, f->s, f->x, f->y, f->z, f->n,
                           (size_t)orig[p] * f->n, (size_t)lagf[p] * f->n, f->L, a * a, q0, dq, nshell, ndir, ngroup,
                           dirs, part);
        hipLaunchKernelGGL(k_fr_reduce, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, f->s, part, FR_BLOCKS,
                           (int)ne, cs);
        AC4_HIP(hipGetLastError());
        AC4_HIP(hipMemcpyAsync(phases + (size_t)p * ne, cs, sizeof(double) * ne, hipMemcpyDeviceToHost, f->s));
    }
    AC4_HIP(hipStreamSynchronize(f->s));
    s4_reduce_host(phases, npairs, ndir, nshell, (double)f->n, s4_ref, s4_dir);
    return AC4_OK;
}
