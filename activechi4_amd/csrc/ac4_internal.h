// ac4_internal.h -- shared helpers of libac4 (errors, Philox, wave reductions).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdarg.h>
#include <stdio.h>
#include <vector>
#include "ac4.h"
#include "ac4_host.h"

#define AC4_HIP(call)                                                              \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            ac4_set_error("%s:%d %s -> %s", __FILE__, __LINE__, #call,             \
                          hipGetErrorString(e_));                                  \
            return AC4_EHIP;                                                       \
        }                                                                          \
    } while (0)

#define AC4_REQUIRE(cond, code, ...)                                               \
    do {                                                                           \
        if (!(cond)) {                                                             \
            ac4_set_error(__VA_ARGS__);                                            \
            return (code);                                                         \
        }                                                                          \
    } while (0)

// Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC 2011): 10 rounds of two 32x32->64 multiplies with the
// published multipliers and Weyl key increments.
__host__ __device__ inline void ac4_philox(const uint32_t c_in[4], const uint32_t k_in[2],
                                           uint32_t out[4]) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t c0 = c_in[0], c1 = c_in[1], c2 = c_in[2], c3 = c_in[3];
    uint32_t k0 = k_in[0], k1 = k_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += W0; k1 += W1; }
        uint64_t p0 = (uint64_t)M0 * c0, p1 = (uint64_t)M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// 53-bit uniform in [0,1) from two 32-bit words (27 + 26 bits).
__host__ __device__ inline double ac4_u01(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

// Uniform number i of shuffle `s`, component `comp` of a replica keyed (k0,k1):
// w/2^32 with a 32-bit word, the resolution of gsl_rng_uniform for mt19937
// (the stream MDcode/active.c:59 draws from).  nact*u is then exact and below
// nact for any nact < 2^21, so the swap index stays in range.
__host__ __device__ inline double ac4_shuffle_uniform(uint32_t k0, uint32_t k1, uint64_t s,
                                                      int comp, int i) {
    uint32_t c[4] = {(uint32_t)i, (uint32_t)comp, (uint32_t)s, (uint32_t)(s >> 32)};
    uint32_t k[2] = {k0, k1}, o[4];
    ac4_philox(c, k, o);
    return (double)o[0] * (1.0 / 4294967296.0);
}

__device__ inline double ac4_wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
__device__ inline double ac4_wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
    return v;
}
__device__ inline int ac4_wave_sum_i(int v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// sin and cos of 2*pi*f for a phase f given in turns (|f| < 2^28).  The
// nearest quarter turn q = rint(4f) is taken off exactly (f - q/4 is exact
// by Sterbenz's lemma), leaving x = 2*pi*(f - q/4) in [-pi/4, pi/4]; sin x
// and cos x are their Taylor polynomials to x^15 and x^16, whose remainders
// are below 5e-17 and 3e-18 there, in Horner form.  The quarter turns then
// rotate (sin x, cos x).  Every operation is written out (explicit fma), so
// host and device evaluate it identically.
__host__ __device__ inline void ac4_sincos_turns(double f, double *s_out, double *c_out) {
    const double q = rint(4.0 * f);
    const double x = fma(-0.25, q, f) * 6.283185307179586;
    const double z = x * x;
    double ps = -1.0 / 1307674368000.0;              // -1/15!
    ps = fma(z, ps, 1.0 / 6227020800.0);             //  1/13!
    ps = fma(z, ps, -1.0 / 39916800.0);              // -1/11!
    ps = fma(z, ps, 1.0 / 362880.0);                 //  1/9!
    ps = fma(z, ps, -1.0 / 5040.0);                  // -1/7!
    ps = fma(z, ps, 1.0 / 120.0);                    //  1/5!
    ps = fma(z, ps, -1.0 / 6.0);                     // -1/3!
    const double sx = fma(x * z, ps, x);
    double pc = 1.0 / 20922789888000.0;              //  1/16!
    pc = fma(z, pc, -1.0 / 87178291200.0);           // -1/14!
    pc = fma(z, pc, 1.0 / 479001600.0);              //  1/12!
    pc = fma(z, pc, -1.0 / 3628800.0);               // -1/10!
    pc = fma(z, pc, 1.0 / 40320.0);                  //  1/8!
    pc = fma(z, pc, -1.0 / 720.0);                   // -1/6!
    pc = fma(z, pc, 1.0 / 24.0);                     //  1/4!
    pc = fma(z, pc, -0.5);                           // -1/2!
    const double cx = fma(z, pc, 1.0);
    // rotate by k = q mod 4 quarter turns: (s, c) -> (c, -s) -> (-s, -c) -> (-c, s)
    const int k = (int)q & 3;
    const double a = (k & 1) ? cx : sx, b = (k & 1) ? sx : cx;
    *s_out = (k & 2) ? -a : a;
    *c_out = ((k + 1) & 2) ? -b : b;
}

// Device view of a frame store (ac4_frames.hip) for the analyses built on it:
// frame k of particle i is x[k*n + i]; step[k] is its step index.
struct ac4_frames_view {
    int n, count;
    double L;
    const double *x, *y, *z;
    const long long *step;
    hipStream_t s;
};
int ac4_frames_view_get(const ac4_frames *f, ac4_frames_view *v);

// Frame step numbers as the reference's int frame indices (fidx[k]).
int ac4_frames_steps(const ac4_frames_view &v, std::vector<int> &fidx);

// Device allocations of one call, freed on every return path.
struct AcDevBuf {
    std::vector<void *> p;
    template <class T>
    int get(T **out, size_t count) {
        void *q = nullptr;
        if (hipMalloc(&q, count ? sizeof(T) * count : 8) != hipSuccess) {
            ac4_set_error("cannot allocate %zu bytes of device memory", sizeof(T) * count);
            return AC4_EHIP;
        }
        p.push_back(q);
        *out = (T *)q;
        return AC4_OK;
    }
    ~AcDevBuf() {
        for (void *q : p) (void)hipFree(q);
    }
};
