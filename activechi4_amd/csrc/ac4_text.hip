// ac4_text.hip -- data_coor.out text parsed on the GPU.
//
// Replaces the per-frame fscanf loop of Analysis/read_all_data.c:65-100
// (LoadFramePP: N lines "%lf %lf %lf %d" per frame, the text MDcode/print.c:10
// writes as "%.5lf\t%.5lf\t%.5lf\t%d").  The caller reads the file in
// chunks into a pinned buffer the handle owns; one call moves a chunk to the
// device, finds every line end (per-block newline counts, one scan, then
// each block writes its line ends in order) and converts one line per
// thread.  Whole frames of n lines come out as device arrays, ready for
// ac4_ana_push_frame_device / ac4_frames_push_device.
//
// Values equal the host reader's (ac4_parse_frame, itself equal to strtod):
// a number "[-]digits[.digits]" with at most 15 significant digits is the
// integer D of its digits divided by 10^f, one correctly rounded division,
// which is the value strtod returns for that decimal.  A line outside that
// form (exponents, longer mantissas, other text) is flagged, and the frames
// of the call are then parsed on the host by ac4_parse_frame instead.
#include <hip/hip_runtime.h>
#include <limits.h>
#include <string.h>
#include "ac4_internal.h"

#define TX_BLK 256
#define TX_PER 64                      // bytes per thread (four 16-byte loads)
#define TX_SEG (TX_BLK * TX_PER)       // bytes per block: 16 KiB
#define TX_SCAN 1024                   // threads of the block-count scan

struct ac4_textframes {
    int n = 0;
    long long cap = 0;                  // bytes of one chunk
    long long lines_cap = 0;            // line ends one chunk can hold
    char *h_text = nullptr;             // pinned host chunk (the caller fills it)
    char *d_text = nullptr;
    int *d_bcnt = nullptr, *d_boff = nullptr, *d_lend = nullptr, *d_meta = nullptr;  // meta: total, first bad
    double *d_x = nullptr, *d_y = nullptr, *d_z = nullptr;
    int *d_t = nullptr;
    int max_frames = 0;                 // frames the parsed arrays hold
    hipStream_t s = nullptr;
};

// Newlines of one 16 KiB segment (bytes at or past len do not count).
__global__ void __launch_bounds__(TX_BLK) k_tx_count(const char *__restrict__ text, long long len, int *bcnt) {
    const long long b0 = (long long)blockIdx.x * TX_SEG + (long long)threadIdx.x * TX_PER;
    int c = 0;
    if (b0 + TX_PER <= len) {
        const uint4 *p = (const uint4 *)(text + b0);
#pragma unroll
        for (int k = 0; k < TX_PER / 16; k++) {
            const uint4 v = p[k];
            const unsigned int w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int q = 0; q < 4; q++)
#pragma unroll
                for (int by = 0; by < 4; by++) c += ((w[q] >> (8 * by)) & 0xffu) == (unsigned)'\n';
        }
    } else {
        for (long long b = b0; b < len && b < b0 + TX_PER; b++) c += text[b] == '\n';
    }
    __shared__ int s_w[TX_BLK / 64];
    const int wsum = ac4_wave_sum_i(c);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = wsum;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < TX_BLK / 64; w++) t += s_w[w];
        bcnt[blockIdx.x] = t;
    }
}

// Exclusive scan of the nb segment counts (one block); meta[0] = total.
__global__ void __launch_bounds__(TX_SCAN) k_tx_scan(const int *__restrict__ bcnt, int nb, int *boff, int *meta) {
    __shared__ int s_w[TX_SCAN / 64];
    __shared__ int s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int base = 0; base < nb; base += TX_SCAN) {
        const int i = base + threadIdx.x;
        const int own = i < nb ? bcnt[i] : 0;
        int v = own;
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(v, o, 64);
            if (lane >= o) v += t;
        }
        if (lane == 63) s_w[w] = v;
        __syncthreads();
        int acc = s_carry + v - own;
        for (int k = 0; k < w; k++) acc += s_w[k];
        if (i < nb) boff[i] = acc;
        __syncthreads();
        if (threadIdx.x == TX_SCAN - 1) s_carry = acc + own;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        meta[0] = s_carry;
        meta[1] = INT_MAX;
    }
}

// Line ends in order: the byte index of every newline at boff[blk] + its
// rank in the segment (threads' counts scanned in the block), up to lcap.
__global__ void __launch_bounds__(TX_BLK) k_tx_ends(const char *__restrict__ text, long long len,
                                                    const int *__restrict__ boff, int *lend, long long lcap) {
    const long long b0 = (long long)blockIdx.x * TX_SEG + (long long)threadIdx.x * TX_PER;
    const long long b1 = b0 + TX_PER < len ? b0 + TX_PER : len;
    int c = 0;
    for (long long b = b0; b < b1; b++) c += text[b] == '\n';
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int v = c;
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    __shared__ int s_w[TX_BLK / 64];
    if (lane == 63) s_w[w] = v;
    __syncthreads();
    long long r = (long long)boff[blockIdx.x] + v - c;
    for (int k = 0; k < w; k++) r += s_w[k];
    for (long long b = b0; b < b1; b++)
        if (text[b] == '\n') {
            if (r < lcap) lend[r] = (int)b;
            r++;
        }
}

__constant__ double c_tx_p10[16] = {1e0, 1e1, 1e2, 1e3, 1e4, 1e5, 1e6, 1e7, 1e8, 1e9, 1e10, 1e11, 1e12, 1e13, 1e14, 1e15};

// "[-|+]digits[.digits]" after spaces or tabs, at most 15 significant
// digits, not followed by a digit or an exponent: true and the value.
__device__ __forceinline__ bool tx_num(const char *__restrict__ t, long long &p, long long e, double &out) {
    while (p < e && (t[p] == ' ' || t[p] == '\t')) p++;
    bool neg = false;
    if (p < e && (t[p] == '-' || t[p] == '+')) { neg = t[p] == '-'; p++; }
    unsigned long long d = 0;
    int nd = 0, nf = 0;
    while (p < e && t[p] >= '0' && t[p] <= '9' && nd < 16) { d = d * 10ull + (unsigned)(t[p] - '0'); p++; nd++; }
    if (p < e && t[p] == '.') {
        p++;
        while (p < e && t[p] >= '0' && t[p] <= '9' && nd < 16) {
            d = d * 10ull + (unsigned)(t[p] - '0');
            p++;
            nd++;
            nf++;
        }
    }
    if (nd == 0 || nd > 15) return false;
    if (p < e && ((t[p] >= '0' && t[p] <= '9') || t[p] == 'e' || t[p] == 'E')) return false;
    const double v = (double)d / c_tx_p10[nf];
    out = neg ? -v : v;
    return true;
}

// strtol's reading of the type column: spaces, a sign, 1 to 9 digits.
__device__ __forceinline__ bool tx_int(const char *__restrict__ t, long long &p, long long e, int &out) {
    while (p < e && (t[p] == ' ' || t[p] == '\t')) p++;
    bool neg = false;
    if (p < e && (t[p] == '-' || t[p] == '+')) { neg = t[p] == '-'; p++; }
    int v = 0, nd = 0;
    while (p < e && t[p] >= '0' && t[p] <= '9' && nd < 10) { v = v * 10 + (t[p] - '0'); p++; nd++; }
    if (nd == 0 || nd > 9) return false;
    out = neg ? -v : v;
    return true;
}

// One line per thread: line l runs from the byte after line end l-1 to line
// end l.  A line outside the fast form records its index in meta[1].
__global__ void __launch_bounds__(TX_BLK) k_tx_parse(const char *__restrict__ text, const int *__restrict__ lend,
                                                     int nlines, double *x, double *y, double *z, int *ty,
                                                     int *meta) {
    const int l = blockIdx.x * TX_BLK + threadIdx.x;
    if (l >= nlines) return;
    long long p = l == 0 ? 0 : (long long)lend[l - 1] + 1;
    const long long e = lend[l];
    double a = 0.0, b = 0.0, c = 0.0;
    int k = 0;
    const bool ok = tx_num(text, p, e, a) && tx_num(text, p, e, b) && tx_num(text, p, e, c) && tx_int(text, p, e, k);
    if (!ok) atomicMin(&meta[1], l);
    x[l] = a;
    y[l] = b;
    z[l] = c;
    ty[l] = k;
}

extern "C" int ac4_textframes_destroy(ac4_textframes *t) {
    if (!t) return AC4_OK;
    if (t->s) (void)hipStreamSynchronize(t->s);
    if (t->h_text) (void)hipHostFree(t->h_text);
    void *d[] = {t->d_text, t->d_bcnt, t->d_boff, t->d_lend, t->d_meta, t->d_x, t->d_y, t->d_z, t->d_t};
    for (void *q : d)
        if (q) (void)hipFree(q);
    if (t->s) (void)hipStreamDestroy(t->s);
    delete t;
    return AC4_OK;
}

extern "C" int ac4_textframes_create(int n, long long chunk_bytes, ac4_textframes **out) {
    AC4_REQUIRE(n > 0 && out && chunk_bytes >= 0, AC4_EARG, "textframes_create: bad arguments");
    int ndev = 0;
    AC4_REQUIRE(hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0, AC4_EHIP,
                "textframes_create: no HIP device visible (libac4 has no CPU path)");
    // a chunk holds at least one frame of lines up to 128 bytes, and at most 2 GiB
    long long cap = chunk_bytes > 0 ? chunk_bytes : (64ll << 20);
    if (cap < 128ll * n) cap = 128ll * n;
    AC4_REQUIRE(cap < (1ll << 31), AC4_EARG, "textframes_create: chunk of %lld bytes exceeds 2 GiB", cap);
    cap = (cap + TX_SEG - 1) / TX_SEG * TX_SEG;
    ac4_textframes *t = new ac4_textframes();
    t->n = n;
    t->cap = cap;
    t->lines_cap = cap / 8 + 1;  // a line is at least 8 bytes ("0 0 0 0\n")
    t->max_frames = (int)(t->lines_cap / n);
    const long long nb = cap / TX_SEG, ml = (long long)t->max_frames * n;
    auto fail = [&](int e) { ac4_textframes_destroy(t); return e; };
    if (hipStreamCreate(&t->s) != hipSuccess) return fail(AC4_EHIP);
    if (hipHostMalloc((void **)&t->h_text, (size_t)cap, hipHostMallocDefault) != hipSuccess) return fail(AC4_EHIP);
    if (hipMalloc(&t->d_text, (size_t)cap) != hipSuccess || hipMalloc(&t->d_bcnt, sizeof(int) * nb) != hipSuccess ||
        hipMalloc(&t->d_boff, sizeof(int) * nb) != hipSuccess ||
        hipMalloc(&t->d_lend, sizeof(int) * (size_t)t->lines_cap) != hipSuccess ||
        hipMalloc(&t->d_meta, sizeof(int) * 2) != hipSuccess || hipMalloc(&t->d_x, sizeof(double) * ml) != hipSuccess ||
        hipMalloc(&t->d_y, sizeof(double) * ml) != hipSuccess || hipMalloc(&t->d_z, sizeof(double) * ml) != hipSuccess ||
        hipMalloc(&t->d_t, sizeof(int) * ml) != hipSuccess) {
        ac4_set_error("textframes_create: cannot allocate the buffers of a %lld-byte chunk", cap);
        return fail(AC4_EHIP);
    }
    *out = t;
    return AC4_OK;
}

extern "C" char *ac4_textframes_buffer(ac4_textframes *t, long long *cap) {
    if (!t) return nullptr;
    if (cap) *cap = t->cap;
    return t->h_text;
}

extern "C" int ac4_textframes_parse(ac4_textframes *t, long long len, int max_frames, int *nframes,
                                    long long *used) {
    AC4_REQUIRE(t && nframes && used && len >= 0 && len <= t->cap && max_frames >= 0, AC4_EARG,
                "textframes_parse: bad arguments");
    *nframes = 0;
    *used = 0;
    if (len == 0 || max_frames == 0) return AC4_OK;
    hipStream_t s = t->s;
    const int nb = (int)((len + TX_SEG - 1) / TX_SEG);
    AC4_HIP(hipMemcpyAsync(t->d_text, t->h_text, (size_t)len, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_tx_count, dim3(nb), dim3(TX_BLK), 0, s, t->d_text, len, t->d_bcnt);
    hipLaunchKernelGGL(k_tx_scan, dim3(1), dim3(TX_SCAN), 0, s, t->d_bcnt, nb, t->d_boff, t->d_meta);
    hipLaunchKernelGGL(k_tx_ends, dim3(nb), dim3(TX_BLK), 0, s, t->d_text, len, t->d_boff, t->d_lend, t->lines_cap);
    AC4_HIP(hipGetLastError());
    int total = 0;
    AC4_HIP(hipMemcpyAsync(&total, t->d_meta, sizeof(int), hipMemcpyDeviceToHost, s));
    AC4_HIP(hipStreamSynchronize(s));
    const long long lines = total < t->lines_cap ? total : t->lines_cap;
    int nf = (int)(lines / t->n);
    if (nf > max_frames) nf = max_frames;
    if (nf > t->max_frames) nf = t->max_frames;
    if (nf == 0) return AC4_OK;  // not one whole frame in the chunk yet
    const int nl = nf * t->n;
    hipLaunchKernelGGL(k_tx_parse, dim3((nl + TX_BLK - 1) / TX_BLK), dim3(TX_BLK), 0, s, t->d_text, t->d_lend, nl,
                       t->d_x, t->d_y, t->d_z, t->d_t, t->d_meta);
    AC4_HIP(hipGetLastError());
    int meta[2] = {0, 0}, last = 0;
    AC4_HIP(hipMemcpyAsync(meta, t->d_meta, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
    AC4_HIP(hipMemcpyAsync(&last, t->d_lend + nl - 1, sizeof(int), hipMemcpyDeviceToHost, s));
    AC4_HIP(hipStreamSynchronize(s));
    if (meta[1] < nl) {
        // a line outside the fast form: these frames through the host reader
        const size_t fb = sizeof(double) * t->n;
        std::vector<double> hx(t->n), hy(t->n), hz(t->n);
        std::vector<int> ht(t->n);
        long long pos = 0;
        for (int f = 0; f < nf; f++) {
            const long long u = ac4_parse_frame(t->h_text + pos, len - pos, t->n, hx.data(), hy.data(), hz.data(),
                                                ht.data());
            if (u < 0) return (int)u;
            const size_t o = (size_t)f * t->n;
            AC4_HIP(hipMemcpyAsync(t->d_x + o, hx.data(), fb, hipMemcpyHostToDevice, s));
            AC4_HIP(hipMemcpyAsync(t->d_y + o, hy.data(), fb, hipMemcpyHostToDevice, s));
            AC4_HIP(hipMemcpyAsync(t->d_z + o, hz.data(), fb, hipMemcpyHostToDevice, s));
            AC4_HIP(hipMemcpyAsync(t->d_t + o, ht.data(), sizeof(int) * t->n, hipMemcpyHostToDevice, s));
            AC4_HIP(hipStreamSynchronize(s));
            pos += u;
        }
    }
    *nframes = nf;
    *used = (long long)last + 1;
    return AC4_OK;
}

extern "C" int ac4_textframes_get(ac4_textframes *t, int k, const double **x, const double **y, const double **z,
                                  const int **type) {
    AC4_REQUIRE(t && k >= 0 && k < t->max_frames, AC4_EARG, "textframes_get: bad arguments");
    const size_t o = (size_t)k * t->n;
    if (x) *x = t->d_x + o;
    if (y) *y = t->d_y + o;
    if (z) *z = t->d_z + o;
    if (type) *type = t->d_t + o;
    return AC4_OK;
}

extern "C" int ac4_textframes_copy(ac4_textframes *t, int k, double *x, double *y, double *z, int *type) {
    AC4_REQUIRE(t && k >= 0 && k < t->max_frames && x && y && z && type, AC4_EARG, "textframes_copy: bad arguments");
    const size_t o = (size_t)k * t->n, b = sizeof(double) * t->n;
    AC4_HIP(hipMemcpyAsync(x, t->d_x + o, b, hipMemcpyDeviceToHost, t->s));
    AC4_HIP(hipMemcpyAsync(y, t->d_y + o, b, hipMemcpyDeviceToHost, t->s));
    AC4_HIP(hipMemcpyAsync(z, t->d_z + o, b, hipMemcpyDeviceToHost, t->s));
    AC4_HIP(hipMemcpyAsync(type, t->d_t + o, sizeof(int) * t->n, hipMemcpyDeviceToHost, t->s));
    AC4_HIP(hipStreamSynchronize(t->s));
    return AC4_OK;
}
