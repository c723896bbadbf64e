// Detection on a spectrum line (reference src/detect.c:4-114) on gfx950:
// the median (a robust noise floor) and detect_channels, for lines in device
// memory (the served display lines) and for the drop-in.
//
// One block of DT threads per line; a line of up to 4096 bins and its mask
// are held in LDS (longer ones use a float scratch of n per line in global
// memory):
//   smooth   the 5-bin moving average with edge padding, summed in the
//            reference's order (x[i-2] + ... + x[i+2], from 0.0f) times 0.2f;
//   floor    optionally the median of the smoothed line (signal_id's floor),
//            by an exact radix select on the floats' order keys;
//   mask     smoothed > floor + threshold, as 1.0f / 0.0f in the scratch;
//   bridge   every zero run with a one on both sides and fewer than gap
//            zeros becomes ones. The reference walks the runs in order, but
//            it only ever turns a run's own zeros to ones, so each run's
//            fate depends on the original mask alone: here each thread owns
//            a contiguous segment and learns the nearest one on either side
//            of it from two block scans (wavefront shuffles, then the
//            sixteen wave totals);
//   runs     the ones runs at least min_run wide, numbered in bin order by a
//            block scan; the thread that owns a run's first bin writes its
//            record: peak (the reference's scan from line[start]), the
//            power-weighted centroid (FP64, the reference's order), snr.
// The first `cap` records in bin order are written and the total found is
// reported; ordering by SNR is the host's (the reference sorts with qsort).
#include <hip/hip_runtime.h>
#include "atkgpu_dev.h"
#include "dev_common.hpp"

#pragma clang fp contract(off)

namespace {

constexpr int DT = 1024;

// float -> unsigned key with the same order (negatives reversed below positives)
__device__ __forceinline__ uint32_t ord_key(float v) {
    const uint32_t u = __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_float(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

struct SelectLds {
    uint32_t hist[256];
    unsigned long long k;
    uint32_t prefix;
};

// inclusive scans within a wavefront (64 lanes)
template <class T, class Op>
__device__ __forceinline__ T wave_scan_up(T v, Op op) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const T w = __shfl_up(v, o, 64);
        if (lane >= o) v = op(v, w);
    }
    return v;
}
template <class T, class Op>
__device__ __forceinline__ T wave_scan_down(T v, Op op) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const T w = __shfl_down(v, o, 64);
        if (lane + o < 64) v = op(v, w);
    }
    return v;
}

// exclusive block scans over the DT threads' values, in thread order
// (prefix) or reverse thread order (suffix); `part` is 16 slots of LDS; the
// caller must not reuse `part` before a barrier
constexpr int NW = DT / 64;
template <class T, class Op>
__device__ T block_excl_prefix(T v, T ident, Op op, T *part) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const T incl = wave_scan_up(v, op);
    T ex = __shfl_up(incl, 1, 64);
    if (lane == 0) ex = ident;
    if (lane == 63) part[w] = incl;
    __syncthreads();
    if (w == 0) {
        T t = lane < NW ? part[lane] : ident;
        t = wave_scan_up(t, op);
        if (lane < NW) part[lane] = t;
    }
    __syncthreads();
    return w > 0 ? op(part[w - 1], ex) : ex;
}
template <class T, class Op>
__device__ T block_excl_suffix(T v, T ident, Op op, T *part) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const T incl = wave_scan_down(v, op);
    T ex = __shfl_down(incl, 1, 64);
    if (lane == 63) ex = ident;
    if (lane == 0) part[w] = incl;
    __syncthreads();
    if (w == 0) {
        T t = lane < NW ? part[lane] : ident;
        // suffix over the NW wave totals
#pragma unroll
        for (int o = 1; o < NW; o <<= 1) {
            const T u = __shfl_down(t, o, 64);
            if (lane + o < NW) t = op(t, u);
        }
        if (lane < NW) part[lane] = t;
    }
    __syncthreads();
    return w + 1 < NW ? op(part[w + 1], ex) : ex;
}

struct OpMax { __device__ long long operator()(long long a, long long b) const { return a > b ? a : b; } };
struct OpMin { __device__ long long operator()(long long a, long long b) const { return a < b ? a : b; } };
struct OpAdd { __device__ unsigned long long operator()(unsigned long long a, unsigned long long b) const { return a + b; } };

// the k-th smallest (0-based) of x[0..n): four 8-bit passes over the order
// key, most significant first; one wavefront finds each pass's digit from the
// 256-bin histogram (a scan of 4-bin lane sums and a ballot); every thread
// of the block calls it and gets the value
__device__ __forceinline__ float block_select(const float *x, size_t n, size_t k, SelectLds &s) {
    uint32_t prefix = 0, pmask = 0;
    if (threadIdx.x == 0) s.k = k;
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (int b = threadIdx.x; b < 256; b += blockDim.x) s.hist[b] = 0;
        __syncthreads();
        for (size_t i = threadIdx.x; i < n; i += blockDim.x) {
            const uint32_t key = ord_key(x[i]);
            if ((key & pmask) == prefix) atomicAdd(&s.hist[(key >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (threadIdx.x < 64) {
            const int lane = threadIdx.x;
            const uint32_t c0 = s.hist[4 * lane], c1 = s.hist[4 * lane + 1], c2 = s.hist[4 * lane + 2],
                           c3 = s.hist[4 * lane + 3];
            const unsigned long long mine = (unsigned long long)c0 + c1 + c2 + c3;
            const unsigned long long incl = wave_scan_up(mine, OpAdd());
            const unsigned long long kk = s.k;
            const unsigned long long hit = __ballot(incl > kk);
            const int L = __ffsll((long long)hit) - 1;
            if (lane == L) {
                unsigned long long rem = kk - (incl - mine);
                uint32_t b = 4u * (uint32_t)lane;
                if (rem >= c0) {
                    rem -= c0; ++b;
                    if (rem >= c1) {
                        rem -= c1; ++b;
                        if (rem >= c2) { rem -= c2; ++b; }
                    }
                }
                s.k = rem;
                s.prefix = prefix | (b << shift);
            }
        }
        __syncthreads();
        prefix = s.prefix;
        pmask |= 255u << shift;
    }
    __syncthreads();
    return key_float(prefix);
}

// the median as the reference defines it: the middle value, or for even n
// 0.5f * (lower middle + upper middle)
__device__ __forceinline__ float block_median(const float *x, size_t n, SelectLds &s) {
    if (n & 1) return block_select(x, n, n / 2, s);
    const float hi = block_select(x, n, n / 2, s);
    const float lo = block_select(x, n, n / 2 - 1, s);
    return 0.5f * (lo + hi);
}

// run [s, e) of the bridged mask starting in this thread's segment
__device__ __forceinline__ size_t run_end(const float *m, size_t s, size_t n) {
    size_t e = s;
    while (e < n && m[e] != 0.0f) ++e;
    return e;
}

// lines up to this many bins (every display line) are held in LDS, line and
// mask, 32 KiB in all; the instantiation for them indexes the LDS arrays by
// name, so no access goes through a generic (flat) pointer into LDS
constexpr int LDS_BINS = 4096;

template <bool IN_LDS>
__global__ __launch_bounds__(DT) void k_detect(const float *__restrict__ lines, const float *const *__restrict__ ptrs,
                                               size_t lstride, size_t n,
                                               float floor_db, int compute_floor, float threshold_db, int gap,
                                               int min_run, float *__restrict__ scratch,
                                               atkgpu_chan *__restrict__ out, size_t cap,
                                               unsigned long long *__restrict__ found, float *__restrict__ floors) {
    __shared__ SelectLds sel;
    __shared__ long long part[NW];
    __shared__ unsigned long long upart[NW];
    __shared__ float s_floor;
    __shared__ float x_lds[IN_LDS ? LDS_BINS : 1], m_lds[IN_LDS ? LDS_BINS : 1];
    const float *xg = ptrs ? ptrs[blockIdx.x] : lines + blockIdx.x * lstride;
    const float *x;
    float *sm;
    if constexpr (IN_LDS) {
        for (size_t i = threadIdx.x; i < n; i += DT) x_lds[i] = xg[i];
        x = x_lds;
        sm = m_lds;
    } else {
        x = xg;
        sm = scratch + blockIdx.x * n;
    }
    atkgpu_chan *o = out + blockIdx.x * cap;
    const int t = threadIdx.x;
    __syncthreads();

    for (size_t i = t; i < n; i += DT) {
        float v;
        if (n < 5) {
            v = x[i];
        } else {
            float s = 0.0f;
            for (int d = -2; d <= 2; ++d) {
                long long j = (long long)i + d;
                if (j < 0) j = 0;
                if (j >= (long long)n) j = (long long)n - 1;
                s += x[j];
            }
            v = s * 0.2f;
        }
        sm[i] = v;
    }
    __syncthreads();
    if (compute_floor) {
        const float f = block_median(sm, n, sel);
        if (t == 0) s_floor = f;
    } else if (t == 0) {
        s_floor = floor_db;
    }
    __syncthreads();
    const float fl = s_floor;
    const float thr = fl + threshold_db;
    for (size_t i = t; i < n; i += DT) sm[i] = sm[i] > thr ? 1.0f : 0.0f;
    __syncthreads();

    // bridge: each thread owns [a0, a1)
    const size_t L = (n + DT - 1) / DT;
    const size_t a0 = (size_t)t * L < n ? (size_t)t * L : n;
    const size_t a1 = a0 + L < n ? a0 + L : n;
    long long first = (long long)n, last = -1;
    for (size_t i = a0; i < a1; ++i)
        if (sm[i] != 0.0f) {
            if (first == (long long)n) first = (long long)i;
            last = (long long)i;
        }
    // the last one strictly left of this segment, the first strictly right
    const long long left = block_excl_prefix(last, -1LL, OpMax(), part);
    __syncthreads();
    const long long right = block_excl_suffix(first, (long long)n, OpMin(), part);
    {
        long long p = left;
        size_t i = a0;
        while (i < a1) {
            if (sm[i] != 0.0f) { p = (long long)i; ++i; continue; }
            size_t j = i;
            while (j < a1 && sm[j] == 0.0f) ++j;
            const long long q = j < a1 ? (long long)j : right;
            if (p >= 0 && q < (long long)n && q - p - 1 < gap)
                for (size_t k = i; k < j; ++k) sm[k] = 1.0f;
            i = j;
        }
    }
    __syncthreads();

    // runs: count this segment's qualifying starts, number them, write
    unsigned long long cnt = 0;
    for (size_t i = a0; i < a1;) {
        if (sm[i] == 0.0f || (i > 0 && sm[i - 1] != 0.0f)) { ++i; continue; }
        const size_t e = run_end(sm, i, n);
        if ((long long)(e - i) >= min_run) ++cnt;
        i = e;
    }
    unsigned long long idx = block_excl_prefix(cnt, 0ULL, OpAdd(), upart);
    const unsigned long long total = upart[NW - 1];
    for (size_t i = a0; i < a1;) {
        if (sm[i] == 0.0f || (i > 0 && sm[i - 1] != 0.0f)) { ++i; continue; }
        const size_t e = run_end(sm, i, n);
        if ((long long)(e - i) >= min_run) {
            if (idx < cap) {
                double wsum = 0.0, csum = 0.0;
                float peak = x[i];
                for (size_t k = i; k < e; ++k) {
                    const double w = pow(10.0, (double)x[k] / 10.0);
                    wsum += w;
                    csum += (double)(k - i) * w;
                    if (x[k] > peak) peak = x[k];
                }
                atkgpu_chan c;
                c.start_bin = (int)i;
                c.end_bin = (int)e;
                c.centroid_bin = (float)((double)i + csum / (wsum + 1e-12));
                c.peak_db = peak;
                c.snr_db = peak - fl;
                o[idx] = c;
            }
            ++idx;
        }
        i = e;
    }
    if (t == 0) {
        found[blockIdx.x] = total;
        if (floors) floors[blockIdx.x] = fl;
    }
}

__global__ __launch_bounds__(DT) void k_median(const float *__restrict__ x, size_t n, float *__restrict__ out) {
    __shared__ SelectLds sel;
    const float m = block_median(x, n, sel);
    if (threadIdx.x == 0) *out = m;
}

}  // namespace

static int detect_launch(void *s, const float *lines, const float *const *ptrs, size_t lstride, size_t n,
                         unsigned nlines, float floor_db, int compute_floor, float threshold_db, int gap, int min_run,
                         float *scratch, atkgpu_chan *out, size_t cap, unsigned long long *found, float *floors) {
    if (n <= (size_t)LDS_BINS)
        hipLaunchKernelGGL(k_detect<true>, dim3(nlines), dim3(DT), 0, (hipStream_t)s, lines, ptrs, lstride, n,
                           floor_db, compute_floor, threshold_db, gap, min_run, scratch, out, cap, found, floors);
    else
        hipLaunchKernelGGL(k_detect<false>, dim3(nlines), dim3(DT), 0, (hipStream_t)s, lines, ptrs, lstride, n,
                           floor_db, compute_floor, threshold_db, gap, min_run, scratch, out, cap, found, floors);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

extern "C" int atkgpu_detect_lines(void *s, const float *lines, size_t lstride, size_t n, unsigned nlines,
                                   float floor_db, int compute_floor, float threshold_db, int gap, int min_run,
                                   float *scratch, atkgpu_chan *out, size_t cap, unsigned long long *found,
                                   float *floors) {
    if (!lines || !scratch || !found || n == 0 || nlines == 0 || (!out && cap)) return -1;
    return detect_launch(s, lines, nullptr, lstride, n, nlines, floor_db, compute_floor, threshold_db, gap, min_run,
                         scratch, out, cap, found, floors);
}

extern "C" int atkgpu_detect_line_set(void *s, const float *const *ptrs, size_t n, unsigned nlines,
                                      float threshold_db, int gap, int min_run, atkgpu_chan *out, size_t cap,
                                      unsigned long long *found, float *floors) {
    // held in LDS only: every line of the set is at most LDS_BINS bins
    if (!ptrs || !found || n == 0 || n > (size_t)LDS_BINS || nlines == 0 || (!out && cap)) return -1;
    hipLaunchKernelGGL(k_detect<true>, dim3(nlines), dim3(DT), 0, (hipStream_t)s, nullptr, ptrs, 0, n, 0.0f, 1,
                       threshold_db, gap, min_run, nullptr, out, cap, found, floors);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

extern "C" int atkgpu_median(void *s, const float *x, size_t n, float *out) {
    if (!x || !out || n == 0) return -1;
    hipLaunchKernelGGL(k_median, dim3(1), dim3(DT), 0, (hipStream_t)s, x, n, out);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}
