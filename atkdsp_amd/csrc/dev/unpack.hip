// Raw receiver samples -> cf32 (reference src/convert.c:33-111) and the
// one-pole DC block (src/convert.c:7-17) on gfx950.
//
// The conversion is one subtract and one multiply per component with the
// constants the host formed exactly as the reference does
// (c' = c + off/k in float, src/convert.c:47, :57, :68), so every output is
// bit-identical to the reference. Contraction is off in this file: nothing
// here may become an FMA.
#include <hip/hip_runtime.h>
#include <atomic>
#include <mutex>
#include "atkgpu_dev.h"
#include "dev_common.hpp"
#include "atkgpu.h"

#pragma clang fp contract(off)

#define UNPACK_BLOCK 256

struct cf { float re, im; };

template <int FMT>
__device__ __forceinline__ cf load_conv(const void *raw, size_t i, float cre, float cim, float k,
                                        float ore, float oim) {
    cf o;
    if constexpr (FMT == FMT_CU8) {
        const uchar2 b = ((const uchar2 *)raw)[i];
        o.re = ((float)b.x - cre) * k;
        o.im = ((float)b.y - cim) * k;
    } else if constexpr (FMT == FMT_CI8) {
        const char2 b = ((const char2 *)raw)[i];
        o.re = ((float)(signed char)b.x - cre) * k;
        o.im = ((float)(signed char)b.y - cim) * k;
    } else if constexpr (FMT == FMT_CI16 || FMT == FMT_CI16Q11) {
        const short2 b = ((const short2 *)raw)[i];
        o.re = ((float)b.x - cre) * k;
        o.im = ((float)b.y - cim) * k;
    } else {
        const float2 b = ((const float2 *)raw)[i];
        o.re = b.x - ore;
        o.im = b.y - oim;
    }
    return o;
}

// grid = ATKGPU_MEAN_PARTS blocks; each block writes its FP64 partial sums
template <int FMT>
__global__ void __launch_bounds__(UNPACK_BLOCK)
k_unpack(const void *__restrict__ raw, size_t n, float cre, float cim, float k, float ore, float oim,
         int copy_cf32, cf *__restrict__ out, double *__restrict__ partials) {
    double sr = 0.0, si = 0.0;
    const size_t stride = (size_t)gridDim.x * UNPACK_BLOCK;
    for (size_t i = blockIdx.x * (size_t)UNPACK_BLOCK + threadIdx.x; i < n; i += stride) {
        cf o;
        if (FMT == FMT_CF32 && copy_cf32) {
            const cf b = ((const cf *)raw)[i];   // off == 0: a copy, as the reference's memcpy
            o = b;
        } else {
            o = load_conv<FMT>(raw, i, cre, cim, k, ore, oim);
        }
        out[i] = o;
        sr += (double)o.re;
        si += (double)o.im;
    }
    if (!partials) return;
    __shared__ double shr[UNPACK_BLOCK], shi[UNPACK_BLOCK];
    shr[threadIdx.x] = sr;
    shi[threadIdx.x] = si;
    __syncthreads();
    for (int w = UNPACK_BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            shr[threadIdx.x] += shr[threadIdx.x + w];
            shi[threadIdx.x] += shi[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = shr[0];
        partials[2 * blockIdx.x + 1] = shi[0];
    }
}

// one block: fixed-order tree over the partials into the FP64 sums of out
__global__ void __launch_bounds__(256)
k_mean_final(const double *__restrict__ partials, int nparts, double *__restrict__ sum2) {
    __shared__ double shr[256], shi[256];
    double sr = 0.0, si = 0.0;
    for (int i = threadIdx.x; i < nparts; i += 256) {
        sr += partials[2 * i];
        si += partials[2 * i + 1];
    }
    shr[threadIdx.x] = sr;
    shi[threadIdx.x] = si;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            shr[threadIdx.x] += shr[threadIdx.x + w];
            shi[threadIdx.x] += shi[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sum2[0] = shr[0];
        sum2[1] = shi[0];
    }
}

extern "C" int atkgpu_unpack(void *s, const void *raw, int fmt, size_t n, float cre, float cim,
                             float k, float off_re, float off_im, void *out, double *partials,
                             double *sum2) {
    if (!n) return 0;
    hipStream_t st = (hipStream_t)s;
    const unsigned grid = grid_for(n, UNPACK_BLOCK, ATKGPU_MEAN_PARTS);
    const int copy = (off_re == 0.0f && off_im == 0.0f) ? 1 : 0;
    double *parts = sum2 ? partials : nullptr;
    switch (fmt) {
    case FMT_CU8:
        hipLaunchKernelGGL(k_unpack<FMT_CU8>, dim3(grid), dim3(UNPACK_BLOCK), 0, st, raw, n, cre, cim,
                           k, off_re, off_im, 0, (cf *)out, parts);
        break;
    case FMT_CI8:
        hipLaunchKernelGGL(k_unpack<FMT_CI8>, dim3(grid), dim3(UNPACK_BLOCK), 0, st, raw, n, cre, cim,
                           k, off_re, off_im, 0, (cf *)out, parts);
        break;
    case FMT_CI16:
    case FMT_CI16Q11:
        hipLaunchKernelGGL(k_unpack<FMT_CI16>, dim3(grid), dim3(UNPACK_BLOCK), 0, st, raw, n, cre,
                           cim, k, off_re, off_im, 0, (cf *)out, parts);
        break;
    case FMT_CF32:
        hipLaunchKernelGGL(k_unpack<FMT_CF32>, dim3(grid), dim3(UNPACK_BLOCK), 0, st, raw, n, cre,
                           cim, k, off_re, off_im, copy, (cf *)out, parts);
        break;
    default:
        return -1;
    }
    if (LAUNCH_OK()) return -4;
    if (sum2) {
        hipLaunchKernelGGL(k_mean_final, dim3(1), dim3(256), 0, st, partials, (int)grid, sum2);
        if (LAUNCH_OK()) return -4;
    }
    return 0;
}

// ---- one-pole DC block as a chunked scan --------------------------------------
// The recurrence st <- st + a(x - st) is affine in st: over a chunk of len
// samples, st_out = (1-a)^len st_in + f, where f is the chunk run from zero.
// Pass 1 forms f per chunk in FP64; pass 2 chains the chunks in FP64 (one
// block); pass 3 re-runs every chunk in FP32 with the reference's own
// operation order from its FP64 start state. Output differs from the serial
// FP32 pole only through the start states, by far less than 2e-6.
#define POLE_CHUNK 256

__global__ void k_pole_chunks(const cf *__restrict__ x, size_t n, float alpha, double *__restrict__ f) {
    const size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const size_t nch = (n + POLE_CHUNK - 1) / POLE_CHUNK;
    if (c >= nch) return;
    const size_t a = c * POLE_CHUNK, b = a + POLE_CHUNK < n ? a + POLE_CHUNK : n;
    const double al = (double)alpha;
    double sr = 0.0, si = 0.0;
    for (size_t i = a; i < b; ++i) {
        sr += al * ((double)x[i].re - sr);
        si += al * ((double)x[i].im - si);
    }
    f[2 * c] = sr;
    f[2 * c + 1] = si;
}

// one block of 1024: each thread owns a run of chunks; a serial pass over the
// 1024 run summaries gives each run its start state. start[c] = state entering chunk c.
__global__ void __launch_bounds__(1024)
k_pole_scan(const double *__restrict__ f, size_t n, float alpha, const float *__restrict__ state,
            double *__restrict__ start) {
    __shared__ double D[1024], Fr[1024], Fi[1024];
    const size_t nch = (n + POLE_CHUNK - 1) / POLE_CHUNK;
    const size_t per = (nch + 1023) / 1024;
    const size_t c0 = threadIdx.x * per, c1 = c0 + per < nch ? c0 + per : nch;
    const double one_m = 1.0 - (double)alpha;
    double d = 1.0, fr = 0.0, fi = 0.0;
    for (size_t c = c0; c < c1; ++c) {
        const size_t len = (c + 1) * POLE_CHUNK <= n ? POLE_CHUNK : n - c * POLE_CHUNK;
        const double dc = pow(one_m, (double)len);
        d *= dc;
        fr = dc * fr + f[2 * c];
        fi = dc * fi + f[2 * c + 1];
    }
    D[threadIdx.x] = d;
    Fr[threadIdx.x] = fr;
    Fi[threadIdx.x] = fi;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sr = (double)state[0], si = (double)state[1];
        for (int t = 0; t < 1024; ++t) {
            const double nr = D[t] * sr + Fr[t], ni = D[t] * si + Fi[t];
            D[t] = sr;     // the run's start state
            Fr[t] = si;
            sr = nr;
            si = ni;
        }
    }
    __syncthreads();
    double sr = D[threadIdx.x], si = Fr[threadIdx.x];
    for (size_t c = c0; c < c1; ++c) {
        start[2 * c] = sr;
        start[2 * c + 1] = si;
        const size_t len = (c + 1) * POLE_CHUNK <= n ? POLE_CHUNK : n - c * POLE_CHUNK;
        const double dc = pow(one_m, (double)len);
        sr = dc * sr + f[2 * c];
        si = dc * si + f[2 * c + 1];
    }
}

__global__ void k_pole_apply(cf *__restrict__ x, size_t n, float alpha, const double *__restrict__ start,
                             float *__restrict__ state) {
    const size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const size_t nch = (n + POLE_CHUNK - 1) / POLE_CHUNK;
    if (c >= nch) return;
    const size_t a = c * POLE_CHUNK, b = a + POLE_CHUNK < n ? a + POLE_CHUNK : n;
    float sr = (float)start[2 * c], si = (float)start[2 * c + 1];
    for (size_t i = a; i < b; ++i) {
        cf v = x[i];
        sr += alpha * (v.re - sr);
        si += alpha * (v.im - si);
        v.re -= sr;
        v.im -= si;
        x[i] = v;
    }
    if (c == nch - 1) {
        state[0] = sr;
        state[1] = si;
    }
}

extern "C" int atkgpu_dc_pole(void *s, void *x, size_t n, float alpha, float *state, double *scratch) {
    if (!n) return 0;
    hipStream_t st = (hipStream_t)s;
    const size_t nch = (n + POLE_CHUNK - 1) / POLE_CHUNK;
    double *f = scratch, *start = scratch + 2 * nch;
    const unsigned g = (unsigned)((nch + 255) / 256);
    hipLaunchKernelGGL(k_pole_chunks, dim3(g), dim3(256), 0, st, (const cf *)x, n, alpha, f);
    hipLaunchKernelGGL(k_pole_scan, dim3(1), dim3(1024), 0, st, (const double *)f, n, alpha,
                       (const float *)state, start);
    hipLaunchKernelGGL(k_pole_apply, dim3(g), dim3(256), 0, st, (cf *)x, n, alpha,
                       (const double *)start, state);
    return LAUNCH_OK();
}

// ---- unpack with the block mean into device buffers, allocating nothing -------------
// The scratch of the mean: the FP64 partials of one pass's blocks (reused pass
// after pass, in stream order) and the FP64 sums of each pass (a pass is
// ATKGPU_MEAN_PASS samples, as in atkdsp_unpack_dc). One copy per GPU,
// resolved to its device address once per device and held until the process
// ends. Every slot a call reads, the same call wrote first.
#define NOALLOC_PASSES 512
static_assert((size_t)NOALLOC_PASSES * ATKGPU_MEAN_PASS == ATKGPU_NOALLOC_MAX_SAMPLES,
              "the header's maximum is the scratch's pass count times the pass length");
static __device__ __align__(256) double scratch[2 * ATKGPU_MEAN_PARTS + 2 * NOALLOC_PASSES];

#define NOALLOC_MAX_DEV 64
static std::atomic<double *> g_scratch_at[NOALLOC_MAX_DEV];
static std::mutex g_scratch_mu;

// the scratch's device address on the current device (resolved on first use
// there); nullptr when it cannot be had
static double *scratch_here() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEV) return nullptr;
    double *p = g_scratch_at[dev].load(std::memory_order_acquire);
    if (p) return p;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    p = g_scratch_at[dev].load(std::memory_order_relaxed);
    if (!p) {
        void *q = nullptr;
        if (hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)) != hipSuccess || !q) {
            (void)hipGetLastError();        // leave no error for the next launch check
            return nullptr;
        }
        p = (double *)q;
        g_scratch_at[dev].store(p, std::memory_order_release);
    }
    return p;
}

// called by the library's init on the current device
extern "C" int atkgpu_noalloc_setup(void) { return scratch_here() ? 0 : -4; }

// the mean from the pass sums, added in pass order from zero, then scaled and
// offset as the pooled entry point does on the host (no contraction here)
__global__ void k_mean_from_passes(const double *__restrict__ sums, int passes, unsigned long long n, float ore,
                                   float oim, double *__restrict__ mean) {
    if (threadIdx.x != 0) return;
    if (n == 0) {
        mean[0] = (double)ore;
        mean[1] = (double)oim;
        return;
    }
    double sr = 0.0, si = 0.0;
    for (int j = 0; j < passes; ++j) {
        sr += sums[2 * j];
        si += sums[2 * j + 1];
    }
    const double inv = 1.0 / (double)n;
    mean[0] = sr * inv + (double)ore;
    mean[1] = si * inv + (double)oim;
}

extern "C" ptrdiff_t atkgpu_unpack_dc_noalloc(void *stream, const void *raw, size_t nbytes, int fmt,
                                              atkdsp_cf32 *out, size_t out_cap, float off_re, float off_im,
                                              double *mean) {
    if (!raw || !out) return ATKDSP_E_ARG;
    size_t bps = 0;
    switch (fmt) {
    case FMT_CU8: case FMT_CI8: bps = 2; break;
    case FMT_CI16: case FMT_CI16Q11: bps = 4; break;
    case FMT_CF32: bps = 8; break;
    default: return ATKDSP_E_ARG;
    }
    size_t n = nbytes / bps;
    if (n > out_cap) n = out_cap;
    if (n > ATKGPU_NOALLOC_MAX_SAMPLES) return ATKDSP_E_CAP;
    double *scr = scratch_here();
    if (!scr) return ATKDSP_E_FFT;
    double *parts = scr, *sums = scr + 2 * ATKGPU_MEAN_PARTS;
    float cre, cim, k;
    atkgpu_unpack_consts(fmt, off_re, off_im, &cre, &cim, &k);
    const size_t passes = (n + ATKGPU_MEAN_PASS - 1) / ATKGPU_MEAN_PASS;
    for (size_t j = 0; j < passes; ++j) {
        const size_t a = j * ATKGPU_MEAN_PASS;
        const size_t c = n - a < ATKGPU_MEAN_PASS ? n - a : ATKGPU_MEAN_PASS;
        if (atkgpu_unpack(stream, (const char *)raw + a * bps, fmt, c, cre, cim, k, off_re, off_im, out + a,
                          parts, mean ? sums + 2 * j : nullptr))
            return ATKDSP_E_FFT;
    }
    if (mean) {
        hipLaunchKernelGGL(k_mean_from_passes, dim3(1), dim3(64), 0, (hipStream_t)stream, sums, (int)passes,
                           (unsigned long long)n, off_re, off_im, mean);
        if (LAUNCH_OK()) return ATKDSP_E_FFT;
    }
    return (ptrdiff_t)n;
}
