// Display helpers (reference src/display.c:4-37) and the NaN-aware sweep
// stitch (src/detect.c:116-134) on gfx950. All three reproduce the reference
// exactly, including NaN handling and which of two equal values is kept.
#include <hip/hip_runtime.h>
#include "atkgpu_dev.h"
#include "dev_common.hpp"

#pragma clang fp contract(off)

// ---- decimate_max -------------------------------------------------------------
// Column c covers [c*per, (c+1)*per) (the last column takes the remainder).
// The reference scans `in[i] > best` from best = in[a]: a NaN in first place
// stays, any other NaN is skipped, and among equal values the first is kept.
// A parallel form with the same result: if in[a] is NaN the answer is in[a];
// else the largest non-NaN value, ties to the lowest index.
__device__ __forceinline__ void dm_pick(float &bv, long long &bi, float v, long long i) {
    // candidate (v, i) beats (bv, bi) when v > bv, or v == bv at a lower index
    if (v != v) return;
    if (bv != bv || v > bv || (v == bv && i < bi)) { bv = v; bi = i; }
}

// one thread a column (short columns)
__global__ void k_decimate_thread(const float *__restrict__ in, size_t n, float *__restrict__ out,
                                  size_t m) {
    const size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (c >= m) return;
    if (m >= n) {
        out[c] = in[c < n ? c : n - 1];
        return;
    }
    const size_t per = n / m;
    const size_t a = c * per, b = (c + 1 == m) ? n : a + per;
    float best = in[a];
    for (size_t i = a + 1; i < b; ++i)
        if (in[i] > best) best = in[i];
    out[c] = best;
}

// one wavefront a column (long columns): lane-strided scan, then a shuffle
// reduction with the (value, first index) rule
__global__ void __launch_bounds__(256)
k_decimate_wave(const float *__restrict__ in, size_t n, float *__restrict__ out, size_t m) {
    const size_t c = blockIdx.x * 4 + threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    if (c >= m) return;
    const size_t per = n / m;
    const size_t a = c * per, b = (c + 1 == m) ? n : a + per;
    const float first = in[a];
    float bv = __builtin_nanf("");
    long long bi = (long long)b;
    for (size_t i = a + lane; i < b; i += 64) dm_pick(bv, bi, in[i], (long long)i);
    for (int off = 32; off > 0; off >>= 1) {
        const float ov = __shfl_xor(bv, off, 64);
        const long long oi = __shfl_xor(bi, off, 64);
        dm_pick(bv, bi, ov, oi);
    }
    if (lane == 0) out[c] = (first != first) ? first : bv;
}

// ---- db_to_pixels -------------------------------------------------------------
__global__ void k_db_to_pixels(const float *__restrict__ db, size_t n, float lo, float hi,
                               const uint32_t *__restrict__ lut, uint32_t *__restrict__ out) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float span = hi - lo;
    const float k = span > 0.0f ? 255.0f / span : 0.0f;
    const float v = db[i];
    if (v != v) { out[i] = lut[0]; return; }
    float t = (v - lo) * k;
    if (t < 0.0f) t = 0.0f;
    if (t > 255.0f) t = 255.0f;
    out[i] = lut[(int)(t + 0.5f)];
}

// ---- stitch_max ------------------------------------------------------------------
// Segment bins map to output bins monotonically, so the bins that land on one
// output bin are a contiguous run. The first bin of each run applies the
// whole run in ascending order: the reference's sequential result, without
// atomics and with no two threads writing one output.
__device__ __forceinline__ bool st_map(size_t i, double seg_lo, double seg_step, double keep_lo,
                                       double keep_hi, double out_lo, double out_hi, double out_step,
                                       size_t nout, long long *k) {
    const double f = seg_lo + ((double)i + 0.5) * seg_step;
    if (f < keep_lo || f > keep_hi) return false;
    if (f < out_lo || f > out_hi) return false;
    long long kk = (long long)((f - out_lo) / out_step);
    if (kk < 0) kk = 0;
    if (kk >= (long long)nout) kk = (long long)nout - 1;
    *k = kk;
    return true;
}

__global__ void k_stitch(const float *__restrict__ seg, size_t nseg, double seg_lo, double seg_step,
                         double keep_lo, double keep_hi, double out_lo, double out_step,
                         float *__restrict__ out, size_t nout) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= nseg) return;
    const double out_hi = out_lo + out_step * (double)nout;
    long long k;
    if (!st_map(i, seg_lo, seg_step, keep_lo, keep_hi, out_lo, out_hi, out_step, nout, &k)) return;
    long long kp;
    if (i > 0 && st_map(i - 1, seg_lo, seg_step, keep_lo, keep_hi, out_lo, out_hi, out_step, nout, &kp) &&
        kp == k)
        return;                                   // not the head of its run
    float cur = out[k];
    for (size_t j = i; j < nseg; ++j) {
        long long kj;
        if (!st_map(j, seg_lo, seg_step, keep_lo, keep_hi, out_lo, out_hi, out_step, nout, &kj) || kj != k)
            break;
        const float v = seg[j];
        if (cur != cur || v > cur) cur = v;
    }
    out[k] = cur;
}

extern "C" int atkgpu_decimate_max(void *s, const float *in, size_t n, float *out, size_t m) {
    if (!n || !m) return 0;
    hipStream_t st = (hipStream_t)s;
    if (m >= n || n / m <= 64) {
        hipLaunchKernelGGL(k_decimate_thread, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, in, n,
                           out, m);
    } else {
        hipLaunchKernelGGL(k_decimate_wave, dim3((unsigned)((m + 3) / 4)), dim3(256), 0, st, in, n, out,
                           m);
    }
    return LAUNCH_OK();
}

extern "C" int atkgpu_db_to_pixels(void *s, const float *db, size_t n, float lo, float hi,
                                   const uint32_t *lut, uint32_t *out) {
    if (!n) return 0;
    hipLaunchKernelGGL(k_db_to_pixels, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)s,
                       db, n, lo, hi, lut, out);
    return LAUNCH_OK();
}

extern "C" int atkgpu_stitch_max(void *s, const float *seg, size_t nseg, double seg_lo,
                                 double seg_step, double keep_lo, double keep_hi, double out_lo,
                                 double out_step, float *out, size_t nout) {
    if (!nseg || !nout) return 0;
    hipLaunchKernelGGL(k_stitch, dim3((unsigned)((nseg + 255) / 256)), dim3(256), 0, (hipStream_t)s,
                       seg, nseg, seg_lo, seg_step, keep_lo, keep_hi, out_lo, out_step, out, nout);
    return LAUNCH_OK();
}
