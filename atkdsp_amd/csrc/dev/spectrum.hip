// Spectrum lines from batched frame transforms (reference src/fft.c:155-431):
// fftshift k = (i + n/2) mod n, linear power |X|^2/n^2, per-bin detector
// across frames in frame order, then dB. Thread per display bin, frames in
// ascending order, so a line does not depend on how frames are split into
// launches. Contraction is off so each formula rounds as it reads.
#include <hip/hip_runtime.h>
#include <math.h>
#include "atkgpu_dev.h"
#include "dev_common.hpp"

#pragma clang fp contract(off)

struct c2 { float x, y; };

// reduce: float power and float detector line (src/fft.c:177-191)
__global__ void k_spec_accum(const c2 *__restrict__ X, size_t n, size_t frames, int det, int first,
                             float *__restrict__ line) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t k = (i + n / 2) % n;
    const float inv2 = 1.0f / ((float)n * (float)n);
    float acc = first ? 0.0f : line[i];
    for (size_t f = 0; f < frames; ++f) {
        const c2 v = X[f * n + k];
        const float pw = (v.x * v.x + v.y * v.y) * inv2;
        if (first && f == 0) { acc = pw; continue; }
        if (det == 0) { if (pw > acc) acc = pw; }
        else if (det == 1) { if (pw < acc) acc = pw; }
        else acc += pw;
    }
    line[i] = acc;
}

__global__ void k_spec_reduce_db(const float *__restrict__ line, size_t n, int det, long long frames,
                                 float *__restrict__ out) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float scale = det == 2 ? 1.0f / (float)frames : 1.0f;
    const float mag = sqrtf(line[i] * scale);
    out[i] = 20.0f * log10f(mag + 1e-12f);
}

// stats: FP64 power; max/min kept as float (rounding is monotonic, so the
// float of the FP64 max equals the max of the rounded values), sums in FP64
// (src/fft.c:317-331)
__global__ void k_spec_stats_accum(const c2 *__restrict__ X, size_t n, size_t frames, int first,
                                   float *__restrict__ mx, float *__restrict__ mn,
                                   double *__restrict__ su, double *__restrict__ s2) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t k = (i + n / 2) % n;
    const double inv2 = 1.0 / ((double)n * (double)n);
    double amx = 0.0, amn = 0.0, asu = 0.0, as2 = 0.0;
    if (!first) { amx = mx[i]; amn = mn[i]; asu = su[i]; as2 = s2[i]; }
    for (size_t f = 0; f < frames; ++f) {
        const c2 v = X[f * n + k];
        const double re = v.x, im = v.y;
        const double pw = (re * re + im * im) * inv2;
        if (first && f == 0) { amx = amn = asu = pw; as2 = pw * pw; continue; }
        if (pw > amx) amx = pw;
        if (pw < amn) amn = pw;
        asu += pw;
        as2 += pw * pw;
    }
    // keep the exact FP64 extremes across launches by storing their float
    // rounding only at the end of each launch: the max of the rounded values
    // is the rounding of the max
    mx[i] = (float)amx;
    mn[i] = (float)amn;
    su[i] = asu;
    s2[i] = as2;
}

__global__ void k_spec_stats_db(size_t n, long long frames, const float *__restrict__ mx,
                                const float *__restrict__ mn, const double *__restrict__ su,
                                const double *__restrict__ s2, float *max_db, float *avg_db,
                                float *min_db, float *sk) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double M = (double)frames;
    if (max_db) max_db[i] = 10.0f * log10f(mx[i] + 1e-12f);
    if (min_db) min_db[i] = 10.0f * log10f(mn[i] + 1e-12f);
    if (avg_db) avg_db[i] = 10.0f * log10f((float)(su[i] / M) + 1e-12f);
    if (sk) {
        const double den = su[i] * su[i];
        sk[i] = (frames < 2 || den <= 0.0) ? 1.0f : (float)(M * s2[i] / den);
    }
}

// one frame, |X|/n in dB, fftshifted (src/fft.c:155-164)
__global__ void k_power_db(const c2 *__restrict__ X, size_t n, float *__restrict__ out) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t k = (i + n / 2) % n;
    const float inv = 1.0f / (float)n;
    const c2 v = X[k];
    const float mag = sqrtf(v.x * v.x + v.y * v.y) * inv;
    out[i] = 20.0f * log10f(mag + 1e-12f);
}

extern "C" int atkgpu_spec_accum(void *s, const void *X, size_t n, size_t frames, int detector,
                                 int first, float *line) {
    if (!frames) return 0;
    hipLaunchKernelGGL(k_spec_accum, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)s,
                       (const c2 *)X, n, frames, detector, first, line);
    return LAUNCH_OK();
}

extern "C" int atkgpu_spec_reduce_db(void *s, const float *line, size_t n, int detector,
                                     long long frames, float *out_db) {
    hipLaunchKernelGGL(k_spec_reduce_db, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)s, line, n, detector, frames, out_db);
    return LAUNCH_OK();
}

extern "C" int atkgpu_spec_stats_accum(void *s, const void *X, size_t n, size_t frames, int first,
                                       float *mx, float *mn, double *su, double *s2) {
    if (!frames) return 0;
    hipLaunchKernelGGL(k_spec_stats_accum, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)s, (const c2 *)X, n, frames, first, mx, mn, su, s2);
    return LAUNCH_OK();
}

extern "C" int atkgpu_spec_stats_db(void *s, size_t n, long long frames, const float *mx,
                                    const float *mn, const double *su, const double *s2,
                                    float *max_db, float *avg_db, float *min_db, float *sk) {
    hipLaunchKernelGGL(k_spec_stats_db, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)s, n, frames, mx, mn, su, s2, max_db, avg_db, min_db, sk);
    return LAUNCH_OK();
}

extern "C" int atkgpu_power_db(void *s, const void *X, size_t n, float *out_db) {
    hipLaunchKernelGGL(k_power_db, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)s,
                       (const c2 *)X, n, out_db);
    return LAUNCH_OK();
}
