// Spectrum lines from batched frame transforms (reference src/fft.c:155-431):
// fftshift k = (i + n/2) mod n, linear power |X|^2/n^2, per-bin detector
// across frames, then dB. A launch's frames are split into up to 32
// contiguous groups, one block row each (256 bins a block), so that a 4096-bin
// line keeps about 512 blocks busy; each group's partial line is then merged
// into the running line in group order. Max and min are exact in any order;
// the average's sum is grouped (within the display tolerance of the
// reference's frame-order float sum). Contraction is off so each formula
// rounds as it reads.
#include <hip/hip_runtime.h>
#include <math.h>
#include "atkgpu_dev.h"
#include "dev_common.hpp"

#pragma clang fp contract(off)

struct c2 { float x, y; };

// frame groups of a launch: at most 32, at most 2^17 / n (so the partials
// fit atkgpu_spec_part_entries(n)), every group holding a frame
static inline void spec_groups(size_t n, size_t frames, int *ng, size_t *fpg) {
    size_t g = ((size_t)1 << 17) / n;
    if (g < 1) g = 1;
    if (g > 32) g = 32;
    if (g > frames) g = frames;
    const size_t f = (frames + g - 1) / g;
    *fpg = f;
    *ng = (int)((frames + f - 1) / f);
}

// doubles of partials a launch may write: 4 planes (the stats' max, min,
// sum, sum of squares) of at most 32 groups, at most max(n, 2^17) entries each
extern "C" size_t atkgpu_spec_part_entries(size_t n) {
    const size_t a = (size_t)32 * n, b = n > ((size_t)1 << 17) ? n : ((size_t)1 << 17);
    return 4 * (a < b ? a : b);
}

__device__ __forceinline__ float spec_merge(int det, float acc, float pw) {
    if (det == 0) return pw > acc ? pw : acc;
    if (det == 1) return pw < acc ? pw : acc;
    return acc + pw;
}

// reduce, step 1: group g's frames [g*fpg, min(frames, (g+1)*fpg)) of bin i,
// float power and float detector in frame order (src/fft.c:177-191)
__global__ void __launch_bounds__(256) k_spec_part(const c2 *__restrict__ X, size_t n, size_t frames, size_t fpg,
                                                   int det, float *__restrict__ part) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t k = (i + n / 2) % n;
    const float inv2 = 1.0f / ((float)n * (float)n);
    const size_t f0 = blockIdx.y * fpg;
    const size_t f1 = f0 + fpg < frames ? f0 + fpg : frames;
    float acc = 0.0f;
    for (size_t f = f0; f < f1; ++f) {
        const c2 v = X[f * n + k];
        const float pw = (v.x * v.x + v.y * v.y) * inv2;
        acc = f == f0 ? pw : spec_merge(det, acc, pw);
    }
    part[blockIdx.y * n + i] = acc;
}

// reduce, step 2: the running line (unless first) and the groups in order
__global__ void __launch_bounds__(256) k_spec_merge(const float *__restrict__ part, size_t n, int ng, int det,
                                                    int first, float *__restrict__ line) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    float acc = first ? part[i] : spec_merge(det, line[i], part[i]);
    for (int g = 1; g < ng; ++g) acc = spec_merge(det, acc, part[(size_t)g * n + i]);
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

// stats: FP64 power; max/min kept as float across launches (rounding is
// monotonic, so the float of the FP64 max equals the max of the rounded
// values), sums in FP64 (src/fft.c:317-331). Step 1: each frame group's FP64
// max, min, sum and sum of squares into part[4][ng][n].
__global__ void __launch_bounds__(256) k_spec_stats_part(const c2 *__restrict__ X, size_t n, size_t frames,
                                                         size_t fpg, int ng, double *__restrict__ part) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t k = (i + n / 2) % n;
    const double inv2 = 1.0 / ((double)n * (double)n);
    const size_t f0 = blockIdx.y * fpg;
    const size_t f1 = f0 + fpg < frames ? f0 + fpg : frames;
    double amx = 0.0, amn = 0.0, asu = 0.0, as2 = 0.0;
    for (size_t f = f0; f < f1; ++f) {
        const c2 v = X[f * n + k];
        const double re = v.x, im = v.y;
        const double pw = (re * re + im * im) * inv2;
        if (f == f0) { amx = amn = asu = pw; as2 = pw * pw; continue; }
        if (pw > amx) amx = pw;
        if (pw < amn) amn = pw;
        asu += pw;
        as2 += pw * pw;
    }
    const size_t plane = (size_t)ng * n, at = blockIdx.y * n + i;
    part[at] = amx;
    part[plane + at] = amn;
    part[2 * plane + at] = asu;
    part[3 * plane + at] = as2;
}

// step 2: the running planes (unless first) and the groups in order; the
// extremes are stored as their float rounding at the end of each launch
__global__ void __launch_bounds__(256) k_spec_stats_merge(const double *__restrict__ part, size_t n, int ng,
                                                          int first, float *__restrict__ mx, float *__restrict__ mn,
                                                          double *__restrict__ su, double *__restrict__ s2) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t plane = (size_t)ng * n;
    double amx = part[i], amn = part[plane + i], asu = part[2 * plane + i], as2 = part[3 * plane + i];
    if (!first) {
        const double pmx = mx[i], pmn = mn[i];
        amx = pmx > amx ? pmx : amx;
        amn = pmn < amn ? pmn : amn;
        asu = su[i] + asu;
        as2 = s2[i] + as2;
    }
    for (int g = 1; g < ng; ++g) {
        const size_t at = (size_t)g * n + i;
        if (part[at] > amx) amx = part[at];
        if (part[plane + at] < amn) amn = part[plane + at];
        asu += part[2 * plane + at];
        as2 += part[3 * plane + at];
    }
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
                                 int first, float *line, void *part) {
    if (!frames) return 0;
    int ng;
    size_t fpg;
    spec_groups(n, frames, &ng, &fpg);
    const unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_spec_part, dim3(nb, (unsigned)ng), dim3(256), 0, (hipStream_t)s, (const c2 *)X, n, frames,
                       fpg, detector, (float *)part);
    if (LAUNCH_OK()) return -4;
    hipLaunchKernelGGL(k_spec_merge, dim3(nb), dim3(256), 0, (hipStream_t)s, (const float *)part, n, ng, detector,
                       first, line);
    return LAUNCH_OK();
}

extern "C" int atkgpu_spec_reduce_db(void *s, const float *line, size_t n, int detector,
                                     long long frames, float *out_db) {
    hipLaunchKernelGGL(k_spec_reduce_db, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)s, line, n, detector, frames, out_db);
    return LAUNCH_OK();
}

extern "C" int atkgpu_spec_stats_accum(void *s, const void *X, size_t n, size_t frames, int first,
                                       float *mx, float *mn, double *su, double *s2, void *part) {
    if (!frames) return 0;
    int ng;
    size_t fpg;
    spec_groups(n, frames, &ng, &fpg);
    const unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_spec_stats_part, dim3(nb, (unsigned)ng), dim3(256), 0, (hipStream_t)s, (const c2 *)X, n,
                       frames, fpg, ng, (double *)part);
    if (LAUNCH_OK()) return -4;
    hipLaunchKernelGGL(k_spec_stats_merge, dim3(nb), dim3(256), 0, (hipStream_t)s, (const double *)part, n, ng,
                       first, mx, mn, su, s2);
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
