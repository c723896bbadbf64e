// The CFAR frame test (reference include/atkdsp.h:1733-1800, src/cfar.c:523-590)
// on gfx950: for each frame of a call and each bin, in FP64 as the reference,
//   raw    power from dB, 10^(v/10), values below -300 dB read as -300;
//   X      the mean of the raw power of the last k_frames frames, summed
//          oldest to newest;
//   T      the reference statistic over the bin's N reference cells, taken
//          from a table (left cells nearest first, right nearest first, then
//          any borrowed from the far side): CA their mean in table order; GO /
//          SO the larger / smaller of the two side means (a bin whose window
//          borrowed falls back to CA with CA's alpha); OS the os_k-th
//          smallest; a NaN cell makes T NaN;
//   thr    alpha T, noise T / E[T], hit X > thr.
// One thread per (frame, bin); a frame's raw rows sit after the k-1 rows of
// the frames before it, so a call needs no history kernel.
#include <hip/hip_runtime.h>
#include "atkgpu_dev.h"

#pragma clang fp contract(off)

namespace {

constexpr double LN10_10 = 0.23025850929940458;   // ln(10) / 10

__global__ void k_cfar_raw(const float *__restrict__ db, size_t n, double *__restrict__ raw) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = (double)db[i];
    if (v < -300.0) v = -300.0;
    raw[i] = exp(v * LN10_10);
}

__global__ void k_cfar_test(const double *__restrict__ raw, int nbins, int nframes, int k, int mode, int N,
                            int os_k, double alpha, double alpha_ca, double et, const int *__restrict__ cells,
                            const unsigned char *__restrict__ ok, const unsigned char *__restrict__ sided,
                            double *__restrict__ X, double *__restrict__ thr, double *__restrict__ noise,
                            unsigned char *__restrict__ hits) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= nbins || j >= nframes) return;
    // X of every bin of frame j (rows j .. j+k-1 of raw, oldest first)
    const double *r0 = raw + (size_t)j * nbins;
    auto avg = [&](int b) {
        double s = 0.0;
        for (int r = 0; r < k; ++r) s += r0[(size_t)r * nbins + b];
        return s / k;
    };
    const double x = avg(i);
    const size_t o = (size_t)j * nbins + i;
    X[o] = x;
    if (!ok[i]) {
        thr[o] = __builtin_nan("");
        noise[o] = __builtin_nan("");
        hits[o] = 0;
        return;
    }
    const int *c = cells + (size_t)i * N;
    double z, a = alpha, scale = et;
    if (mode == 1) {
        // the os_k-th smallest: count, for each cell, the cells below it (ties
        // broken by position); a NaN cell makes T NaN
        double v[512];
        bool nan = false;
        for (int q = 0; q < N; ++q) {
            v[q] = avg(c[q]);
            nan = nan || (v[q] != v[q]);
        }
        z = __builtin_nan("");
        if (!nan) {
            for (int q = 0; q < N; ++q) {
                int below = 0;
                for (int p = 0; p < N; ++p) below += (v[p] < v[q]) || (v[p] == v[q] && p < q);
                if (below == os_k - 1) { z = v[q]; break; }
            }
        }
    } else if ((mode == 2 || mode == 3) && sided[i]) {
        const int n = N / 2;
        double L = 0.0, R = 0.0;
        for (int q = 0; q < n; ++q) L += avg(c[q]);
        for (int q = n; q < N; ++q) R += avg(c[q]);
        L /= n;
        R /= n;
        if (L != L || R != R) z = __builtin_nan("");
        else z = mode == 2 ? (L > R ? L : R) : (L < R ? L : R);
    } else {
        double s = 0.0;
        for (int q = 0; q < N; ++q) s += avg(c[q]);
        z = s / N;
        if (mode == 2 || mode == 3) { a = alpha_ca; scale = 1.0; }
    }
    const double t = a * z;
    thr[o] = t;
    noise[o] = z / scale;
    hits[o] = x > t ? 1 : 0;
}

}  // namespace

extern "C" int atkgpu_cfar_raw(void *s, const float *db, size_t n, double *raw) {
    if (!db || !raw || !n) return -1;
    hipLaunchKernelGGL(k_cfar_raw, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)s, db, n, raw);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

extern "C" int atkgpu_cfar_test(void *s, const double *raw, int nbins, int nframes, int k, int mode, int N, int os_k,
                                double alpha, double alpha_ca, double et, const int *cells, const unsigned char *ok,
                                const unsigned char *sided, double *X, double *thr, double *noise,
                                unsigned char *hits) {
    if (!raw || nbins < 1 || nframes < 1 || k < 1 || N < 2 || N > 512 || !cells || !ok || !sided) return -1;
    hipLaunchKernelGGL(k_cfar_test, dim3((unsigned)((nbins + 127) / 128), (unsigned)nframes), dim3(128), 0,
                       (hipStream_t)s, raw, nbins, nframes, k, mode, N, os_k, alpha, alpha_ca, et, cells, ok, sided,
                       X, thr, noise, hits);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}
