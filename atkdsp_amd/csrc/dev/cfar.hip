// The CFAR frame test (reference include/atkdsp.h:1733-1800, src/cfar.c:523-590)
// on gfx950: for each frame of a call and each bin, in FP64 as the reference,
//   raw    power from dB, 10^(v/10), values below -300 dB read as -300;
//   X      the mean of the raw power of the last k_frames frames, summed
//          oldest to newest (with one frame, X is the raw row itself);
//   T      the reference statistic over the bin's N reference cells, taken
//          from a table (left cells nearest first, right nearest first, then
//          any borrowed from the far side): CA their mean in table order; GO /
//          SO the larger / smaller of the two side means (a bin whose window
//          borrowed falls back to CA with CA's alpha); OS the os_k-th
//          smallest; a NaN cell makes T NaN;
//   thr    alpha T, noise T / E[T], hit X > thr.
// One thread per (frame, bin). A frame's raw rows sit after the k-1 rows of
// the frames before it, so a call needs no history kernel. A block first
// stages its 128 bins' averages and n_guard + n_ref on either side in LDS
// (one coalesced read); a bin whose table is the plain window (flag `reg`:
// left cells i-g-1, i-g-2, ..., right cells i+g+1, ..., none masked or
// borrowed) takes its cells from there in table order, with no table reads;
// other bins gather through the table. OS with up to 64
// cells sorts them in registers (a bitonic network over the next power of two,
// padded with +inf, which cannot change the k-th smallest of the N); above
// 64 it ranks them by counting in per-lane scratch (each cell's rank = the
// cells below it, ties by position). Both give the os_k-th smallest value.
#include <hip/hip_runtime.h>
#include "atkgpu_dev.h"

#pragma clang fp contract(off)

namespace {

constexpr double LN10_10 = 0.23025850929940458;   // ln(10) / 10
constexpr int TB = 128;                            // threads (bins) per block
constexpr int OS_REG_N = 64;
constexpr int TILE_R = 512;                        // n_guard + n_ref at most

__global__ void k_cfar_raw(const float *__restrict__ db, size_t n, double *__restrict__ raw) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = (double)db[i];
    if (v < -300.0) v = -300.0;
    raw[i] = exp(v * LN10_10);
}

// X of frame j = rows j .. j+k-1 of raw, oldest first
__global__ void k_cfar_avg(const double *__restrict__ raw, int nbins, int nframes, int k, double *__restrict__ X) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= nbins || j >= nframes) return;
    const double *r0 = raw + (size_t)j * nbins + i;
    double s = 0.0;
    for (int r = 0; r < k; ++r) s += r0[(size_t)r * nbins];
    X[(size_t)j * nbins + i] = s / k;
}

// the os_k-th smallest (1-based) of N <= P values held in registers: a
// bitonic sort of the P registers, every index a compile-time constant
template <int P, typename Cell>
__device__ __forceinline__ double os_select_reg(Cell cell, int N, int os_k) {
    double v[P];
    bool nan = false;
#pragma unroll
    for (int q = 0; q < P; ++q) {
        v[q] = q < N ? cell(q < N ? q : 0) : __builtin_inf();
        nan = nan || (v[q] != v[q]);
    }
    if (nan) return __builtin_nan("");
#pragma unroll
    for (int k = 2; k <= P; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (int i = 0; i < P; ++i) {
                const int l = i ^ j;
                if (l > i) {
                    const double a = v[i], b = v[l];
                    const double lo = a < b ? a : b, hi = a < b ? b : a;
                    const bool up = (i & k) == 0;
                    v[i] = up ? lo : hi;
                    v[l] = up ? hi : lo;
                }
            }
    double z = v[0];
#pragma unroll
    for (int q = 1; q < P; ++q) z = q == os_k - 1 ? v[q] : z;
    return z;
}

// the statistic T of one bin over its N cells (cell(q), q in table order) and
// the factor and scale it takes: CA the mean; GO / SO the larger / smaller side
// mean, CA with CA's alpha and unit scale where a side borrowed; OS the os_k-th
// smallest (OSF >= 4: sorted in OSF registers; 2: ranked in scratch)
template <int OSF, typename Cell>
__device__ __forceinline__ void cfar_stat(Cell cell, int mode, int N, int os_k, bool sided, double alpha_ca,
                                          double &z, double &a, double &scale) {
    const int nr = N / 2;
    if constexpr (OSF >= 4) {
        z = os_select_reg<OSF>(cell, N, os_k);
    } else if constexpr (OSF == 2) {
        double v[512];
        bool nan = false;
        for (int q = 0; q < N; ++q) {
            const double x = cell(q);
            v[q] = x;
            nan = nan || (x != x);
        }
        z = __builtin_nan("");
        if (!nan) {
            for (int q = 0; q < N; ++q) {
                const double vq = v[q];
                int below = 0;
                for (int p = 0; p < N; ++p) {
                    const double vp = v[p];
                    below += (vp < vq) || (vp == vq && p < q);
                }
                if (below == os_k - 1) { z = vq; break; }
            }
        }
    } else if ((mode == 2 || mode == 3) && sided) {
        double L = 0.0, R = 0.0;
        for (int q = 0; q < nr; ++q) L += cell(q);
        for (int q = nr; q < N; ++q) R += cell(q);
        L /= nr;
        R /= nr;
        if (L != L || R != R) z = __builtin_nan("");
        else z = mode == 2 ? (L > R ? L : R) : (L < R ? L : R);
    } else {
        double s = 0.0;
        for (int q = 0; q < N; ++q) s += cell(q);
        z = s / N;
        if (mode == 2 || mode == 3) { a = alpha_ca; scale = 1.0; }
    }
}

// OSF: 0 CA / GO / SO only; 4..64 OS with its cells sorted in OSF registers;
// 2 OS ranked in scratch
template <int OSF>
__global__ __launch_bounds__(TB) void k_cfar_test(const double *__restrict__ Xall, int nbins, int nframes, int mode,
                                                  int N, int os_k, int n_guard, double alpha, double alpha_ca,
                                                  double et, const int *__restrict__ cells,
                                                  const unsigned char *__restrict__ ok,
                                                  const unsigned char *__restrict__ sided,
                                                  const unsigned char *__restrict__ reg, double *__restrict__ thr,
                                                  double *__restrict__ noise, unsigned char *__restrict__ hits) {
    __shared__ double tile[TB + 2 * TILE_R];
    const int i = blockIdx.x * TB + threadIdx.x;
    const int j = blockIdx.y;
    const int nr = N / 2, R = n_guard + nr;                 // R <= TILE_R (checked by the launcher)
    const double *X = Xall + (size_t)j * nbins;
    const int t0 = blockIdx.x * TB - R;                     // tile[t] = X[t0 + t]
    for (int t = threadIdx.x; t < TB + 2 * R; t += TB) {
        const int g = t0 + t;
        tile[t] = (g >= 0 && g < nbins) ? X[g] : 0.0;
    }
    __syncthreads();
    if (i >= nbins || j >= nframes) return;
    const size_t o = (size_t)j * nbins + i;
    if (!ok[i]) {
        thr[o] = __builtin_nan("");
        noise[o] = __builtin_nan("");
        hits[o] = 0;
        return;
    }
    const int *c = cells + (size_t)i * N;
    const int li = i - t0;                                  // this bin in the tile
    double z, a = alpha, scale = et;
    if (reg[i]) {
        // the plain window from the tile: cell q < nr at i-g-1-q, q >= nr at
        // i+g+1+(q-nr). The empty asm statements keep this branch apart from
        // the table branch, so that its reads stay LDS reads.
        asm volatile("" ::: "memory");
        cfar_stat<OSF>(
            [&](int q) -> double { return tile[q < nr ? li - n_guard - 1 - q : li + n_guard + 1 + (q - nr)]; },
            mode, N, os_k, sided[i] != 0, alpha_ca, z, a, scale);
        asm volatile("" ::: "memory");
    } else {
        cfar_stat<OSF>([&](int q) -> double { return X[c[q]]; }, mode, N, os_k, sided[i] != 0, alpha_ca, z, a,
                       scale);
    }
    const double t = a * z;
    thr[o] = t;
    noise[o] = z / scale;
    hits[o] = tile[li] > t ? 1 : 0;
}

// out[2q], out[2q+1] = X[idx[q]], noise[idx[q]]: the values the tracker reads,
// at the hit bins only (idx: frame * nbins + bin)
__global__ void k_cfar_gather(const double *__restrict__ X, const double *__restrict__ noise,
                              const unsigned *__restrict__ idx, size_t n, double *__restrict__ out) {
    const size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (q >= n) return;
    const size_t i = idx[q];
    out[2 * q] = X[i];
    out[2 * q + 1] = noise[i];
}

}  // namespace

extern "C" int atkgpu_cfar_gather(void *s, const double *X, const double *noise, const unsigned *idx, size_t n,
                                  double *out) {
    if (!X || !noise || !idx || !out) return -1;
    if (!n) return 0;
    hipLaunchKernelGGL(k_cfar_gather, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)s, X, noise, idx,
                       n, out);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

extern "C" int atkgpu_cfar_raw(void *s, const float *db, size_t n, double *raw) {
    if (!db || !raw || !n) return -1;
    hipLaunchKernelGGL(k_cfar_raw, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)s, db, n, raw);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

extern "C" int atkgpu_cfar_test(void *s, const double *raw, int nbins, int nframes, int k, int mode, int N, int os_k,
                                int n_guard, double alpha, double alpha_ca, double et, const int *cells,
                                const unsigned char *ok, const unsigned char *sided, const unsigned char *reg,
                                double *X, double *thr, double *noise, unsigned char *hits) {
    if (!raw || nbins < 1 || nframes < 1 || nframes > 65535 || k < 1 || N < 2 || N > 512 || (N & 1) || !cells ||
        !ok || !sided || !reg || !thr || !noise || !hits || (k > 1 && !X) || (mode == 1 && (os_k < 1 || os_k > N)) ||
        n_guard < 0 || n_guard + N / 2 > TILE_R)
        return -1;
    const dim3 grid((unsigned)((nbins + TB - 1) / TB), (unsigned)nframes);
    const double *Xin = raw;
    if (k > 1) {
        hipLaunchKernelGGL(k_cfar_avg, grid, dim3(TB), 0, (hipStream_t)s, raw, nbins, nframes, k, X);
        Xin = X;
    }
    auto kern = mode != 1 ? k_cfar_test<0>
                : N <= 4  ? k_cfar_test<4>
                : N <= 8  ? k_cfar_test<8>
                : N <= 16 ? k_cfar_test<16>
                : N <= 32 ? k_cfar_test<32>
                : N <= OS_REG_N ? k_cfar_test<64>
                : k_cfar_test<2>;
    hipLaunchKernelGGL(kern, grid, dim3(TB), 0, (hipStream_t)s, Xin, nbins, nframes, mode, N, os_k, n_guard, alpha,
                       alpha_ca, et, cells, ok, sided, reg, thr, noise, hits);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}
