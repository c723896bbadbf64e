// Polyphase filterbank kernels, general shapes (reference src/channelizer.c:218-303),
// plus the W1 covariance and W2 power accumulates (src/channelizer.c:420-471).
//
// The general path is fold -> batched FFT (fft.hip) -> rotate/cal/transpose:
//   fold[q]  = sum_k h[kM+q] * w[kM+q], k ascending, FP32 multiply then add
//              (the reference's operation order, src/channelizer.c:258-268)
//   X        = M-point forward FFT of fold
//   y[k]     = X[k] * exp(-2 pi i k shift/M), shift = ((b+1) hop) mod M with
//              the int64 hop counter b (src/channelizer.c:270-292), times cal
//   out      = channel-major [element][channel][hop] (src/channelizer.c:294-300)
#include <hip/hip_runtime.h>
#include "atkgpu_dev.h"
#include "dev_common.hpp"

#pragma clang fp contract(off)

struct c2 { float x, y; };

#define FOLD_P 8   // hops a thread folds

__global__ void __launch_bounds__(256)
k_pfb_fold(const c2 *__restrict__ sbuf, size_t spitch, size_t m, unsigned M, unsigned hop, unsigned K,
           const float *__restrict__ taps, c2 *__restrict__ fold) {
    const unsigned q = blockIdx.x * 256 + threadIdx.x;
    const size_t i0 = (size_t)blockIdx.y * FOLD_P;
    const unsigned e = blockIdx.z;
    if (q >= M) return;
    const c2 *s = sbuf + (size_t)e * spitch;
    for (int p = 0; p < FOLD_P; ++p) {
        const size_t i = i0 + p;
        if (i >= m) break;
        const c2 *w = s + i * hop + q;
        float ar = taps[q] * w[0].x;
        float ai = taps[q] * w[0].y;
        for (unsigned k = 1; k < K; ++k) {
            const float h = taps[(size_t)k * M + q];
            const c2 x = w[(size_t)k * M];
            ar = ar + h * x.x;
            ai = ai + h * x.y;
        }
        fold[((size_t)e * m + i) * M + q] = {ar, ai};
    }
}

// 32x32 tiles: read X hop-major (coalesced over channels), write out
// channel-major (coalesced over hops)
__global__ void __launch_bounds__(256)
k_pfb_rotate_store(const c2 *__restrict__ X, size_t m, unsigned M, unsigned hop, long long b0,
                   const c2 *__restrict__ tw, const c2 *__restrict__ cal, unsigned cal_nel,
                   c2 *__restrict__ out) {
    __shared__ c2 tile[32][33];
    const unsigned e = blockIdx.z;
    const unsigned k0 = blockIdx.x * 32;
    const size_t i0 = (size_t)blockIdx.y * 32;
    const unsigned tx = threadIdx.x % 32, ty = threadIdx.x / 32;
    for (unsigned r = ty; r < 32; r += 8) {
        const size_t i = i0 + r;
        const unsigned k = k0 + tx;
        if (i < m && k < M) {
            c2 v = X[((size_t)e * m + i) * M + k];
            const long long b = b0 + (long long)i;
            const unsigned shift = (unsigned)((((b + 1) * (long long)hop) % (long long)M + M) % M);
            if (shift) {
                const unsigned idx = (unsigned)(((unsigned long long)k * shift) % M);
                const c2 w = tw[idx];
                v = {v.x * w.x - v.y * w.y, v.x * w.y + v.y * w.x};
            }
            if (cal && e < cal_nel) {
                const c2 c = cal[(size_t)e * M + k];
                v = {v.x * c.x - v.y * c.y, v.x * c.y + v.y * c.x};
            }
            tile[r][tx] = v;
        }
    }
    __syncthreads();
    for (unsigned r = ty; r < 32; r += 8) {
        const unsigned k = k0 + r;
        const size_t i = i0 + tx;
        if (i < m && k < M) out[((size_t)e * M + k) * m + i] = tile[tx][r];
    }
}

extern "C" int atkgpu_pfb_fold(void *s, const void *sbuf, size_t spitch, unsigned nel, size_t m,
                               unsigned M, unsigned hop, unsigned K, const float *taps, void *fold) {
    if (!m) return 0;
    dim3 grid((M + 255) / 256, (unsigned)((m + FOLD_P - 1) / FOLD_P), nel);
    hipLaunchKernelGGL(k_pfb_fold, grid, dim3(256), 0, (hipStream_t)s, (const c2 *)sbuf, spitch, m, M,
                       hop, K, taps, (c2 *)fold);
    return LAUNCH_OK();
}

extern "C" int atkgpu_pfb_rotate_store(void *s, const void *X, unsigned nel, size_t m, unsigned M,
                                       unsigned hop, long long b0, const void *twM, const void *cal,
                                       unsigned cal_nel, void *out) {
    if (!m) return 0;
    dim3 grid((M + 31) / 32, (unsigned)((m + 31) / 32), nel);
    hipLaunchKernelGGL(k_pfb_rotate_store, grid, dim3(256), 0, (hipStream_t)s, (const c2 *)X, m, M, hop,
                       b0, (const c2 *)twM, (const c2 *)cal, cal_nel, (c2 *)out);
    return LAUNCH_OK();
}

// ---- W1: R[k][a][b] += sum_i y_a(i) conj(y_b(i)) for gated k ------------------
// One block a channel, one thread an (a, b) pair; hops in ascending order and
// each term added to R as it is formed, the reference's order
// (src/channelizer.c:425-437), so R matches it to the bit.
#define COV_TILE 64
__global__ void __launch_bounds__(256)
k_pfb_cov(const c2 *__restrict__ y, unsigned nel, unsigned nchan, size_t m, size_t ypitch,
          const uint8_t *__restrict__ gate, c2 *__restrict__ R, uint32_t *__restrict__ counts) {
    const unsigned k = blockIdx.x;
    if (gate && !gate[k]) return;
    __shared__ c2 ys[32][COV_TILE];
    const unsigned pairs = nel * nel;
    c2 *Rk = R + (size_t)k * pairs;
    // up to 256 pairs per pass (nel <= 16); larger nel loops over pair groups
    for (unsigned pg = 0; pg < pairs; pg += 256) {
        const unsigned pr = pg + threadIdx.x;
        const unsigned a = pr / nel, b = pr % nel;
        float rr = 0.f, ri = 0.f;
        if (pr < pairs) { rr = Rk[pr].x; ri = Rk[pr].y; }
        for (size_t t0 = 0; t0 < m; t0 += COV_TILE) {
            const unsigned tn = (unsigned)(m - t0 < COV_TILE ? m - t0 : COV_TILE);
            __syncthreads();
            for (unsigned idx = threadIdx.x; idx < nel * COV_TILE && nel <= 32; idx += 256) {
                const unsigned el = idx / COV_TILE, ii = idx % COV_TILE;
                if (ii < tn) ys[el][ii] = y[((size_t)el * nchan + k) * ypitch + t0 + ii];
            }
            __syncthreads();
            if (pr < pairs && nel <= 32) {
                for (unsigned ii = 0; ii < tn; ++ii) {
                    const c2 ya = ys[a][ii], yb = ys[b][ii];
                    rr += ya.x * yb.x + ya.y * yb.y;
                    ri += ya.y * yb.x - ya.x * yb.y;
                }
            } else if (pr < pairs) {
                for (unsigned ii = 0; ii < tn; ++ii) {
                    const c2 ya = y[((size_t)a * nchan + k) * ypitch + t0 + ii];
                    const c2 yb = y[((size_t)b * nchan + k) * ypitch + t0 + ii];
                    rr += ya.x * yb.x + ya.y * yb.y;
                    ri += ya.y * yb.x - ya.x * yb.y;
                }
            }
        }
        if (pr < pairs) Rk[pr] = {rr, ri};
    }
    if (threadIdx.x == 0) counts[k] += (uint32_t)m;
}

__global__ void k_pfb_cov_reset(c2 *__restrict__ R, uint32_t *__restrict__ counts, unsigned nel,
                                unsigned nchan, const uint8_t *__restrict__ gate) {
    const unsigned k = blockIdx.x;
    if (gate && !gate[k]) return;
    const unsigned pairs = nel * nel;
    for (unsigned p = threadIdx.x; p < pairs; p += blockDim.x) R[(size_t)k * pairs + p] = {0.f, 0.f};
    if (threadIdx.x == 0) counts[k] = 0;
}

// ---- W2: per (element, channel) mean |y|^2 summed in FP64 (src/channelizer.c:453-471)
// One block a channel, one wavefront an element: lanes sum hops i = lane,
// lane + 64, ... in FP64, a wavefront reduction in FP64 finishes the sum,
// then the element means are averaged in element order as the reference
// does. Only the order of the FP64 hop sum differs from the reference's
// ascending loop (a change of about 1e-16 relative, before rounding to float).
__global__ void __launch_bounds__(1024) k_pfb_power(const c2 *__restrict__ y, unsigned nel, unsigned nchan,
                                                    size_t m, size_t ypitch, float *__restrict__ pwr,
                                                    float *__restrict__ mean_pwr) {
    const unsigned k = blockIdx.x;
    const unsigned w = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    __shared__ double es[16];
    const double inv = m ? 1.0 / (double)m : 0.0;
    double msum = 0.0;                         // thread 0's, in element order
    for (unsigned a0 = 0; a0 < nel; a0 += nw) {
        const unsigned a = a0 + w;
        if (a < nel) {
            const c2 *ya = y + ((size_t)a * nchan + k) * ypitch;
            double s = 0.0;
            for (size_t i = lane; i < m; i += 64) {
                const c2 v = ya[i];
                s += (double)v.x * v.x + (double)v.y * v.y;
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
            if (lane == 0) {
                es[w] = s * inv;
                if (pwr) pwr[(size_t)a * nchan + k] = (float)(s * inv);
            }
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (unsigned e = a0; e < nel && e < a0 + nw; ++e) msum += es[e - a0];
        __syncthreads();
    }
    if (threadIdx.x == 0 && mean_pwr) mean_pwr[k] = (float)(msum / (double)nel);
}

extern "C" int atkgpu_pfb_cov(void *s, const void *y, unsigned nel, unsigned nchan, size_t m, size_t ypitch,
                              const uint8_t *gate, void *R, uint32_t *counts) {
    if (!nchan) return 0;
    hipLaunchKernelGGL(k_pfb_cov, dim3(nchan), dim3(256), 0, (hipStream_t)s, (const c2 *)y, nel, nchan, m,
                       ypitch, gate, (c2 *)R, counts);
    return LAUNCH_OK();
}

extern "C" int atkgpu_pfb_cov_reset(void *s, void *R, uint32_t *counts, unsigned nel, unsigned nchan,
                                    const uint8_t *gate) {
    if (!nchan) return 0;
    hipLaunchKernelGGL(k_pfb_cov_reset, dim3(nchan), dim3(64), 0, (hipStream_t)s, (c2 *)R, counts, nel,
                       nchan, gate);
    return LAUNCH_OK();
}

extern "C" int atkgpu_pfb_power(void *s, const void *y, unsigned nel, unsigned nchan, size_t m, size_t ypitch,
                                float *pwr, float *mean_pwr) {
    if (!nchan) return 0;
    if (nel < 1) return -1;
    hipLaunchKernelGGL(k_pfb_power, dim3(nchan), dim3(64 * (nel < 16 ? nel : 16)), 0, (hipStream_t)s,
                       (const c2 *)y, nel, nchan, m, ypitch, pwr, mean_pwr);
    return LAUNCH_OK();
}
