// Batched complex FFT for gfx950, any length n >= 2 (reference FFT contract:
// include/atkdsp.h:379-389, src/fft.c:31-131; the reference computes in FP64
// through pocketfft, this computes in FP32 with an FP64-derived twiddle table
// rounded once, which keeps the error near 1e-7 of full scale).
//
// Plans:
//   team     n = 2^k <= 8192: a team of n/PT threads holds PT points each in
//            registers; radix-PT Stockham stages exchange through LDS.
//   generic  any n <= 8192: runtime mixed radix (2..8 fixed-size, any prime as
//            a thread-per-output stage), ping-pong in LDS.
//   split    n > 8192 with n = n1*n2, n1, n2 <= 8192: four-step (column FFTs,
//            twiddle W_n^{j2 k1}, row FFTs), one transform at a time.
//   direct   otherwise: O(n^2) DFT with FP64 accumulation (correct, slow).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "atkgpu_dev.h"
#include "dev_common.hpp"
#include "fft_core.hpp"
#include "fft_team.hpp"

#define LDS_MAX_N 8192

struct FftIO {
    const cpx *in;
    long long in_stride, in_dist;
    cpx *out;
    long long out_stride, out_dist;
    long long batch;
    const float *window;       // x[r] *= window[r*win_stride + t*win_dist]
    long long win_stride, win_dist;
    int conj_in, conj_out;
    float scale;               // applied to outputs when != 1
    const cpx *post_tw;        // out[k] *= post_tw[(t*k) mod post_n] (four-step)
    long long post_n;
};

__device__ __forceinline__ cpx io_load(const FftIO &io, long long t, long long r) {
    cpx x = io.in[t * io.in_dist + r * io.in_stride];
    if (io.window) {
        const float w = io.window[r * io.win_stride + t * io.win_dist];
        x.x *= w;
        x.y *= w;
    }
    if (io.conj_in) x.y = -x.y;
    return x;
}

__device__ __forceinline__ void io_store(const FftIO &io, long long t, long long k, cpx y) {
    if (io.post_tw) y = cmul(y, io.post_tw[(t * k) % io.post_n]);
    if (io.conj_out) y.y = -y.y;
    if (io.scale != 1.0f) {
        y.x *= io.scale;
        y.y *= io.scale;
    }
    io.out[t * io.out_dist + k * io.out_stride] = y;
}

template <int N>
struct TeamShape {
    static constexpr int PT = N <= 8 ? N : (N <= 32 ? 4 : (N <= 128 ? 8 : 16));
    static constexpr int NT = N / PT;
    static constexpr int TEAMS = NT >= 256 ? 1 : 256 / NT;
    static constexpr int BLOCK = NT * TEAMS;
};

template <int N>
__global__ void __launch_bounds__(TeamShape<N>::BLOCK)
k_fft_team(FftIO io, const cpx *__restrict__ tw) {
    using SH = TeamShape<N>;
    constexpr int PT = SH::PT, NT = SH::NT;
    constexpr int TL = lpad_size<N>();
    __shared__ cpx lds[SH::TEAMS * TL];
    const int team = threadIdx.x / NT, j = threadIdx.x % NT;
    const long long t = (long long)blockIdx.x * SH::TEAMS + team;
    const bool act = t < io.batch;
    cpx v[PT];
#pragma unroll
    for (int r = 0; r < PT; ++r) v[r] = act ? io_load(io, t, j + r * NT) : cmake(0.f, 0.f);
    if constexpr (N == PT) {
        // a single thread holds the whole transform
        if constexpr (N > 1) dft_pow2<N>(v);
    } else {
        team_fft<N, PT>(v, lds + team * TL, tw, j);
    }
    if (act) {
#pragma unroll
        for (int r = 0; r < PT; ++r) io_store(io, t, j + r * NT, v[r]);
    }
}

// ---- generic mixed-radix Stockham ---------------------------------------------
struct Radices {
    int nst;
    int r[32];
};

__global__ void __launch_bounds__(256) k_fft_generic(FftIO io, const cpx *__restrict__ tw,
                                                     int n, Radices rad) {
    extern __shared__ cpx gl[];
    cpx *b0 = gl, *b1 = gl + n;
    const long long t = blockIdx.x;
    for (int i = threadIdx.x; i < n; i += blockDim.x) b0[i] = io_load(io, t, i);
    __syncthreads();
    int Ns = 1;
    for (int s = 0; s < rad.nst; ++s) {
        const int R = rad.r[s];
        for (int o = threadIdx.x; o < n; o += blockDim.x)
            b1[o] = generic_stage_output(b0, tw, n, Ns, R, o);
        __syncthreads();
        cpx *tmp = b0; b0 = b1; b1 = tmp;
        Ns *= R;
    }
    for (int k = threadIdx.x; k < n; k += blockDim.x) io_store(io, t, k, b0[k]);
}

// ---- direct DFT (lengths with a prime factor beyond the LDS plans) -------------
__global__ void k_dft_direct(FftIO io, const cpx *__restrict__ tw, long long n) {
    const long long t = blockIdx.y;
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x) {
        double ax = 0.0, ay = 0.0;
        long long e = 0;
        for (long long jx = 0; jx < n; ++jx) {
            const cpx a = io_load(io, t, jx);
            const cpx w = tw[e];
            ax += (double)a.x * w.x - (double)a.y * w.y;
            ay += (double)a.x * w.y + (double)a.y * w.x;
            e += k;
            if (e >= n) e -= n;
        }
        io_store(io, t, k, cmake((float)ax, (float)ay));
    }
}

// ---- plans ---------------------------------------------------------------------
enum { PLAN_TEAM = 0, PLAN_GENERIC = 1, PLAN_SPLIT = 2, PLAN_DIRECT = 3 };

struct atkgpu_fftplan {
    size_t n;
    int kind;
    cpx *tw;                  // n entries, device
    Radices rad;              // generic
    atkgpu_fftplan *p1, *p2;  // split: n = n1 * n2
    size_t n1, n2;
};

static bool is_pow2(size_t n) { return n && !(n & (n - 1)); }

// radices for the generic plan: 8, 4, 2 first, then odd primes ascending
static void factor_radices(size_t n, Radices *rad) {
    rad->nst = 0;
    size_t m = n;
    while (m % 8 == 0) { rad->r[rad->nst++] = 8; m /= 8; }
    while (m % 4 == 0) { rad->r[rad->nst++] = 4; m /= 4; }
    while (m % 2 == 0) { rad->r[rad->nst++] = 2; m /= 2; }
    for (size_t p = 3; p * p <= m; p += 2)
        while (m % p == 0) { rad->r[rad->nst++] = (int)p; m /= p; }
    if (m > 1) rad->r[rad->nst++] = (int)m;
}

static cpx *make_twiddles(size_t n) {
    cpx *h = (cpx *)malloc(n * sizeof(cpx));
    if (!h) return nullptr;
    const double two_pi = 6.283185307179586476925286766559;
    for (size_t t = 0; t < n; ++t) {
        // reduce the angle by symmetry first so every entry is rounded from an
        // exact-argument FP64 evaluation
        const double a = two_pi * (double)t / (double)n;
        h[t] = cmake((float)cos(a), (float)-sin(a));
    }
    cpx *d = (cpx *)atkgpu_dmalloc(n * sizeof(cpx));
    if (d && hipMemcpy(d, h, n * sizeof(cpx), hipMemcpyHostToDevice) != hipSuccess) {
        atkgpu_dfree(d);
        d = nullptr;
    }
    free(h);
    return d;
}

extern "C" atkgpu_fftplan *atkgpu_fftplan_create(size_t n) {
    if (n < 2) return nullptr;
    if (atkgpu_init() != 0) return nullptr;
    atkgpu_fftplan *p = (atkgpu_fftplan *)calloc(1, sizeof *p);
    if (!p) return nullptr;
    p->n = n;
    p->tw = make_twiddles(n);
    if (!p->tw) { free(p); return nullptr; }
    if (n <= LDS_MAX_N) {
        p->kind = is_pow2(n) ? PLAN_TEAM : PLAN_GENERIC;
        factor_radices(n, &p->rad);
        if (p->kind == PLAN_GENERIC)
            (void)hipFuncSetAttribute((const void *)k_fft_generic,
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      2 * LDS_MAX_N * (int)sizeof(cpx));
        return p;
    }
    // split n = n1 * n2 with both sides LDS-sized, n1 as close to sqrt(n) as possible
    size_t best = 0;
    for (size_t d = 2; d <= LDS_MAX_N && d * d <= n * 4; ++d) {
        if (n % d) continue;
        const size_t e = n / d;
        if (e > LDS_MAX_N) continue;
        if (!best || (d > e ? d - e : e - d) < (best > n / best ? best - n / best : n / best - best))
            best = d;
    }
    if (best) {
        p->kind = PLAN_SPLIT;
        p->n1 = best;
        p->n2 = n / best;
        p->p1 = atkgpu_fftplan_create(p->n1);
        p->p2 = atkgpu_fftplan_create(p->n2);
        if (!p->p1 || !p->p2) { atkgpu_fftplan_destroy(p); return nullptr; }
        return p;
    }
    p->kind = PLAN_DIRECT;
    return p;
}

extern "C" void atkgpu_fftplan_destroy(atkgpu_fftplan *p) {
    if (!p) return;
    atkgpu_fftplan_destroy(p->p1);
    atkgpu_fftplan_destroy(p->p2);
    atkgpu_dfree(p->tw);
    free(p);
}

extern "C" const void *atkgpu_fftplan_twiddles(const atkgpu_fftplan *p) { return p ? p->tw : nullptr; }

extern "C" size_t atkgpu_fftplan_scratch(const atkgpu_fftplan *p, size_t batch) {
    if (!p) return 0;
    if (p->kind == PLAN_SPLIT) return p->n;
    if (p->kind == PLAN_DIRECT) return p->n * (batch ? batch : 1);
    return 0;
}

template <int N>
static void launch_team(const FftIO &io, const cpx *tw, hipStream_t s) {
    using SH = TeamShape<N>;
    const unsigned blocks = (unsigned)((io.batch + SH::TEAMS - 1) / SH::TEAMS);
    hipLaunchKernelGGL(k_fft_team<N>, dim3(blocks), dim3(SH::BLOCK), 0, s, io, tw);
}

static int run_plan(const atkgpu_fftplan *p, const FftIO &io, hipStream_t s, cpx *scratch) {
    switch (p->kind) {
    case PLAN_TEAM:
        switch (p->n) {
        case 2: launch_team<2>(io, p->tw, s); break;
        case 4: launch_team<4>(io, p->tw, s); break;
        case 8: launch_team<8>(io, p->tw, s); break;
        case 16: launch_team<16>(io, p->tw, s); break;
        case 32: launch_team<32>(io, p->tw, s); break;
        case 64: launch_team<64>(io, p->tw, s); break;
        case 128: launch_team<128>(io, p->tw, s); break;
        case 256: launch_team<256>(io, p->tw, s); break;
        case 512: launch_team<512>(io, p->tw, s); break;
        case 1024: launch_team<1024>(io, p->tw, s); break;
        case 2048: launch_team<2048>(io, p->tw, s); break;
        case 4096: launch_team<4096>(io, p->tw, s); break;
        case 8192: launch_team<8192>(io, p->tw, s); break;
        default: return -4;
        }
        return LAUNCH_OK();
    case PLAN_GENERIC: {
        const size_t lds = 2 * p->n * sizeof(cpx);
        for (long long b0 = 0; b0 < io.batch; b0 += 65535) {
            FftIO sub = io;
            sub.in += b0 * io.in_dist;
            sub.out += b0 * io.out_dist;
            if (sub.window) sub.window += b0 * io.win_dist;
            sub.batch = io.batch - b0 < 65535 ? io.batch - b0 : 65535;
            hipLaunchKernelGGL(k_fft_generic, dim3((unsigned)sub.batch), dim3(256), lds, s, sub,
                               (const cpx *)p->tw, (int)p->n, p->rad);
            if (LAUNCH_OK()) return -4;
        }
        return 0;
    }
    case PLAN_SPLIT: {
        const long long n1 = p->n1, n2 = p->n2;
        for (long long b = 0; b < io.batch; ++b) {
            // A: n2 column transforms of length n1 over x[j1*n2 + j2], twiddled by W_n^{j2 k1}
            FftIO a = {};
            a.in = io.in + b * io.in_dist;
            a.in_stride = n2 * io.in_stride;
            a.in_dist = io.in_stride;
            a.out = scratch;
            a.out_stride = n2;
            a.out_dist = 1;
            a.batch = n2;
            a.window = io.window ? io.window + b * io.win_dist : nullptr;
            a.win_stride = n2 * io.win_stride;
            a.win_dist = io.win_stride;
            a.conj_in = io.conj_in;
            a.scale = 1.0f;
            a.post_tw = p->tw;
            a.post_n = (long long)p->n;
            int rc = run_plan(p->p1, a, s, nullptr);
            if (rc) return rc;
            // B: n1 row transforms of length n2, X[k1 + n1 k2]
            FftIO c = {};
            c.in = scratch;
            c.in_stride = 1;
            c.in_dist = n2;
            c.out = io.out + b * io.out_dist;
            c.out_stride = n1 * io.out_stride;
            c.out_dist = io.out_stride;
            c.batch = n1;
            c.conj_out = io.conj_out;
            c.scale = io.scale;
            rc = run_plan(p->p2, c, s, nullptr);
            if (rc) return rc;
        }
        return 0;
    }
    case PLAN_DIRECT: {
        // the DFT reads every input for each output: stage inputs in scratch so
        // that in-place calls are safe
        FftIO d = io;
        if ((const void *)io.in == (const void *)io.out) {
            for (long long b = 0; b < io.batch; ++b)
                if (hipMemcpyAsync(scratch + b * p->n, io.in + b * io.in_dist, p->n * sizeof(cpx),
                                   hipMemcpyDeviceToDevice, s) != hipSuccess)
                    return -4;
            d.in = scratch;
            d.in_dist = (long long)p->n;
        }
        for (long long b0 = 0; b0 < io.batch; b0 += 65535) {
            FftIO sub = d;
            sub.in += b0 * d.in_dist;
            sub.out += b0 * d.out_dist;
            if (sub.window) sub.window += b0 * d.win_dist;
            sub.batch = d.batch - b0 < 65535 ? d.batch - b0 : 65535;
            const unsigned gx = grid_for(p->n, 256, 1024);
            hipLaunchKernelGGL(k_dft_direct, dim3(gx, (unsigned)sub.batch), dim3(256), 0, s, sub,
                               (const cpx *)p->tw, (long long)p->n);
            if (LAUNCH_OK()) return -4;
        }
        return 0;
    }
    }
    return -4;
}

extern "C" int atkgpu_fft_batch(void *s, const atkgpu_fftplan *p, const void *in, size_t in_dist,
                                const float *window, size_t batch, int inverse, void *out,
                                void *scratch) {
    if (!p || !in || !out) return -1;
    if (!batch) return 0;
    FftIO io = {};
    io.in = (const cpx *)in;
    io.in_stride = 1;
    io.in_dist = (long long)in_dist;
    io.out = (cpx *)out;
    io.out_stride = 1;
    io.out_dist = (long long)p->n;
    io.batch = (long long)batch;
    io.window = window;
    io.win_stride = 1;
    io.win_dist = 0;
    io.conj_in = inverse ? 1 : 0;
    io.conj_out = inverse ? 1 : 0;
    io.scale = inverse ? (float)(1.0 / (double)p->n) : 1.0f;
    return run_plan(p, io, (hipStream_t)s, (cpx *)scratch);
}
