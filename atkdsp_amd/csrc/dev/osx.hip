// Kernels of the overlap-save extractor (reference src/channelizer.c:473-788)
// and of the arbitrary-ratio resampler (src/resample.c:100-252).
//
// The extractor's forward transform of every block (nfft points, blocks hop
// = nfft/2 apart in the stream row) is the batched FFT of fft.hip. Then, for
// every channel of a sub-call at once (one descriptor per channel):
//   k_osx_fold_all  the roll by the channel's integer bin and the product with
//                   its filter's frequency response, folded to M = nfft/decim
//                   points: Yf[k] = sum_j X[(k + j*M + sh) mod nfft] *
//                   H[k + j*M] (src/channelizer.c:695-705), in float; when M
//                   is small, 256/M threads split an output's sum (strided j)
//                   and add their parts in a fixed order
//   (the M-point inverse FFT of fft.hip per channel, 1/M applied)
//   k_osx_mix_all   the ov/decim transient dropped, 1/decim, and the mix to
//                   absolute time in FP64: phase -2 pi (a0 + delta*nabs/fs) with
//                   a0 = ((sh*start) mod nfft)/nfft, start = bi*hop - ov and
//                   nabs = bi*hop + k*decim (src/channelizer.c:707-728)
//   k_osx_rs_all    the channels' resamplers (below), then their histories
// The resampler (also the drop-in atkdsp_arb_resampler):
//   k_rs        outputs K0 .. K0+count-1 of one call: output K sits at input
//               position base + K*step - in_off; a 2*Q-tap FIR over the
//               phase-b taps and the forward differences toward the next
//               phase, c = p + mu*d, newest input first; inputs before the
//               call come from the carried history (src/resample.c:206-252)
//   k_rs_hist   the last Q inputs of the call become the next history (into
//               the other of two buffers)
#include <hip/hip_runtime.h>
#include <math.h>
#include "atkgpu_dev.h"
#include "dev_common.hpp"
#include "fft_core.hpp"

struct RsArgs {
    const cpx *in;            // element e's inputs of this call at in + e*ipitch
    long long ipitch;
    const cpx *hist;          // element e's last Q inputs before the call, oldest first
    const float *p, *d;       // P*Q prototype (gain P) and forward differences
    int P, Q;
    double base, step, inoff;
    unsigned long long K0;
    cpx *out;                 // element e's outputs at out + e*opitch
    long long opitch;
    int count;
};

// One output of the resampler. The reference's arithmetic, each operation
// rounded on its own (no contraction): the position base + K*step - in_off,
// the tap c = p + mu*d and acc += c*x, q ascending.
__device__ __forceinline__ cpx rs_output(const cpx *in, const cpx *hist, const float *p, const float *d, int P,
                                         int Q, double base, double step, double inoff, unsigned long long K) {
#pragma clang fp contract(off)
    const double local = base + (double)K * step - inoff;
    const long long i = (long long)floor(local);
    const double frac = local - (double)i;
    const double ph = frac * (double)P;
    int b = (int)ph;
    if (b >= P) b = P - 1;
    const float mu = (float)(ph - (double)b);
    float ar = 0.0f, ai = 0.0f;
    for (int q = 0; q < Q; ++q) {
        const long long mm = i - q;
        cpx x;
        if (mm >= 0) {
            x = in[mm];
        } else {
            const long long hb = (long long)Q + mm;
            if (hb < 0) continue;
            x = hist[hb];
        }
        const float c = p[b + q * P] + mu * d[b + q * P];
        ar += c * x.x;
        ai += c * x.y;
    }
    return cmake(ar, ai);
}

__global__ void __launch_bounds__(256) k_rs(RsArgs a) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int e = blockIdx.y;
    if (t >= a.count) return;
    a.out[(long long)e * a.opitch + t] =
        rs_output(a.in + (long long)e * a.ipitch, a.hist + (long long)e * a.Q, a.p, a.d, a.P, a.Q, a.base, a.step,
                  a.inoff, a.K0 + (unsigned long long)t);
}

// next history (into hout) = the last Q of [hist (Q) ++ in (n)]
__global__ void k_rs_hist(const cpx *__restrict__ in, long long ipitch, long long n, const cpx *__restrict__ hin,
                          cpx *__restrict__ hout, int Q) {
    const int e = blockIdx.y;
    for (int j = blockIdx.x * 256 + threadIdx.x; j < Q; j += gridDim.x * 256) {
        const long long src = n + j;              // index into hist ++ in
        hout[(long long)e * Q + j] = src >= Q ? in[(long long)e * ipitch + (src - Q)] : hin[(long long)e * Q + src];
    }
}

extern "C" int atkgpu_rs_run(void *s, const void *in, size_t ipitch, size_t n, unsigned nel, const void *hist_in,
                             void *hist_out, const float *p, const float *d, unsigned P, unsigned Q, double base,
                             double step, unsigned long long K0, unsigned long long in_off, void *out, size_t opitch,
                             size_t count) {
    if (!nel) return 0;
    if (count) {
        if (count > 0x7fffffffu) return -1;
        RsArgs a;
        a.in = (const cpx *)in;
        a.ipitch = (long long)ipitch;
        a.hist = (const cpx *)hist_in;
        a.p = p;
        a.d = d;
        a.P = (int)P;
        a.Q = (int)Q;
        a.base = base;
        a.step = step;
        a.inoff = (double)in_off;
        a.K0 = K0;
        a.out = (cpx *)out;
        a.opitch = (long long)opitch;
        a.count = (int)count;
        hipLaunchKernelGGL(k_rs, dim3((unsigned)((count + 255) / 256), nel), dim3(256), 0, (hipStream_t)s, a);
        if (LAUNCH_OK()) return -4;
    }
    const unsigned gx = (Q + 255) / 256;
    hipLaunchKernelGGL(k_rs_hist, dim3(gx, nel), dim3(256), 0, (hipStream_t)s, (const cpx *)in, (long long)ipitch,
                       (long long)n, (const cpx *)hist_in, (cpx *)hist_out, (int)Q);
    return LAUNCH_OK();
}

// ---- every channel of a sub-call ------------------------------------------------------
constexpr int OSX_FOLD_CPB = 4;               // channels a fold block takes

// block (t, g): block t = e*m + i of the forward transforms, channels
// [g*OSX_FOLD_CPB, (g+1)*OSX_FOLD_CPB)
__global__ void __launch_bounds__(256) k_osx_fold_all(const cpx *__restrict__ X, const atkgpu_osx_chan *chans,
                                                      int nch, int m, long long xpitch, int nfft) {
#pragma clang fp contract(off)
    __shared__ atkgpu_osx_chan dl[OSX_FOLD_CPB];
    __shared__ float red_r[256], red_i[256];
    const int t = blockIdx.x;                 // e*m + i
    const int e = t / m, i = t % m;
    const cpx *x = X + (long long)e * xpitch + (long long)i * nfft;
    {
        const int c0 = blockIdx.y * OSX_FOLD_CPB;
        const int nc = min(OSX_FOLD_CPB, nch - c0);
        __syncthreads();
        // the descriptors may sit in host memory: fetch them together
        {
            const int words = nc * (int)(sizeof(atkgpu_osx_chan) / 4);
            const unsigned *src = (const unsigned *)(chans + c0);
            unsigned *dst = (unsigned *)dl;
            for (int w = threadIdx.x; w < words; w += 256) dst[w] = src[w];
        }
        __syncthreads();
        for (int c = 0; c < nc; ++c) {
            const int M = dl[c].M, decim = dl[c].decim, sh = dl[c].sh;
            const cpx *H = (const cpx *)dl[c].H;
            cpx *Y = (cpx *)dl[c].Y + (long long)t * M;
            // tpk threads per output, consecutive lanes on consecutive outputs
            // (coalesced loads); part p sums j = p, p + tpk, ...
            const int tpk = M >= 256 ? 1 : 256 / M;
            const int kstep = 256 / tpk;
            const int part = threadIdx.x / kstep;
            for (int k0 = 0; k0 < M; k0 += kstep) {
                const int k = k0 + threadIdx.x % kstep;
                float ar = 0.0f, ai = 0.0f;
#pragma unroll 8
                for (int j = part; j < decim; j += tpk) {
                    const int bin = k + j * M;
                    const cpx xv = x[(bin + sh) & (nfft - 1)];
                    const cpx hv = H[bin];
                    ar += xv.x * hv.x - xv.y * hv.y;
                    ai += xv.x * hv.y + xv.y * hv.x;
                }
                if (tpk == 1) {
                    Y[k] = cmake(ar, ai);
                    continue;
                }
                red_r[threadIdx.x] = ar;
                red_i[threadIdx.x] = ai;
                __syncthreads();
                if (part == 0) {                          // the parts in order
                    float sr = red_r[threadIdx.x], si = red_i[threadIdx.x];
                    for (int q = 1; q < tpk; ++q) {
                        sr += red_r[threadIdx.x + q * kstep];
                        si += red_i[threadIdx.x + q * kstep];
                    }
                    Y[k] = cmake(sr, si);
                }
                __syncthreads();
            }
        }
    }
}

__global__ void __launch_bounds__(256) k_osx_mix_all(const atkgpu_osx_chan *chans, int m, int hop, int ov, int nfft,
                                                     long long blk0, double fs) {
#pragma clang fp contract(off)
    const atkgpu_osx_chan &d = chans[blockIdx.y];
    const int t = blockIdx.x;                 // e*m + i
    const int e = t / m, i = t % m;
    const int M = d.M, ovd = d.ovd, nvalid = d.nvalid, decim = d.decim, sh = d.sh;
    const double delta = d.delta, invd = 1.0 / (double)decim;
    const cpx *ys = (const cpx *)d.Y + (long long)t * M + ovd;
    cpx *v = (cpx *)d.V + (long long)e * d.vpitch + (long long)i * nvalid;
    const long long bi = blk0 + i;
    const long long start = bi * hop - ov;
    long long smod = ((long long)sh * start) % nfft;
    if (smod < 0) smod += nfft;
    const double a0 = (double)smod / (double)nfft;
    for (int k = threadIdx.x; k < nvalid; k += 256) {
        const long long nabs = bi * hop + (long long)k * decim;
        const double ph = -2.0 * M_PI * (a0 + delta * (double)nabs / fs);
        double si, cr;
        sincos(ph, &si, &cr);
        const cpx y = ys[k];
        const double yr = (double)y.x * invd, yi = (double)y.y * invd;
        v[k] = cmake((float)(yr * cr - yi * si), (float)(yr * si + yi * cr));
    }
}

__global__ void __launch_bounds__(256) k_osx_rs_all(const atkgpu_osx_chan *chans, int m) {
    const atkgpu_osx_chan &d = chans[blockIdx.y];
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int e = blockIdx.z;
    if (!d.rs || t >= d.count) return;
    const cpx *in = (const cpx *)d.V + (long long)e * d.vpitch;
    ((cpx *)d.out)[(long long)e * d.ocap + t] =
        rs_output(in, (const cpx *)d.hist_in + (long long)e * d.Q, d.p, d.d, d.P, d.Q, d.base, d.step, d.inoff,
                  d.K0 + (unsigned long long)t);
}

__global__ void k_osx_rs_hist_all(const atkgpu_osx_chan *chans, int m) {
    const atkgpu_osx_chan &d = chans[blockIdx.y];
    const int e = blockIdx.z;
    if (!d.rs) return;
    const long long n = (long long)m * d.nvalid;
    const cpx *in = (const cpx *)d.V + (long long)e * d.vpitch;
    const cpx *hin = (const cpx *)d.hist_in + (long long)e * d.Q;
    cpx *hout = (cpx *)d.hist_out + (long long)e * d.Q;
    for (int j = blockIdx.x * 256 + threadIdx.x; j < d.Q; j += gridDim.x * 256) {
        const long long src = n + j;
        hout[j] = src >= d.Q ? in[src - d.Q] : hin[src];
    }
}

extern "C" int atkgpu_osx_fold_all(void *s, const void *X, const void *chans, int nch, unsigned nel, size_t m,
                                   size_t xpitch, unsigned nfft) {
    if (nch <= 0 || !nel || !m) return 0;
    const size_t blocks = (size_t)nel * m;
    const unsigned groups = (unsigned)((nch + OSX_FOLD_CPB - 1) / OSX_FOLD_CPB);
    if (blocks > 0x7fffffffu || groups > 65535) return -1;
    hipLaunchKernelGGL(k_osx_fold_all, dim3((unsigned)blocks, groups), dim3(256), 0, (hipStream_t)s,
                       (const cpx *)X, (const atkgpu_osx_chan *)chans, nch, (int)m, (long long)xpitch, (int)nfft);
    return LAUNCH_OK();
}

extern "C" int atkgpu_osx_mix_all(void *s, const void *chans, int nch, unsigned nel, size_t m, int max_nvalid,
                                  unsigned hop, unsigned ov, unsigned nfft, long long blk0, double fs) {
    if (nch <= 0 || !nel || !m || max_nvalid <= 0) return 0;
    const size_t blocks = (size_t)nel * m;
    if (blocks > 0x7fffffffu || nch > 65535) return -1;
    hipLaunchKernelGGL(k_osx_mix_all, dim3((unsigned)blocks, (unsigned)nch), dim3(256), 0, (hipStream_t)s,
                       (const atkgpu_osx_chan *)chans, (int)m, (int)hop, (int)ov, (int)nfft, blk0, fs);
    return LAUNCH_OK();
}

extern "C" int atkgpu_osx_rs_all(void *s, const void *chans, int nch, unsigned nel, size_t m, int max_count,
                                 int max_q) {
    if (nch <= 0 || !nel || !m || nch > 65535 || nel > 65535) return 0;
    const atkgpu_osx_chan *c = (const atkgpu_osx_chan *)chans;
    if (max_count > 0) {
        hipLaunchKernelGGL(k_osx_rs_all, dim3((unsigned)((max_count + 255) / 256), (unsigned)nch, nel), dim3(256), 0,
                           (hipStream_t)s, c, (int)m);
        if (LAUNCH_OK()) return -4;
    }
    if (max_q > 0) {
        hipLaunchKernelGGL(k_osx_rs_hist_all, dim3((unsigned)((max_q + 255) / 256), (unsigned)nch, nel), dim3(256),
                           0, (hipStream_t)s, c, (int)m);
    }
    return LAUNCH_OK();
}
