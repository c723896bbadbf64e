// Kernels of the overlap-save extractor (reference src/channelizer.c:473-788)
// and of the arbitrary-ratio resampler (src/resample.c:100-252).
//
// The extractor's forward transform of every block (nfft points, blocks hop
// = nfft/2 apart in the stream row) is the batched FFT of fft.hip. Per channel:
//   k_osx_fold  the roll by the channel's integer bin and the product with its
//               filter's frequency response, folded to M = nfft/decim points:
//               Yf[k] = sum_j X[(k + j*M + sh) mod nfft] * H[k + j*M], j
//               ascending, in float (src/channelizer.c:695-705)
//   (the M-point inverse FFT of fft.hip, 1/M applied)
//   k_osx_mix   the ov/decim transient dropped, 1/decim, and the mix to
//               absolute time in FP64: phase -2 pi (a0 + delta*nabs/fs) with
//               a0 = ((sh*start) mod nfft)/nfft, start = bi*hop - ov and
//               nabs = bi*hop + k*decim (src/channelizer.c:707-728)
//   k_rs        the resampler's outputs K0 .. K0+count-1 of one call: output K
//               sits at input position base + K*step - in_off; a 2*Q-tap FIR
//               over the phase-b taps and the forward differences toward the
//               next phase, c = p + mu*d, newest input first; inputs before
//               the call come from the carried history (src/resample.c:206-252)
//   k_rs_hist   the last Q inputs of the call become the next history (into
//               the other of two buffers)
#include <hip/hip_runtime.h>
#include <math.h>
#include "atkgpu_dev.h"
#include "dev_common.hpp"
#include "fft_core.hpp"

// blocks of `t` (element e, block i) at X + t*nfft, t = e*m + i
__global__ void __launch_bounds__(256) k_osx_fold(const cpx *__restrict__ X, const cpx *__restrict__ H,
                                                  cpx *__restrict__ Yf, int nfft, int M, int decim, int sh,
                                                  int kblocks) {
#pragma clang fp contract(off)
    const int t = blockIdx.x / kblocks;
    const int k = (blockIdx.x % kblocks) * 256 + threadIdx.x;
    if (k >= M) return;
    const cpx *x = X + (long long)t * nfft;
    float ar = 0.0f, ai = 0.0f;
    for (int j = 0; j < decim; ++j) {
        const int bin = k + j * M;
        const cpx xv = x[(bin + sh) & (nfft - 1)];
        const cpx hv = H[bin];
        ar += xv.x * hv.x - xv.y * hv.y;
        ai += xv.x * hv.y + xv.y * hv.x;
    }
    Yf[(long long)t * M + k] = cmake(ar, ai);
}

struct OsxMixArgs {
    const cpx *ys;            // [nel*m][M] inverse transforms
    cpx *valid;               // element e, block i, sample k at valid + e*vpitch + i*nvalid + k
    long long vpitch;
    int m, M, ovd, nvalid, decim, hop, ov, nfft, sh;
    long long blk0;           // global index of block 0
    double delta, fs, invd;
    int kblocks;
};

__global__ void __launch_bounds__(256) k_osx_mix(OsxMixArgs a) {
#pragma clang fp contract(off)
    const int t = blockIdx.x / a.kblocks;          // e*m + i
    const int k = (blockIdx.x % a.kblocks) * 256 + threadIdx.x;
    if (k >= a.nvalid) return;
    const int e = t / a.m, i = t % a.m;
    const long long bi = a.blk0 + i;
    const long long start = bi * a.hop - a.ov;
    long long smod = ((long long)a.sh * start) % a.nfft;
    if (smod < 0) smod += a.nfft;
    const double a0 = (double)smod / (double)a.nfft;
    const long long nabs = bi * a.hop + (long long)k * a.decim;
    const double ph = -2.0 * M_PI * (a0 + a.delta * (double)nabs / a.fs);
    double si, cr;
    sincos(ph, &si, &cr);
    const cpx y = a.ys[(long long)t * a.M + a.ovd + k];
    const double yr = (double)y.x * a.invd, yi = (double)y.y * a.invd;
    a.valid[(long long)e * a.vpitch + (long long)i * a.nvalid + k] =
        cmake((float)(yr * cr - yi * si), (float)(yr * si + yi * cr));
}

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

// The reference's arithmetic, each operation rounded on its own (no
// contraction): the position base + K*step - in_off, the tap c = p + mu*d and
// acc += c*x, q ascending.
__global__ void __launch_bounds__(256) k_rs(RsArgs a) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int e = blockIdx.y;
    if (t >= a.count) return;
    const unsigned long long K = a.K0 + (unsigned long long)t;
    const double local = a.base + (double)K * a.step - a.inoff;
    const long long i = (long long)floor(local);
    const double frac = local - (double)i;
    const double ph = frac * (double)a.P;
    int b = (int)ph;
    if (b >= a.P) b = a.P - 1;
    const float mu = (float)(ph - (double)b);
    const cpx *in = a.in + (long long)e * a.ipitch;
    const cpx *hist = a.hist + (long long)e * a.Q;
    float ar = 0.0f, ai = 0.0f;
    for (int q = 0; q < a.Q; ++q) {
        const long long mm = i - q;
        cpx x;
        if (mm >= 0) {
            x = in[mm];
        } else {
            const long long hb = (long long)a.Q + mm;
            if (hb < 0) continue;
            x = hist[hb];
        }
        const float c = a.p[b + q * a.P] + mu * a.d[b + q * a.P];
        ar += c * x.x;
        ai += c * x.y;
    }
    a.out[(long long)e * a.opitch + t] = cmake(ar, ai);
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

extern "C" int atkgpu_osx_fold(void *s, const void *X, const void *H, void *Yf, size_t nt, unsigned nfft, unsigned M,
                               unsigned decim, unsigned sh) {
    if (!nt) return 0;
    const int kb = (int)((M + 255) / 256);
    const size_t blocks = nt * (size_t)kb;
    if (blocks > 0x7fffffffu) return -1;
    hipLaunchKernelGGL(k_osx_fold, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)s, (const cpx *)X,
                       (const cpx *)H, (cpx *)Yf, (int)nfft, (int)M, (int)decim, (int)sh, kb);
    return LAUNCH_OK();
}

extern "C" int atkgpu_osx_mix(void *s, const void *ys, void *valid, size_t vpitch, unsigned nel, size_t m, unsigned M,
                              unsigned ovd, unsigned decim, unsigned hop, unsigned ov, unsigned nfft, unsigned sh,
                              long long blk0, double delta, double fs) {
    if (!nel || !m) return 0;
    OsxMixArgs a;
    a.ys = (const cpx *)ys;
    a.valid = (cpx *)valid;
    a.vpitch = (long long)vpitch;
    a.m = (int)m;
    a.M = (int)M;
    a.ovd = (int)ovd;
    a.nvalid = (int)(M - ovd);
    a.decim = (int)decim;
    a.hop = (int)hop;
    a.ov = (int)ov;
    a.nfft = (int)nfft;
    a.sh = (int)sh;
    a.blk0 = blk0;
    a.delta = delta;
    a.fs = fs;
    a.invd = 1.0 / (double)decim;
    a.kblocks = (a.nvalid + 255) / 256;
    const size_t blocks = (size_t)nel * m * (size_t)a.kblocks;
    if (blocks > 0x7fffffffu) return -1;
    hipLaunchKernelGGL(k_osx_mix, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)s, a);
    return LAUNCH_OK();
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
