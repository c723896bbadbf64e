// Kernels of the device-resident front-end chain: many independent streams
// of one shape, each launch covering every stream.
//
// Each stream keeps its samples in a row of a slab buffer:
//   [kept prefix (keep samples) | this block (n samples)]
// The prefix is the last `keep` samples of the stream so far, so the filterbank
// finds its history (ov + lc samples) and the display its unfinished frame
// tail (< nfft samples) just before the block. After a block the last keep
// samples move to the other slab buffer.
//
//   k_disp_unpack    the display reduction straight from the raw block: each
//                    frame's samples are converted as they are loaded
//                    (bit-exact unpack_dc, reference src/convert.c:33-111,
//                    with the DC offset from the stream's previous-block mean,
//                    device state); each converted sample is also written
//                    into the row once, by the frame whose first hop holds it
//                    (the last frame takes the rest of the block), and summed
//                    in FP64 for the block mean; then window, team FFT,
//                    |X|^2/n^2, per-bin detector in registers (reference
//                    src/fft.c:177-191), partial lines
//   k_chain_unpack   raw -> cf32 into the row with the FP64 partial sums, for
//                    a block that completes no display frame
//   k_chain_mean     per stream: FP64 mean -> next block's offset
//   k_disp_final     partial lines in fixed order -> dB line (src/fft.c:286-292)
//   k_disp_post      decimate_max (src/display.c:19-37) and the NaN-aware
//                    stitch of the decimated line (src/detect.c:116-134)
//   k_chain_keep     last keep samples -> the other slab buffer
#include <hip/hip_runtime.h>
#include <math.h>
#include "atkgpu_dev.h"
#include "dev_common.hpp"
#include "fft_core.hpp"
#include "fft_team.hpp"
#include "disp.hpp"

// ---- unpack with a per-stream offset ---------------------------------------------
struct ChainUnpackArgs {
    const unsigned char *raw;   // stream s at raw + s*raw_pitch
    long long raw_pitch;
    int fmt;
    long long n;
    cpx *row0;                  // stream s block start: row0 + s*spitch
    long long spitch;
    const float *off;           // per stream offset (re, im); null = 0
    double *partials;           // [S][parts][2]
    int parts;
};

template <int FMT>
__device__ __forceinline__ cpx chain_conv(const unsigned char *raw, long long i, float cre, float cim, float k,
                                          float ore, float oim) {
#pragma clang fp contract(off)
    cpx o;
    if constexpr (FMT == FMT_CU8) {
        const uchar2 b = ((const uchar2 *)raw)[i];
        o.x = ((float)b.x - cre) * k;
        o.y = ((float)b.y - cim) * k;
    } else if constexpr (FMT == FMT_CI8) {
        const char2 b = ((const char2 *)raw)[i];
        o.x = ((float)(signed char)b.x - cre) * k;
        o.y = ((float)(signed char)b.y - cim) * k;
    } else if constexpr (FMT == FMT_CI16) {
        const short2 b = ((const short2 *)raw)[i];
        o.x = ((float)b.x - cre) * k;
        o.y = ((float)b.y - cim) * k;
    } else {
        const float2 b = ((const float2 *)raw)[i];
        o.x = b.x - ore;
        o.y = b.y - oim;
    }
    return o;
}

template <int FMT>
__global__ void __launch_bounds__(256) k_chain_unpack(ChainUnpackArgs a, float c0, float kk) {
#pragma clang fp contract(off)
    const int s = blockIdx.y;
    const float ore = a.off ? a.off[2 * s] : 0.0f, oim = a.off ? a.off[2 * s + 1] : 0.0f;
    // the folded constants exactly as src/convert.c:47, :57, :68 (IEEE division)
    const float cre = c0 + ore / kk, cim = c0 + oim / kk;
    const unsigned char *raw = a.raw + (long long)s * a.raw_pitch;
    cpx *row = a.row0 + (long long)s * a.spitch;
    double sr = 0.0, si = 0.0;
    const long long stride = (long long)gridDim.x * 256;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < a.n; i += stride) {
        const cpx o = chain_conv<FMT>(raw, i, cre, cim, kk, ore, oim);
        // streamed past the caches: the row is read back by the filterbank
        // and display kernels only after every block is written
        unsigned long long bits;
        __builtin_memcpy(&bits, &o, 8);
        __builtin_nontemporal_store(bits, (unsigned long long *)&row[i]);
        sr += (double)o.x;
        si += (double)o.y;
    }
    __shared__ double shr[256], shi[256];
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
        double *p = a.partials + ((long long)s * a.parts + blockIdx.x) * 2;
        p[0] = shr[0];
        p[1] = shi[0];
    }
}

// per stream: sum the partials in order, mean = sum/n + off; with update != 0
// the offset for the next block becomes (float)mean
__global__ void k_chain_mean(const double *__restrict__ partials, int parts, long long n, float *off,
                             double *__restrict__ mean, int update, int S) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    double sr = 0.0, si = 0.0;
    for (int b = 0; b < parts; ++b) {
        sr += partials[((long long)s * parts + b) * 2];
        si += partials[((long long)s * parts + b) * 2 + 1];
    }
    const double inv = n ? 1.0 / (double)n : 0.0;
    const double ore = off ? (double)off[2 * s] : 0.0, oim = off ? (double)off[2 * s + 1] : 0.0;
    const double mr = sr * inv + ore, mi = si * inv + oim;
    mean[2 * s] = mr;
    mean[2 * s + 1] = mi;
    if (update && off) {
        off[2 * s] = (float)mr;
        off[2 * s + 1] = (float)mi;
    }
}

extern "C" int atkgpu_chain_unpack(void *st, const void *raw, size_t raw_pitch, int fmt, size_t n, void *row0,
                                   size_t spitch, float *off, double *partials, int parts, double *mean,
                                   int update, unsigned S) {
    if (!n || !S) return 0;
    ChainUnpackArgs a;
    a.raw = (const unsigned char *)raw;
    a.raw_pitch = (long long)raw_pitch;
    a.fmt = fmt;
    a.n = (long long)n;
    a.row0 = (cpx *)row0;
    a.spitch = (long long)spitch;
    a.off = off;
    a.partials = partials;
    a.parts = parts;
    hipStream_t s = (hipStream_t)st;
    dim3 grid((unsigned)parts, S);
    float c0 = 0.0f, kk = 1.0f;
    switch (fmt) {
    case FMT_CU8:
        c0 = 127.5f; kk = 1.0f / 127.5f;
        hipLaunchKernelGGL(k_chain_unpack<FMT_CU8>, grid, dim3(256), 0, s, a, c0, kk);
        break;
    case FMT_CI8:
        kk = 1.0f / 128.0f;
        hipLaunchKernelGGL(k_chain_unpack<FMT_CI8>, grid, dim3(256), 0, s, a, c0, kk);
        break;
    case FMT_CI16:
    case FMT_CI16Q11:
        kk = fmt == FMT_CI16 ? 1.0f / 32768.0f : 1.0f / 2048.0f;
        hipLaunchKernelGGL(k_chain_unpack<FMT_CI16>, grid, dim3(256), 0, s, a, c0, kk);
        break;
    case FMT_CF32:
        hipLaunchKernelGGL(k_chain_unpack<FMT_CF32>, grid, dim3(256), 0, s, a, c0, kk);
        break;
    default:
        return -1;
    }
    if (LAUNCH_OK()) return -4;
    hipLaunchKernelGGL(k_chain_mean, dim3((S + 63) / 64), dim3(64), 0, s, (const double *)partials, parts,
                       (long long)n, off, mean, update, (int)S);
    return LAUNCH_OK();
}

__global__ void k_disp_final(const float *__restrict__ partial, int nfb, int N, int det, long long frames,
                             float *__restrict__ line_db);

// ---- the display reduction from the raw block, unpacking as it goes ----------------------------
// Row positions are counted from frame 0's first sample: [0, dlen) is the
// kept cf32 tail of earlier blocks, [dlen, dlen + n) this block, raw sample i
// at position dlen + i. Frame f owns the block samples of its first hop
// [f fhop, (f+1) fhop); the last frame owns the rest of the block.
// one raw sample at index i (32-bit) of a block, converted exactly as
// k_chain_unpack does
template <int FMT>
__device__ __forceinline__ cpx raw_conv(const unsigned char *raw, unsigned i, float cre, float cim, float k,
                                        float ore, float oim) {
#pragma clang fp contract(off)
    cpx o;
    if constexpr (FMT == FMT_CU8) {
        const uchar2 b = ((const uchar2 *)raw)[i];
        o.x = ((float)b.x - cre) * k;
        o.y = ((float)b.y - cim) * k;
    } else if constexpr (FMT == FMT_CI8) {
        const char2 b = ((const char2 *)raw)[i];
        o.x = ((float)(signed char)b.x - cre) * k;
        o.y = ((float)(signed char)b.y - cim) * k;
    } else if constexpr (FMT == FMT_CI16) {
        const short2 b = ((const short2 *)raw)[i];
        o.x = ((float)b.x - cre) * k;
        o.y = ((float)b.y - cim) * k;
    } else {
        const float2 b = ((const float2 *)raw)[i];
        o.x = b.x - ore;
        o.y = b.y - oim;
    }
    return o;
}

// Bytes of one raw sample of format FMT.
template <int FMT>
__host__ __device__ constexpr int raw_bps() {
    return FMT == FMT_CU8 || FMT == FMT_CI8 ? 2 : FMT == FMT_CF32 ? 8 : 4;
}

// HALF: fhop = N/2, so a frame inside the block (and not the last) owns
// exactly its first PT/2 sample slots
template <int FMT, bool HALF>
struct DispRawSrc {
    cpx *row;
    const unsigned char *raw;
    __amdgpu_buffer_rsrc_t rr;      // the block's raw bytes
    int fhop, dlen, last;
    float cre, cim, kk, ore, oim;
    // fm: k is a power of two and c' is zero or c'k normal, so ((float)b - c')*k
    // is computed as fma((float)b, k, -c'k): c'k is exact, and scaling by a
    // power of two commutes with rounding, so both give the same float (the
    // result is never subnormal: b - c' is 0, or at least 1/2 in magnitude
    // when |c'| < 1/2, or a nonzero multiple of ulp(c') >= 2^-24 otherwise)
    bool fm;
    float dre, dim;
    double sr, si;
    __device__ __forceinline__ cpx conv(int pos) const {
        return raw_conv<FMT>(raw, (unsigned)(pos - dlen), cre, cim, kk, ore, oim);
    }
    // the raw word of sample i0 + r*NT of the block, through the buffer (a
    // lane offset and a constant offset per r: no vector address arithmetic)
    template <int NT>
    __device__ __forceinline__ unsigned bword(unsigned i0, int r) const {
        constexpr int B = raw_bps<FMT>();
        const int vo = (int)(i0 * B), so = r * NT * B;
        if constexpr (B == 2) return __builtin_amdgcn_raw_buffer_load_b16(rr, vo, so, 0);
        else return __builtin_amdgcn_raw_buffer_load_b32(rr, vo, so, 0);
    }
    __device__ __forceinline__ void wparts(unsigned w, float &x, float &y) const {
        if constexpr (FMT == FMT_CU8) {
            x = (float)(w & 0xffu);
            y = (float)((w >> 8) & 0xffu);
        } else if constexpr (FMT == FMT_CI8) {
            x = (float)(signed char)(w & 0xffu);
            y = (float)(signed char)((w >> 8) & 0xffu);
        } else {
            x = (float)(short)(w & 0xffffu);
            y = (float)(short)(w >> 16);
        }
    }
    // the frame's PT samples j + r*NT, all inside the block (sample i0 first)
    template <int PT, int NT>
    __device__ __forceinline__ void bload(unsigned i0, cpx *v) const {
#pragma clang fp contract(off)
        if constexpr (FMT == FMT_CF32) {
#pragma unroll
            for (int r = 0; r < PT; ++r) {
                const cpx b = __builtin_bit_cast(
                    cpx, __builtin_amdgcn_raw_buffer_load_b64(rr, (int)(i0 * 8), r * NT * 8, 0));
                v[r] = cmake(b.x - ore, b.y - oim);
            }
        } else {
            unsigned w[PT];
#pragma unroll
            for (int r = 0; r < PT; ++r) w[r] = bword<NT>(i0, r);
            if (FMT != FMT_CU8 && fm) {
#pragma unroll
                for (int r = 0; r < PT; ++r) {
                    float x, y;
                    wparts(w[r], x, y);
                    v[r] = cmake(__builtin_fmaf(x, kk, -dre), __builtin_fmaf(y, kk, -dim));
                }
            } else {
#pragma unroll
                for (int r = 0; r < PT; ++r) {
                    float x, y;
                    wparts(w[r], x, y);
                    v[r] = cmake((x - cre) * kk, (y - cim) * kk);
                }
            }
        }
    }
    template <int PT, int NT>
    __device__ __forceinline__ void load(int f, int j, cpx *v) const {
        const int p0 = f * fhop + j;
        if (f * fhop >= dlen) {
            // the whole frame is in this block: raw loads only
            const unsigned i0 = (unsigned)(p0 - dlen);
            bload<PT, NT>(i0, v);
        } else {
#pragma unroll
            for (int r = 0; r < PT; ++r) {
                const int pos = p0 + r * NT;
                v[r] = pos < dlen ? row[pos] : conv(pos);
            }
        }
    }
    __device__ __forceinline__ void take(cpx *dst, cpx x) {
        *dst = x;
        sr += (double)x.x;
        si += (double)x.y;
    }
    template <int PT, int NT>
    __device__ __forceinline__ void own(int f, int j, const cpx *v) {
        const int p0 = f * fhop + j;
        if (HALF && f != last && f * fhop >= dlen) {
            cpx *dst = row + (unsigned)p0;
#pragma unroll
            for (int r = 0; r < PT / 2; ++r) take(dst + r * NT, v[r]);
            return;
        }
#pragma unroll
        for (int r = 0; r < PT; ++r) {
            const int p = j + r * NT, pos = p0 + r * NT;
            if (pos >= dlen && (p < fhop || f == last)) take(row + pos, v[r]);
        }
    }
};

struct DispRawArgs {
    cpx *row0;                  // stream s frame 0 at row0 + s*spitch
    long long spitch;
    const unsigned char *raw;   // stream s block at raw + s*raw_pitch
    long long raw_pitch;
    int dlen, n;                // kept tail before the block; block samples
    const float *off;           // per stream offset (re, im); null = 0
    float c0, kk;
    long long frames, fhop;
    int fpb, nfb;
    const float *win;
    const cpx *tw;
    float *partial;             // [S][nfb][N] display order (fftshifted)
    double *mparts;             // [S][nfb][2] FP64 sums of the samples each block wrote
};

template <int N, int DET, int FMT, bool HALF>
__global__ void __launch_bounds__(256) k_disp_unpack(DispRawArgs a) {
#pragma clang fp contract(off)
    using D = DispShape<N>;
    __shared__ cpx lds[D::TEAMS * D::TL];
    __shared__ float red[D::TEAMS > 1 ? D::TEAMS * N : 1];
    __shared__ cpx twl[N];
    const int s = blockIdx.y, fb = blockIdx.x;
    for (int i = threadIdx.x; i < N; i += 256) twl[i] = a.tw[i];
    __syncthreads();
    DispRawSrc<FMT, HALF> src;
    src.row = a.row0 + (long long)s * a.spitch;
    src.raw = a.raw + (long long)s * a.raw_pitch;
    src.fhop = (int)a.fhop;
    src.dlen = a.dlen;
    src.last = (int)a.frames - 1;
    src.kk = a.kk;
    src.ore = a.off ? a.off[2 * s] : 0.0f;
    src.oim = a.off ? a.off[2 * s + 1] : 0.0f;
    // the folded constants exactly as src/convert.c:47, :57, :68 (IEEE division)
    src.cre = a.c0 + src.ore / a.kk;
    src.cim = a.c0 + src.oim / a.kk;
    src.rr = __builtin_amdgcn_make_buffer_rsrc((void *)src.raw, (short)0, a.n * raw_bps<FMT>(), 0x00020000);
    {
#pragma clang fp contract(off)
        const unsigned kb = __builtin_bit_cast(unsigned, a.kk);
        const bool pow2 = (kb & 0x807fffffu) == 0 && (kb >> 23) != 0 && (kb >> 23) != 255;
        src.dre = src.cre * a.kk;
        src.dim = src.cim * a.kk;
        // c' = 0 exactly, or c'k normal (a c'k that underflowed is not exact)
        auto ok = [](float c, float d) { return c == 0.0f || __builtin_fabsf(d) >= 1.17549435e-38f; };
        src.fm = pow2 && ok(src.cre, src.dre) && ok(src.cim, src.dim);
    }
    src.sr = 0.0;
    src.si = 0.0;
    const int f0 = fb * a.fpb;
    const int f1 = min(f0 + a.fpb, (int)a.frames);
    disp_frames<N, DET, 256>(src, f0, f1, a.win, twl, lds, red, a.partial + ((long long)s * a.nfb + fb) * N);
    if (f1 == (int)a.frames) {
        // the block's samples past the last frame's window
        for (int pos = src.last * src.fhop + N + (int)threadIdx.x; pos < a.dlen + a.n; pos += 256) {
            const cpx x = src.conv(pos);
            src.row[pos] = x;
            src.sr += (double)x.x;
            src.si += (double)x.y;
        }
    }
    // this block's FP64 sums, in a fixed order
    __syncthreads();
    double *sh = (double *)lds;
    sh[threadIdx.x] = src.sr;
    sh[256 + threadIdx.x] = src.si;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sh[threadIdx.x] += sh[threadIdx.x + w];
            sh[256 + threadIdx.x] += sh[256 + threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double *mp = a.mparts + ((long long)s * a.nfb + fb) * 2;
        mp[0] = sh[0];
        mp[1] = sh[256];
    }
}

template <int N, int FMT, bool HALF>
static int launch_disp_unpack_h(const DispRawArgs &a, int det, unsigned S, hipStream_t st) {
    const dim3 g((unsigned)a.nfb, S);
    switch (det) {
    case 0: hipLaunchKernelGGL((k_disp_unpack<N, 0, FMT, HALF>), g, dim3(256), 0, st, a); break;
    case 1: hipLaunchKernelGGL((k_disp_unpack<N, 1, FMT, HALF>), g, dim3(256), 0, st, a); break;
    default: hipLaunchKernelGGL((k_disp_unpack<N, 2, FMT, HALF>), g, dim3(256), 0, st, a); break;
    }
    return LAUNCH_OK();
}

template <int N, int FMT>
static int launch_disp_unpack(const DispRawArgs &a, int det, unsigned S, hipStream_t st) {
    return a.fhop * 2 == N ? launch_disp_unpack_h<N, FMT, true>(a, det, S, st)
                           : launch_disp_unpack_h<N, FMT, false>(a, det, S, st);
}

template <int N>
static int launch_disp_unpack_fmt(const DispRawArgs &a, int fmt, int det, unsigned S, hipStream_t st) {
    switch (fmt) {
    case FMT_CU8: return launch_disp_unpack<N, FMT_CU8>(a, det, S, st);
    case FMT_CI8: return launch_disp_unpack<N, FMT_CI8>(a, det, S, st);
    case FMT_CI16:
    case FMT_CI16Q11: return launch_disp_unpack<N, FMT_CI16>(a, det, S, st);
    case FMT_CF32: return launch_disp_unpack<N, FMT_CF32>(a, det, S, st);
    }
    return -1;
}

// The display reduction of `frames` frames straight from the raw blocks,
// unpacking each block into its row on the way (frames >= 1), then the dB
// line, and the per-stream mean -> next offset (update != 0) from the FP64
// sums the display blocks formed. *nfb_used receives the frame blocks per
// stream (the partial line and mean rows used).
extern "C" int atkgpu_chain_display_unpack(void *st, void *row0, size_t spitch, const void *raw, size_t raw_pitch,
                                           int fmt, size_t n, size_t dlen, float *off, unsigned S, size_t frames,
                                           size_t fhop, unsigned N, int nfb, const float *win, const void *tw,
                                           int det, float *partial, float *line_db, double *mparts, double *mean,
                                           int update) {
    if (!frames || !S || !n) return -1;
    DispRawArgs a;
    a.row0 = (cpx *)row0;
    a.spitch = (long long)spitch;
    a.raw = (const unsigned char *)raw;
    a.raw_pitch = (long long)raw_pitch;
    a.dlen = (int)dlen;
    a.n = (int)n;
    a.off = off;
    a.c0 = 0.0f;
    a.kk = 1.0f;
    switch (fmt) {
    case FMT_CU8: a.c0 = 127.5f; a.kk = 1.0f / 127.5f; break;
    case FMT_CI8: a.kk = 1.0f / 128.0f; break;
    case FMT_CI16: a.kk = 1.0f / 32768.0f; break;
    case FMT_CI16Q11: a.kk = 1.0f / 2048.0f; break;
    case FMT_CF32: break;
    default: return -1;
    }
    a.frames = (long long)frames;
    a.fhop = (long long)fhop;
    a.fpb = (int)((frames + nfb - 1) / nfb);
    a.nfb = (int)((frames + a.fpb - 1) / a.fpb);
    a.win = win;
    a.tw = (const cpx *)tw;
    a.partial = partial;
    a.mparts = mparts;
    hipStream_t s = (hipStream_t)st;
    int rc;
    switch (N) {
    case 1024: rc = launch_disp_unpack_fmt<1024>(a, fmt, det, S, s); break;
    case 2048: rc = launch_disp_unpack_fmt<2048>(a, fmt, det, S, s); break;
    case 4096: rc = launch_disp_unpack_fmt<4096>(a, fmt, det, S, s); break;
    default: return -1;
    }
    if (rc) return rc;
    hipLaunchKernelGGL(k_disp_final, dim3((N + 255) / 256, S), dim3(256), 0, s, (const float *)partial, a.nfb,
                       (int)N, det, (long long)frames, line_db);
    if (LAUNCH_OK()) return -4;
    hipLaunchKernelGGL(k_chain_mean, dim3((S + 63) / 64), dim3(64), 0, s, (const double *)mparts, a.nfb,
                       (long long)n, off, mean, update, (int)S);
    return LAUNCH_OK();
}

// per (stream, bin): combine the frame blocks in order, then dB (reduce's
// 20 log10(sqrt(P) + 1e-12), src/fft.c:286-292)
__global__ void k_disp_final(const float *__restrict__ partial, int nfb, int N, int det, long long frames,
                             float *__restrict__ line_db) {
#pragma clang fp contract(off)
    const int s = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float *p = partial + (long long)s * nfb * N + i;
    float v = p[0];
    for (int b = 1; b < nfb; ++b) {
        const float u = p[(long long)b * N];
        if (det == 0) { if (u > v) v = u; }
        else if (det == 1) { if (u < v) v = u; }
        else v += u;
    }
    const float scale = det == 2 ? 1.0f / (float)frames : 1.0f;
    line_db[(long long)s * N + i] = 20.0f * log10f(sqrtf(v * scale) + 1e-12f);
}

// per stream: decimate the dB line to `cols` columns, then stitch those into
// the stream's trace (NaN = unmeasured)
__global__ void __launch_bounds__(256)
k_disp_post(const float *__restrict__ line_db, int N, int cols, float *__restrict__ dec, float *__restrict__ trace,
            int nout, double seg_lo0, double seg_lo_step, double seg_step, double out_lo, double out_step) {
#pragma clang fp contract(off)
    const int s = blockIdx.x;
    const float *in = line_db + (long long)s * N;
    float *d = dec + (long long)s * cols;
    disp_decimate(in, N, cols, d);
    if (!trace || nout <= 0) return;
    __syncthreads();
    // stitch: segment bin i at seg_lo + (i + 0.5) seg_step, kept over the whole
    // segment span; runs that land on one output bin are applied in order by
    // their first bin's thread
    const double seg_lo = seg_lo0 + seg_lo_step * s;
    const double keep_lo = seg_lo, keep_hi = seg_lo + seg_step * cols;
    const double out_hi = out_lo + out_step * (double)nout;
    float *t = trace + (long long)s * nout;
    auto kmap = [&](int i, long long *k) -> bool {
        const double f = seg_lo + ((double)i + 0.5) * seg_step;
        if (f < keep_lo || f > keep_hi || f < out_lo || f > out_hi) return false;
        long long kk = (long long)((f - out_lo) / out_step);
        if (kk < 0) kk = 0;
        if (kk >= nout) kk = nout - 1;
        *k = kk;
        return true;
    };
    for (int i = threadIdx.x; i < cols; i += 256) {
        long long k, kp;
        if (!kmap(i, &k)) continue;
        if (i > 0 && kmap(i - 1, &kp) && kp == k) continue;
        float cur = t[k];
        for (int q = i; q < cols; ++q) {
            long long kq;
            if (!kmap(q, &kq) || kq != k) break;
            const float v = d[q];
            if (cur != cur || v > cur) cur = v;
        }
        t[k] = cur;
    }
}

extern "C" int atkgpu_chain_post(void *st, const float *line_db, unsigned S, unsigned N, unsigned cols, float *dec,
                                 float *trace, unsigned nout, double seg_lo0, double seg_lo_step, double seg_step,
                                 double out_lo, double out_step) {
    if (!S) return 0;
    hipLaunchKernelGGL(k_disp_post, dim3(S), dim3(256), 0, (hipStream_t)st, line_db, (int)N, (int)cols, dec, trace,
                       (int)nout, seg_lo0, seg_lo_step, seg_step, out_lo, out_step);
    return LAUNCH_OK();
}

// ---- keep the last `keep` samples of every stream row ------------------------------------
__global__ void k_chain_keep(const cpx *__restrict__ src, long long pitch, cpx *__restrict__ dst, long long n) {
    const int s = blockIdx.y;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        dst[(long long)s * pitch + i] = src[(long long)s * pitch + i];
}

// src/dst: stream 0's first sample to move / destination, rows spitch apart
extern "C" int atkgpu_chain_keep(void *st, const void *src, size_t spitch, void *dst, size_t keep, unsigned S) {
    if (!keep || !S) return 0;
    const unsigned gx = grid_for(keep, 256, 64);
    hipLaunchKernelGGL(k_chain_keep, dim3(gx, S), dim3(256), 0, (hipStream_t)st, (const cpx *)src,
                       (long long)spitch, (cpx *)dst, (long long)keep);
    return LAUNCH_OK();
}
