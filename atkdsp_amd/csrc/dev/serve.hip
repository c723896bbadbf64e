// Ragged-batch kernels for the serving loop: every job is one chunk of one
// independent stream (its own length, format constants, buffers). A block
// finds its job by binary search over the jobs' first-block indices.
//   k_srv_unpack  raw -> cf32 with the stream's fixed offset folded into the
//                 per-format constants (bit-exact unpack_dc, reference
//                 src/convert.c:33-111); raw bytes may sit in pinned host
//                 memory mapped for the device (read over PCIe, no copy call)
//   k_srv_copy    moves the kept prefix of each stream to its other row
//   k_srv_disp    display lines of the streams that have them
#include <hip/hip_runtime.h>
#include "atkgpu_dev.h"
#include "dev_common.hpp"
#include "disp.hpp"

template <typename J>
__device__ __forceinline__ int find_job(const J *jobs, int njobs, int b) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].blk0 <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}

struct f2 { float x, y; };

// One format's share of a ragged unpack block: samples [i0, i1) of job jb.
// Whole 16-byte words of raw bytes are loaded at once (a PCIe read of mapped
// host memory moves 1 KiB per wavefront load), the ragged end sample by
// sample. The arithmetic is the reference's per sample, contraction off.
template <int FMT>
__device__ __forceinline__ void srv_unpack_range(const atkgpu_unpack_job &jb, int i0, int i1) {
#pragma clang fp contract(off)
    constexpr int BPS = (FMT == FMT_CU8 || FMT == FMT_CI8) ? 2 : (FMT == FMT_CF32 ? 8 : 4);
    constexpr int SPW = 16 / BPS;                 // samples per 16-byte word
    float c0, kk;
    if constexpr (FMT == FMT_CU8) { c0 = 127.5f; kk = 1.0f / 127.5f; }
    else if constexpr (FMT == FMT_CI8) { c0 = 0.0f; kk = 1.0f / 128.0f; }
    else if constexpr (FMT == FMT_CI16) { c0 = 0.0f; kk = 1.0f / 32768.0f; }
    else if constexpr (FMT == FMT_CI16Q11) { c0 = 0.0f; kk = 1.0f / 2048.0f; }
    else { c0 = 0.0f; kk = 1.0f; }
    // the folded constants as src/convert.c:47, :57, :68 form them
    const float cre = c0 + jb.off_re / kk, cim = c0 + jb.off_im / kk;
    const bool plain = jb.off_re == 0.0f && jb.off_im == 0.0f;
    const unsigned char *raw = (const unsigned char *)jb.raw;
    f2 *dst = (f2 *)jb.dst;
    auto conv = [&](unsigned re_bits, unsigned im_bits) -> f2 {
        f2 o;
        if constexpr (FMT == FMT_CU8) {
            o.x = ((float)(re_bits & 0xffu) - cre) * kk;
            o.y = ((float)(im_bits & 0xffu) - cim) * kk;
        } else if constexpr (FMT == FMT_CI8) {
            o.x = ((float)(signed char)(re_bits & 0xffu) - cre) * kk;
            o.y = ((float)(signed char)(im_bits & 0xffu) - cim) * kk;
        } else if constexpr (FMT == FMT_CI16 || FMT == FMT_CI16Q11) {
            o.x = ((float)(short)(re_bits & 0xffffu) - cre) * kk;
            o.y = ((float)(short)(im_bits & 0xffffu) - cim) * kk;
        } else {
            const float a = __uint_as_float(re_bits), b = __uint_as_float(im_bits);
            if (plain) { o.x = a; o.y = b; }
            else { o.x = a - jb.off_re; o.y = b - jb.off_im; }
        }
        return o;
    };
    const int nw = (i1 - i0) / SPW;
    for (int w = (int)threadIdx.x; w < nw; w += 256) {
        const int i = i0 + w * SPW;
        const uint4 v = *(const uint4 *)(raw + (size_t)i * BPS);
        const unsigned u[4] = {v.x, v.y, v.z, v.w};
        f2 o[SPW];
#pragma unroll
        for (int k = 0; k < SPW; ++k) {
            if constexpr (BPS == 2) {           // two samples per 32-bit word
                const unsigned x = u[k >> 1] >> (16 * (k & 1));
                o[k] = conv(x, x >> 8);
            } else if constexpr (BPS == 4) {
                o[k] = conv(u[k], u[k] >> 16);
            } else {
                o[k] = conv(u[2 * k], u[2 * k + 1]);
            }
        }
        float4 *d4 = (float4 *)(dst + i);
#pragma unroll
        for (int k = 0; k < SPW / 2; ++k) d4[k] = make_float4(o[2 * k].x, o[2 * k].y, o[2 * k + 1].x, o[2 * k + 1].y);
    }
    for (int i = i0 + nw * SPW + (int)threadIdx.x; i < i1; i += 256) {
        const unsigned char *p = raw + (size_t)i * BPS;
        if constexpr (BPS == 2) dst[i] = conv(p[0], p[1]);
        else if constexpr (BPS == 4) dst[i] = conv(*(const unsigned short *)p, *(const unsigned short *)(p + 2));
        else dst[i] = conv(*(const unsigned *)p, *(const unsigned *)(p + 4));
    }
}

// Every format in one launch, and the copies of kept prefixes beside them: a
// block belongs to one job, so the switch is uniform over the block.
// The grid may be smaller than the job blocks: a block then takes every
// gridDim.x-th of them, which bounds the PCIe reads in flight (enough for
// the link's bandwidth; many more only queue ahead of other kernels' HBM
// traffic).
// With per_block set, entry b is block b's own (its tile in blk0): one read
// of the descriptor memory instead of a search through it.
__global__ void __launch_bounds__(256) k_srv_unpack(const atkgpu_unpack_job *__restrict__ jobs, int njobs,
                                                    int nblocks, int per_block) {
  for (int b = blockIdx.x; b < nblocks; b += gridDim.x) {
    const atkgpu_unpack_job jb = jobs[per_block ? b : find_job(jobs, njobs, b)];
    const int i0 = (per_block ? jb.blk0 : b - jb.blk0) * ATKGPU_RAGGED_SPB;
    const int i1 = min(i0 + ATKGPU_RAGGED_SPB, jb.n);
    switch (jb.fmt) {
    case ATKGPU_JOB_COPY: {
        const f2 *src = (const f2 *)jb.raw;
        f2 *dst = (f2 *)jb.dst;
        for (int i = i0 + (int)threadIdx.x; i < i1; i += 256) dst[i] = src[i];
        break;
    }
    case FMT_CU8: srv_unpack_range<FMT_CU8>(jb, i0, i1); break;
    case FMT_CI8: srv_unpack_range<FMT_CI8>(jb, i0, i1); break;
    case FMT_CI16: srv_unpack_range<FMT_CI16>(jb, i0, i1); break;
    case FMT_CI16Q11: srv_unpack_range<FMT_CI16Q11>(jb, i0, i1); break;
    default: srv_unpack_range<FMT_CF32>(jb, i0, i1); break;
    }
  }
}

__global__ void __launch_bounds__(256) k_srv_copy(const atkgpu_copy_job *__restrict__ jobs, int njobs) {
    const int b = blockIdx.x;
    const atkgpu_copy_job jb = jobs[find_job(jobs, njobs, b)];
    const f2 *src = (const f2 *)jb.src;
    f2 *dst = (f2 *)jb.dst;
    const int i0 = (b - jb.blk0) * ATKGPU_RAGGED_SPB;
    const int i1 = min(i0 + ATKGPU_RAGGED_SPB, jb.n);
    for (int i = i0 + (int)threadIdx.x; i < i1; i += 256) dst[i] = src[i];
}

extern "C" int atkgpu_unpack_jobs(void *s, const void *jobs, int njobs, int nblocks, int per_block,
                                  int max_grid) {
    if (!njobs || !nblocks) return 0;
    const int grid = (max_grid > 0 && max_grid < nblocks) ? max_grid : nblocks;
    hipLaunchKernelGGL(k_srv_unpack, dim3(grid), dim3(256), 0, (hipStream_t)s,
                       (const atkgpu_unpack_job *)jobs, njobs, nblocks, per_block);
    return LAUNCH_OK();
}

extern "C" int atkgpu_copy_jobs(void *s, const void *jobs, int njobs, int nblocks) {
    if (!njobs || !nblocks) return 0;
    hipLaunchKernelGGL(k_srv_copy, dim3(nblocks), dim3(256), 0, (hipStream_t)s, (const atkgpu_copy_job *)jobs,
                       njobs);
    return LAUNCH_OK();
}

extern "C" void *atkgpu_hmalloc_mapped(size_t bytes, void **dev_ptr) {
    void *p = nullptr;
    if (bytes == 0) bytes = 256;
    if (hipHostMalloc(&p, bytes, hipHostMallocMapped) != hipSuccess) return nullptr;
    void *d = nullptr;
    if (hipHostGetDevicePointer(&d, p, 0) != hipSuccess) {
        (void)hipHostFree(p);
        return nullptr;
    }
    *dev_ptr = d;
    return p;
}

// ---- served display lines -----------------------------------------------------------
//   k_srv_frames  one FFT team per frame: window, N-point team FFT, |X|^2/n^2
//                 in float (reference src/fft.c:177-183), written in display
//                 order (fftshifted, src/fft.c:190)
//   k_srv_lines   one block per line segment: the detector over the segment's
//                 spectra in frame order after the line's running value
//                 (src/fft.c:184-189), then the running line, or the finished
//                 line's dB (src/fft.c:286-292) through decimate_max
//                 (src/display.c:19-37)
// Both launches are wide (every frame, every segment of the batch at once)
// and no block waits for another.
template <int N>
__global__ void __launch_bounds__(256) k_srv_frames(const atkgpu_frame_blk *__restrict__ blks, const cpx *tw) {
#pragma clang fp contract(off)
    using D = DispShape<N, 256>;
    constexpr int PT = D::PT, NT = D::NT, TL = D::TL;
    __shared__ cpx lds[D::TEAMS * TL];
    __shared__ cpx twl[N];
    const atkgpu_frame_blk jb = blks[blockIdx.x];
    for (int i = threadIdx.x; i < N; i += 256) twl[i] = tw[i];
    const int tm = threadIdx.x / NT, j = threadIdx.x % NT;
    const bool act = tm < jb.nfr;
    const cpx *fr = (const cpx *)jb.first + (act ? tm : 0) * jb.fhop + j;
    cpx v[PT];
#pragma unroll
    for (int r = 0; r < PT; ++r) {
        const float w = jb.win ? jb.win[j + r * NT] : 1.0f;
        const cpx x = fr[r * NT];
        v[r] = act ? cmake(x.x * w, x.y * w) : cmake(0.f, 0.f);
    }
    __syncthreads();
    team_fft<N, PT>(v, lds + tm * TL, twl, j);
    if (!act) return;
    const float inv2 = 1.0f / ((float)N * (float)N);
    float *P = jb.P + (long long)tm * N;
    // FFT bin k = j + r*NT sits at display bin (k + N/2) mod N
#pragma unroll
    for (int r = 0; r < PT; ++r) P[j + ((r + PT / 2) % PT) * NT] = (v[r].x * v[r].x + v[r].y * v[r].y) * inv2;
}

template <int N, int DET>
__global__ void __launch_bounds__(256) k_srv_lines(const atkgpu_line_blk *__restrict__ blks) {
#pragma clang fp contract(off)
    constexpr int NB = N / 256;                // bins a thread: k = threadIdx.x + i*256
    __shared__ float line[N];
    const atkgpu_line_blk jb = blks[blockIdx.x];
    float v[NB];
    int f = 0;
    if (jb.acc_in) {
#pragma unroll
        for (int i = 0; i < NB; ++i) v[i] = jb.acc_in[threadIdx.x + i * 256];
    } else {
#pragma unroll
        for (int i = 0; i < NB; ++i) v[i] = jb.P[threadIdx.x + i * 256];
        f = 1;
    }
    // four frames' loads in flight at a time, merged in frame order
    for (; f + 4 <= jb.nfr; f += 4) {
        const float *p = jb.P + (long long)f * N + threadIdx.x;
        float u[4][NB];
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int i = 0; i < NB; ++i) u[q][i] = p[(long long)q * N + i * 256];
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int i = 0; i < NB; ++i) v[i] = det_merge<DET>(v[i], u[q][i]);
    }
    for (; f < jb.nfr; ++f) {
        const float *p = jb.P + (long long)f * N;
#pragma unroll
        for (int i = 0; i < NB; ++i) v[i] = det_merge<DET>(v[i], p[threadIdx.x + i * 256]);
    }
    if (jb.acc_out) {
#pragma unroll
        for (int i = 0; i < NB; ++i) jb.acc_out[threadIdx.x + i * 256] = v[i];
        return;
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) line[threadIdx.x + i * 256] = disp_db(v[i], jb.scale);
    __syncthreads();
    disp_decimate(line, N, jb.cols, jb.line);
}

extern "C" int atkgpu_disp_teams(unsigned N) {
    switch (N) {
    case 1024: return DispShape<1024, 256>::TEAMS;
    case 2048: return DispShape<2048, 256>::TEAMS;
    case 4096: return DispShape<4096, 256>::TEAMS;
    }
    return 0;
}

extern "C" int atkgpu_disp_frames(void *s, const void *blks, int nblk, unsigned N, const void *tw) {
    if (nblk <= 0) return 0;
    const atkgpu_frame_blk *b = (const atkgpu_frame_blk *)blks;
    const dim3 g(nblk), t(256);
    switch (N) {
    case 1024: hipLaunchKernelGGL(k_srv_frames<1024>, g, t, 0, (hipStream_t)s, b, (const cpx *)tw); break;
    case 2048: hipLaunchKernelGGL(k_srv_frames<2048>, g, t, 0, (hipStream_t)s, b, (const cpx *)tw); break;
    case 4096: hipLaunchKernelGGL(k_srv_frames<4096>, g, t, 0, (hipStream_t)s, b, (const cpx *)tw); break;
    default: return -1;
    }
    return LAUNCH_OK();
}

template <int N>
static int launch_lines(const atkgpu_line_blk *b, int nblk, int det, hipStream_t s) {
    switch (det) {
    case 0: hipLaunchKernelGGL((k_srv_lines<N, 0>), dim3(nblk), dim3(256), 0, s, b); break;
    case 1: hipLaunchKernelGGL((k_srv_lines<N, 1>), dim3(nblk), dim3(256), 0, s, b); break;
    case 2: hipLaunchKernelGGL((k_srv_lines<N, 2>), dim3(nblk), dim3(256), 0, s, b); break;
    default: return -1;
    }
    return LAUNCH_OK();
}

extern "C" int atkgpu_disp_lines(void *s, const void *blks, int nblk, unsigned N, int det) {
    if (nblk <= 0) return 0;
    const atkgpu_line_blk *b = (const atkgpu_line_blk *)blks;
    switch (N) {
    case 1024: return launch_lines<1024>(b, nblk, det, (hipStream_t)s);
    case 2048: return launch_lines<2048>(b, nblk, det, (hipStream_t)s);
    case 4096: return launch_lines<4096>(b, nblk, det, (hipStream_t)s);
    }
    return -1;
}
