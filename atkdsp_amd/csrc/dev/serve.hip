// Ragged-batch kernels for the serving loop: every job is one chunk of one
// independent stream (its own length, format constants, buffers). A block
// finds its job by binary search over the jobs' first-block indices.
//   k_srv_unpack  raw -> cf32 with the stream's fixed offset folded into the
//                 per-format constants (bit-exact unpack_dc, reference
//                 src/convert.c:33-111); raw bytes may sit in pinned host
//                 memory mapped for the device (read over PCIe, no copy call)
//   k_srv_copy    moves the kept prefix of each stream to its other row
#include <hip/hip_runtime.h>
#include "atkgpu_dev.h"
#include "dev_common.hpp"

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

template <int FMT>
__global__ void __launch_bounds__(256) k_srv_unpack(const atkgpu_unpack_job *__restrict__ jobs, int njobs) {
#pragma clang fp contract(off)
    const int b = blockIdx.x;
    const atkgpu_unpack_job jb = jobs[find_job(jobs, njobs, b)];
    float c0, kk;
    if constexpr (FMT == FMT_CU8) { c0 = 127.5f; kk = 1.0f / 127.5f; }
    else if constexpr (FMT == FMT_CI8) { c0 = 0.0f; kk = 1.0f / 128.0f; }
    else if constexpr (FMT == FMT_CI16) { c0 = 0.0f; kk = 1.0f / 32768.0f; }
    else if constexpr (FMT == FMT_CI16Q11) { c0 = 0.0f; kk = 1.0f / 2048.0f; }
    else { c0 = 0.0f; kk = 1.0f; }
    // the folded constants as src/convert.c:47, :57, :68 form them
    const float cre = c0 + jb.off_re / kk, cim = c0 + jb.off_im / kk;
    f2 *dst = (f2 *)jb.dst;
    const int i0 = (b - jb.blk0) * ATKGPU_RAGGED_SPB;
    const int i1 = min(i0 + ATKGPU_RAGGED_SPB, jb.n);
    for (int i = i0 + (int)threadIdx.x; i < i1; i += 256) {
        f2 o;
        if constexpr (FMT == FMT_CU8) {
            const uchar2 v = ((const uchar2 *)jb.raw)[i];
            o.x = ((float)v.x - cre) * kk;
            o.y = ((float)v.y - cim) * kk;
        } else if constexpr (FMT == FMT_CI8) {
            const char2 v = ((const char2 *)jb.raw)[i];
            o.x = ((float)(signed char)v.x - cre) * kk;
            o.y = ((float)(signed char)v.y - cim) * kk;
        } else if constexpr (FMT == FMT_CI16 || FMT == FMT_CI16Q11) {
            const short2 v = ((const short2 *)jb.raw)[i];
            o.x = ((float)v.x - cre) * kk;
            o.y = ((float)v.y - cim) * kk;
        } else {
            const float2 v = ((const float2 *)jb.raw)[i];
            if (jb.off_re == 0.0f && jb.off_im == 0.0f) { o.x = v.x; o.y = v.y; }
            else { o.x = v.x - jb.off_re; o.y = v.y - jb.off_im; }
        }
        dst[i] = o;
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

extern "C" int atkgpu_unpack_jobs(void *s, int fmt, const void *jobs, int njobs, int nblocks) {
    if (!njobs || !nblocks) return 0;
    const atkgpu_unpack_job *j = (const atkgpu_unpack_job *)jobs;
    hipStream_t st = (hipStream_t)s;
    switch (fmt) {
    case FMT_CU8: hipLaunchKernelGGL(k_srv_unpack<FMT_CU8>, dim3(nblocks), dim3(256), 0, st, j, njobs); break;
    case FMT_CI8: hipLaunchKernelGGL(k_srv_unpack<FMT_CI8>, dim3(nblocks), dim3(256), 0, st, j, njobs); break;
    case FMT_CI16: hipLaunchKernelGGL(k_srv_unpack<FMT_CI16>, dim3(nblocks), dim3(256), 0, st, j, njobs); break;
    case FMT_CI16Q11:
        hipLaunchKernelGGL(k_srv_unpack<FMT_CI16Q11>, dim3(nblocks), dim3(256), 0, st, j, njobs);
        break;
    case FMT_CF32: hipLaunchKernelGGL(k_srv_unpack<FMT_CF32>, dim3(nblocks), dim3(256), 0, st, j, njobs); break;
    default: return -1;
    }
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
