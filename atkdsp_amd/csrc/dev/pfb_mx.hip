// Launchers of the filterbank's matrix-core form (pfb_mx.hpp) and its set-up:
// the per-column A operands of the tap matrix, and the raw-history copy.
#include "atkgpu_dev.h"
#include "pfb_mx.hpp"

namespace {

template <int M, int OS, int KOS, int FMT, int ELT, int NTERM>
int launch(const PfbMxArgs &a, unsigned grid, hipStream_t s) {
    hipLaunchKernelGGL((k_pfb_mx<M, OS, KOS, FMT, ELT, NTERM>), dim3(grid), dim3(MxShape<M, OS, KOS>::NTHR), 0, s,
                       a);
    return LAUNCH_OK();
}

// form: 1 = BF16 hi+lo, 2 = BF16, 3 = FP16 hi+lo, 4 = FP16 (ATKGPU_FORM_*).
// The 64-channel critically sampled shape has every (format, form) pair;
// the other shapes have BF16 hi+lo for 8-bit and FP16 hi+lo for 12-bit samples.
template <int M, int OS, int KOS, int FMT, bool FULL>
int by_form(const PfbMxArgs &a, int form, unsigned grid, hipStream_t s) {
    constexpr bool B8 = FMT == FMT_CU8 || FMT == FMT_CI8;     // 8-bit samples are exact in BF16
    if constexpr (B8) {
        if (form == 1) return launch<M, OS, KOS, FMT, MX_BF16, 2>(a, grid, s);
        if constexpr (FULL) {
            if (form == 2) return launch<M, OS, KOS, FMT, MX_BF16, 1>(a, grid, s);
            if (form == 3) return launch<M, OS, KOS, FMT, MX_F16, 2>(a, grid, s);
            if (form == 4 && FMT == FMT_CU8) return launch<M, OS, KOS, FMT, MX_F16, 1>(a, grid, s);
        }
    } else {
        if (form == 3) return launch<M, OS, KOS, FMT, MX_F16, 2>(a, grid, s);
        if constexpr (FULL)
            if (form == 4) return launch<M, OS, KOS, FMT, MX_F16, 1>(a, grid, s);
    }
    return -1;
}

template <int M, int OS, int KOS, bool FULL>
int by_fmt(const PfbMxArgs &a, int fmt, int form, unsigned grid, hipStream_t s) {
    switch (fmt) {
    case FMT_CU8: return by_form<M, OS, KOS, FMT_CU8, FULL>(a, form, grid, s);
    case FMT_CI8: if constexpr (FULL) return by_form<M, OS, KOS, FMT_CI8, FULL>(a, form, grid, s); else return -1;
    case FMT_CI16Q11: return by_form<M, OS, KOS, FMT_CI16Q11, FULL>(a, form, grid, s);
    }
    return -1;
}

int dispatch(const PfbMxArgs &a, unsigned M, unsigned os, unsigned kos, int fmt, int form, unsigned grid,
             hipStream_t s) {
    switch (M * 1000 + os * 100 + kos) {
    case 64108: return by_fmt<64, 1, 8, true>(a, fmt, form, grid, s);
    case 64116: return by_fmt<64, 1, 16, false>(a, fmt, form, grid, s);
    case 64132: return by_fmt<64, 1, 32, false>(a, fmt, form, grid, s);
    case 64216: return by_fmt<64, 2, 16, false>(a, fmt, form, grid, s);
    case 128216: return by_fmt<128, 2, 16, false>(a, fmt, form, grid, s);
    }
    return -1;
}

int group_of(unsigned M, unsigned os, unsigned kos) {
    switch (M * 1000 + os * 100 + kos) {
    case 64108: return MxShape<64, 1, 8>::G;
    case 64116: return MxShape<64, 1, 16>::G;
    case 64132: return MxShape<64, 1, 32>::G;
    case 64216: return MxShape<64, 2, 16>::G;
    case 128216: return MxShape<128, 2, 16>::G;
    }
    return 0;
}

int ksteps_of(unsigned os, unsigned kos) { return (int)((16 / os + kos - 1 + 31) / 32); }

}  // namespace

extern "C" int atkgpu_pfb_mx_supported(unsigned M, unsigned os, unsigned kos, int fmt, int form) {
    if (!group_of(M, os, kos)) return 0;
    const bool full = M == 64 && os == 1 && kos == 8;
    if (fmt == FMT_CU8) return form == 1 || (full && form >= 2 && form <= 4);
    if (fmt == FMT_CI8) return full && (form == 1 || form == 2 || form == 3);
    if (fmt == FMT_CI16Q11) return form == 3 || (full && form == 4);   // 12-bit samples need FP16
    return 0;
}

extern "C" size_t atkgpu_pfb_mx_tpad_count(unsigned M, unsigned os, unsigned kos, int form) {
    const int nterm = (form == 1 || form == 3) ? 2 : 1;
    return (size_t)(M / os) * nterm * os * (32 * ksteps_of(os, kos) + 16);
}

extern "C" int atkgpu_pfb_mx_tpad(void *s, const float *taps, unsigned M, unsigned os, unsigned kos, int form,
                                  float tscale, void *tpad) {
    const int hop = (int)(M / os), tpl = 32 * ksteps_of(os, kos) + 16, nterm = (form == 1 || form == 3) ? 2 : 1;
    const int total = hop * nterm * (int)os * tpl;
    const int blocks = (total + 255) / 256;
    if (form == 1 || form == 2)
        hipLaunchKernelGGL(k_mx_tpad<MX_BF16>, dim3(blocks), dim3(256), 0, (hipStream_t)s, taps, hop, (int)kos,
                           (int)os, tpl, nterm, tscale, (uint16_t *)tpad);
    else
        hipLaunchKernelGGL(k_mx_tpad<MX_F16>, dim3(blocks), dim3(256), 0, (hipStream_t)s, taps, hop, (int)kos,
                           (int)os, tpl, nterm, tscale, (uint16_t *)tpad);
    return LAUNCH_OK();
}

extern "C" int atkgpu_pfb_mx(void *s, const void *raw, size_t spitch, size_t nvalid, unsigned nel, size_t m,
                             long long b0, unsigned M, unsigned os, unsigned kos, int fmt, int form,
                             const void *tpad, const float *tapsum, const float *taps, int zlo, float scale,
                             const void *offs, const void *tw, const void *cal, unsigned cal_nel, void *out,
                             size_t opitch) {
    if (!m || !nel) return 0;
    if (!atkgpu_pfb_mx_supported(M, os, kos, fmt, form)) return -1;
    const int G = group_of(M, os, kos);
    PfbMxArgs a;
    a.raw = (const unsigned char *)raw;
    a.spitch = (long long)spitch;
    a.nvalid = (long long)nvalid;
    a.m = (long long)m;
    a.b0 = b0;
    a.tpad = (const uint16_t *)tpad;
    a.tapsum = tapsum;
    a.taps = taps;
    a.zlo = zlo;
    a.scale = scale;
    a.offs = (const float2 *)offs;
    a.tw = (const cpx *)tw;
    a.cal = (const cpx *)cal;
    a.cal_nel = (int)cal_nel;
    a.out = (cpx *)out;
    a.opitch = (long long)opitch;
    // groups start on multiples of 16/os hops (absolute): the first begins
    // b0 mod (16/os) hops before hop 0
    const long long c0 = b0 % (long long)(16 / os);
    const long long ngroups = ((long long)m + c0 + G - 1) / G;
    // about two resident blocks' worth of groups per CU at a time
    long long gpb = ((long long)nel * ngroups + 2047) / 2048;
    if (gpb < 1) gpb = 1;
    a.groups_per_block = (int)gpb;
    a.nblk_per_el = (int)((ngroups + gpb - 1) / gpb);
    a.nblocks = a.nblk_per_el * (int)nel;
    const unsigned grid = ((unsigned)a.nblocks + 7u) & ~7u;
    return dispatch(a, M, os, kos, fmt, form, grid, (hipStream_t)s);
}

// byte copies of ragged jobs (the raw history of each stream)
__global__ void __launch_bounds__(256) k_copy_bytes(const atkgpu_copy_job *__restrict__ jobs, int njobs) {
    const int b = blockIdx.x;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].blk0 <= b) lo = mid; else hi = mid - 1;
    }
    const atkgpu_copy_job jb = jobs[lo];
    const unsigned char *src = (const unsigned char *)jb.src;
    unsigned char *dst = (unsigned char *)jb.dst;
    const int i0 = (b - jb.blk0) * 4096, i1 = min(i0 + 4096, jb.n);
    for (int i = i0 + (int)threadIdx.x; i < i1; i += 256) dst[i] = src[i];
}

extern "C" int atkgpu_copy_bytes_jobs(void *s, const void *jobs, int njobs, int nblocks) {
    if (!njobs || !nblocks) return 0;
    hipLaunchKernelGGL(k_copy_bytes, dim3(nblocks), dim3(256), 0, (hipStream_t)s, (const atkgpu_copy_job *)jobs,
                       njobs);
    return LAUNCH_OK();
}
