// The fused FP32 filterbank reading raw receiver rows (8-bit and 12/16-bit
// integer samples) and converting them as it loads: the FP32 form on the same
// raw input the matrix-core forms read. Instantiated for the shapes those
// forms cover: 64 and 128 channels, 1x and 2x, 8/16/32 padded tap rows.
#include "atkgpu_dev.h"
#include "pfb_fused.hpp"

namespace {

template <int M, int OS, int IN>
int launch_kos(const PfbFusedArgs &a, int kos, unsigned grid, hipStream_t s) {
    switch (kos) {
    case 8: hipLaunchKernelGGL((k_pfb_fused<M, OS, 8, IN>), dim3(grid), dim3(256), 0, s, a); return LAUNCH_OK();
    case 16: hipLaunchKernelGGL((k_pfb_fused<M, OS, 16, IN>), dim3(grid), dim3(256), 0, s, a); return LAUNCH_OK();
    case 32: hipLaunchKernelGGL((k_pfb_fused<M, OS, 32, IN>), dim3(grid), dim3(256), 0, s, a); return LAUNCH_OK();
    }
    return -1;
}

template <int IN>
int launch_in(const PfbFusedArgs &a, unsigned M, int os, int kos, unsigned grid, hipStream_t s) {
    if (M == 64 && os == 1) return launch_kos<64, 1, IN>(a, kos, grid, s);
    if (M == 64 && os == 2) return launch_kos<64, 2, IN>(a, kos, grid, s);
    if (M == 128 && os == 1) return launch_kos<128, 1, IN>(a, kos, grid, s);
    if (M == 128 && os == 2) return launch_kos<128, 2, IN>(a, kos, grid, s);
    return -1;
}

// the unpack's constants of each format (src/convert.c:33-111)
bool fmt_consts(int fmt, float *c0, float *kk) {
    switch (fmt) {
    case FMT_CU8: *c0 = 127.5f; *kk = 1.0f / 127.5f; return true;
    case FMT_CI8: *c0 = 0.0f; *kk = 1.0f / 128.0f; return true;
    case FMT_CI16: *c0 = 0.0f; *kk = 1.0f / 32768.0f; return true;
    case FMT_CI16Q11: *c0 = 0.0f; *kk = 1.0f / 2048.0f; return true;
    }
    return false;
}

}  // namespace

extern "C" int atkgpu_pfb_fused_raw_supported(unsigned M, unsigned os, unsigned kos, int fmt) {
    float c0, kk;
    return (M == 64 || M == 128) && (os == 1 || os == 2) && (kos == 8 || kos == 16 || kos == 32) &&
           fmt_consts(fmt, &c0, &kk);
}

extern "C" int atkgpu_pfb_fused_raw(void *s, const void *raw, size_t spitch, size_t nvalid, unsigned nel, size_t m,
                                    long long b0, unsigned M, unsigned os, unsigned kos, int fmt, float off_re,
                                    float off_im, int zlo, const float *taps, const void *tw, void *out,
                                    size_t opitch) {
    if (!m || !nel) return 0;
    if (!atkgpu_pfb_fused_raw_supported(M, os, kos, fmt) || opitch > PFB_OPITCH_MAX) return -1;
    const int G = atkgpu_pfb_group(M, os);
    PfbFusedArgs a = {};
    a.s = (const cpx *)raw;
    a.spitch = (long long)spitch;
    a.nvalid = (long long)nvalid;
    a.m = (long long)m;
    a.b0 = b0;
    a.taps = taps;
    a.tw = (const cpx *)tw;
    a.out = (cpx *)out;
    a.opitch = (long long)opitch;
    fmt_consts(fmt, &a.c0, &a.kk);
    a.off_re = off_re;
    a.off_im = off_im;
    a.zlo = zlo;
    // about 8192 blocks in all, as the cf32-row launch at these channel
    // counts (256 x 1,200,000 cu8 samples into 64 channels: 0.67-0.68 ms
    // against 0.73-0.74 ms with 2048 blocks; 16384 and 32768 no better)
    const long long ngroups = ((long long)m + G - 1) / G;
    long long gpb = ((long long)nel * ngroups + 8191) / 8192;
    if (gpb < 1) gpb = 1;
    a.groups_per_block = (int)gpb;
    a.nblk_per_el = (int)((ngroups + gpb - 1) / gpb);
    a.nblocks = a.nblk_per_el * (int)nel;
    const unsigned grid = (unsigned)((a.nblocks + 7) / 8 * 8);
    hipStream_t st = (hipStream_t)s;
    switch (fmt) {
    case FMT_CU8: return launch_in<FMT_CU8>(a, M, (int)os, (int)kos, grid, st);
    case FMT_CI8: return launch_in<FMT_CI8>(a, M, (int)os, (int)kos, grid, st);
    default: return launch_in<FMT_CI16>(a, M, (int)os, (int)kos, grid, st);
    }
}
