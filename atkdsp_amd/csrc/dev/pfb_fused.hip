// Dispatch of the fused filterbank kernel (pfb_fused.hpp) by channel count,
// oversampling and padded tap-row count.
#include "atkgpu_dev.h"
#include "pfb_fused.hpp"

int pfb_fused_m64(const PfbFusedArgs &, int, int, unsigned, hipStream_t);
int pfb_fused_m128(const PfbFusedArgs &, int, int, unsigned, hipStream_t);
int pfb_fused_m256(const PfbFusedArgs &, int, int, unsigned, hipStream_t);
int pfb_fused_m512(const PfbFusedArgs &, int, int, unsigned, hipStream_t);
int pfb_fused_m1024(const PfbFusedArgs &, int, int, unsigned, hipStream_t);
int pfb_fused_m2048(const PfbFusedArgs &, int, int, unsigned, hipStream_t);

static int group_size(unsigned M, unsigned os) {
    switch (M * 4 + os) {
#define G_(m, o) case m * 4 + o: return PfbShape<m, o>::G;
        G_(64, 1) G_(64, 2) G_(128, 1) G_(128, 2) G_(256, 1) G_(256, 2)
        G_(512, 1) G_(512, 2) G_(1024, 1) G_(1024, 2) G_(2048, 1) G_(2048, 2)
#undef G_
    }
    return 0;
}

extern "C" int atkgpu_pfb_fused_kos(unsigned M, unsigned os, unsigned K) {
    if (os != 1 && os != 2) return 0;
    if (M < 64 || M > 2048 || (M & (M - 1))) return 0;
    const unsigned kos = K * os;
    if (kos <= 8) return 8;
    if (kos <= 16) return 16;
    if (kos <= 24) return 24;
    if (kos <= 32) return 32;
    return 0;
}

extern "C" int atkgpu_pfb_fused(void *s, const void *sbuf, size_t spitch, size_t nvalid, unsigned nel,
                                size_t m, long long b0, unsigned M, unsigned os, unsigned kos,
                                const float *taps, const void *tw, const void *cal, unsigned cal_nel,
                                void *out, size_t opitch, size_t ocol) {
    if (!m || !nel) return 0;
    const int G = group_size(M, os);
    if (!G) return -1;
    if (opitch > PFB_OPITCH_MAX) return -1;
    PfbFusedArgs a;
    a.jobs = nullptr;
    a.njobs = 0;
    a.s = (const cpx *)sbuf;
    a.spitch = (long long)spitch;
    a.nvalid = (long long)nvalid;
    a.m = (long long)m;
    a.b0 = b0;
    a.taps = taps;
    a.tw = (const cpx *)tw;
    a.cal = (const cpx *)cal;
    a.cal_nel = (int)cal_nel;
    a.out = (cpx *)out;
    a.opitch = (long long)opitch;
    a.ocol = (long long)ocol;
    const long long ngroups = ((long long)m + G - 1) / G;
    // about 2048 blocks in all at 256 channels and more: four rounds of the
    // 512 that are resident at once (two per CU on 256 CUs); measured at
    // 128 x 2,000,000 samples into 1024 channels against 1024, 1536, 3072
    // and 4096 blocks. The narrow shapes (64 and 128 channels, memory-bound,
    // small groups) run best with about 8192 (64 channels at 256 x 1,200,000
    // samples: 1.00 against 1.05 ms)
    const long long nbt = M <= 128 ? 8192 : 2048;
    long long gpb = ((long long)nel * ngroups + nbt - 1) / nbt;
    if (gpb < 1) gpb = 1;
    a.groups_per_block = (int)gpb;
    a.nblk_per_el = (int)((ngroups + gpb - 1) / gpb);
    a.nblocks = a.nblk_per_el * (int)nel;
    const unsigned grid = (unsigned)((a.nblocks + 7) / 8 * 8);
    hipStream_t st = (hipStream_t)s;
    switch (M) {
    case 64: return pfb_fused_m64(a, (int)os, (int)kos, grid, st);
    case 128: return pfb_fused_m128(a, (int)os, (int)kos, grid, st);
    case 256: return pfb_fused_m256(a, (int)os, (int)kos, grid, st);
    case 512: return pfb_fused_m512(a, (int)os, (int)kos, grid, st);
    case 1024: return pfb_fused_m1024(a, (int)os, (int)kos, grid, st);
    case 2048: return pfb_fused_m2048(a, (int)os, (int)kos, grid, st);
    }
    return -1;
}

static int dispatch_m(const PfbFusedArgs &a, unsigned M, int os, int kos, unsigned grid, hipStream_t st) {
    switch (M) {
    case 64: return pfb_fused_m64(a, os, kos, grid, st);
    case 128: return pfb_fused_m128(a, os, kos, grid, st);
    case 256: return pfb_fused_m256(a, os, kos, grid, st);
    case 512: return pfb_fused_m512(a, os, kos, grid, st);
    case 1024: return pfb_fused_m1024(a, os, kos, grid, st);
    case 2048: return pfb_fused_m2048(a, os, kos, grid, st);
    }
    return -1;
}

extern "C" int atkgpu_pfb_group(unsigned M, unsigned os) { return group_size(M, os); }

// A ragged batch: jobs (device-visible array, job j owns blocks
// [blk0, blk0 + nblk), one group of G hops each) of one (M, os, kos) class.
extern "C" int atkgpu_pfb_fused_jobs(void *s, const void *jobs, int njobs, int nblocks, unsigned M, unsigned os,
                                     unsigned kos, const void *tw) {
    if (!njobs || !nblocks) return 0;
    if (!group_size(M, os)) return -1;
    PfbFusedArgs a = {};
    a.jobs = (const PfbJob *)jobs;
    a.njobs = njobs;
    a.tw = (const cpx *)tw;
    a.nblocks = nblocks;
    a.groups_per_block = 1;
    const unsigned grid = (unsigned)((nblocks + 7) / 8 * 8);
    return dispatch_m(a, M, (int)os, (int)kos, grid, (hipStream_t)s);
}
