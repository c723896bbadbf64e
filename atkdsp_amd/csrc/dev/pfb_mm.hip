// Launcher of the matrix-core filterbank (pfb_mm.hpp): 256, 512 or 1024
// channels at 2x, 256 or 512 at 1x, up to 16 taps per branch (tap rows
// zero-padded to K = 4, 8, 12 or 16, which leaves every finite sum unchanged).
// 1024 channels at 1x are not covered: their windows (no branch sharing)
// overflow the registers.
#include "atkgpu_dev.h"
#include "pfb_mm.hpp"

namespace {

template <int M, int OS, int K>
int launch(const PfbMmArgs &a, unsigned grid, hipStream_t s) {
    hipLaunchKernelGGL((k_pfb_mm<M, OS, K>), dim3(grid), dim3(MmShape<M, OS, K>::NTHR), 0, s, a);
    return LAUNCH_OK();
}

template <int M>
int by_shape(const PfbMmArgs &a, int os, int K, unsigned grid, hipStream_t s) {
    if (os == 2) {
        switch (K) {
        case 4: return launch<M, 2, 4>(a, grid, s);
        case 8: return launch<M, 2, 8>(a, grid, s);
        case 12: return launch<M, 2, 12>(a, grid, s);
        case 16: return launch<M, 2, 16>(a, grid, s);
        }
    } else if constexpr (M <= 512) {
        if (os == 1) {
            switch (K) {
            case 8: return launch<M, 1, 8>(a, grid, s);
            case 16: return launch<M, 1, 16>(a, grid, s);
            }
        }
    }
    return -1;
}

}  // namespace

// The padded tap rows per branch the kernel runs for (M, os, kos), or 0 when
// the shape is not covered (kos = the fused kernel's padded row count).
extern "C" int atkgpu_pfb_mm_k(unsigned M, unsigned os, unsigned kos) {
    if (M != 256 && M != 512 && M != 1024) return 0;
    if (os == 2 && (kos == 8 || kos == 16 || kos == 24 || kos == 32)) return (int)(kos / 2);
    if (os == 1 && M <= 512 && (kos == 8 || kos == 16)) return (int)kos;
    return 0;
}

extern "C" int atkgpu_pfb_mm(void *s, const void *sbuf, size_t spitch, size_t nvalid, unsigned nel, size_t m,
                             long long b0, unsigned M, unsigned os, unsigned kos, const float *taps, const void *tw,
                             const void *cal, unsigned cal_nel, void *out, size_t opitch, size_t ocol) {
    if (!m || !nel) return 0;
    const int K = atkgpu_pfb_mm_k(M, os, kos);
    if (!K) return -1;
    PfbMmArgs a;
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
    const int G = 8;
    a.nchunk = (int)((m + G - 1) / G);
    // one block per CU is resident (LDS): about 256 runs in all, each a
    // contiguous stretch of one element's chunks
    int nrun = (int)((256 + nel - 1) / nel);
    if (nrun > a.nchunk) nrun = a.nchunk;
    if (nrun < 1) nrun = 1;
    a.cpr = (a.nchunk + nrun - 1) / nrun;
    a.nrun = (a.nchunk + a.cpr - 1) / a.cpr;
    a.nblocks = a.nrun * (int)nel;
    const unsigned grid = (unsigned)((a.nblocks + 7) / 8 * 8);
    hipStream_t st = (hipStream_t)s;
    switch (M) {
    case 256: return by_shape<256>(a, (int)os, K, grid, st);
    case 512: return by_shape<512>(a, (int)os, K, grid, st);
    case 1024: return by_shape<1024>(a, (int)os, K, grid, st);
    }
    return -1;
}
