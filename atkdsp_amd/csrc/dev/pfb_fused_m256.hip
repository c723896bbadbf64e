// Instantiations of the fused filterbank kernel for 256 channels.
#include "pfb_fused.hpp"

template <int OS, int KOS>
static int launch(const PfbFusedArgs &a, unsigned grid, hipStream_t s) {
    hipLaunchKernelGGL((k_pfb_fused<256, OS, KOS>), dim3(grid), dim3(256), 0, s, a);
    return LAUNCH_OK();
}

int pfb_fused_m256(const PfbFusedArgs &a, int os, int kos, unsigned grid, hipStream_t s) {
    if (os == 1) {
        switch (kos) {
        case 8: return launch<1, 8>(a, grid, s);
        case 16: return launch<1, 16>(a, grid, s);
        case 24: return launch<1, 24>(a, grid, s);
        case 32: return launch<1, 32>(a, grid, s);
        }
    } else if (os == 2) {
        switch (kos) {
        case 8: return launch<2, 8>(a, grid, s);
        case 16: return launch<2, 16>(a, grid, s);
        case 24: return launch<2, 24>(a, grid, s);
        case 32: return launch<2, 32>(a, grid, s);
        }
    }
    return -1;
}
