// Shared helpers for the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// a launch error (bad configuration) surfaces here; execution faults surface
// at the next synchronisation
static inline int LAUNCH_OK() { return hipGetLastError() == hipSuccess ? 0 : -4; }

// sample formats (values are the ABI's: reference include/atkdsp.h:215-221)
enum { FMT_CU8 = 0, FMT_CI8 = 1, FMT_CI16 = 2, FMT_CI16Q11 = 3, FMT_CF32 = 4 };

static inline unsigned grid_for(size_t work, unsigned per_block, unsigned cap) {
    size_t g = (work + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}
