// The power-of-two team FFT as a device routine (LDS exchanges between
// stages, block-wide barriers: every team of a block must call it together).
#pragma once
#include <hip/hip_runtime.h>
#include "fft_core.hpp"

// ---- team FFT over the stages of TeamFftPlan<N, PT> --------------------------
template <int N, int PT, int STAGE>
__device__ __forceinline__ void team_fft_from(cpx *v, cpx *lds, const cpx *tw, int j, int Ns) {
    using P = TeamFftPlan<N, PT>;
    if constexpr (STAGE < P::S) {
        constexpr bool last = (STAGE == P::NSTAGE - 1);
        team_stage<N, PT, PT>(v, lds, tw, j, Ns, last);
        if constexpr (!last) {
            __syncthreads();
            team_load<N, PT>(v, lds, j);
            __syncthreads();
            team_fft_from<N, PT, STAGE + 1>(v, lds, tw, j, Ns * PT);
        }
    } else if constexpr (P::RL > 1) {
        team_stage<N, PT, P::RL>(v, lds, tw, j, Ns, true);
    }
}

template <int N, int PT>
__device__ __forceinline__ void team_fft(cpx *v, cpx *lds, const cpx *tw, int j) {
    team_fft_from<N, PT, 0>(v, lds, tw, j, 1);
}

