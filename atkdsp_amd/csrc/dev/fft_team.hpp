// The power-of-two team FFT as a device routine. Stages exchange data through
// LDS; between a stage's writes and the next stage's reads the team must
// synchronise:
//   WAVE = false  block-wide barrier (every team of the block calls together);
//   WAVE = true   the team lies inside one wavefront (N/PT <= 64): LDS
//                 operations of a wavefront are performed in order, so only
//                 the compiler's reordering has to be stopped.
#pragma once
#include <hip/hip_runtime.h>
#include "fft_core.hpp"

template <bool WAVE>
__device__ __forceinline__ void team_sync() {
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
    } else {
        __syncthreads();
    }
}

// ---- team FFT over the stages of TeamFftPlan<N, PT> --------------------------
template <int N, int PT, int STAGE, bool WAVE = false>
__device__ __forceinline__ void team_fft_from(cpx *v, cpx *lds, const cpx *tw, int j, int Ns) {
    using P = TeamFftPlan<N, PT>;
    static_assert(!WAVE || N / PT <= 64, "a wave-synchronised team must fit one wavefront");
    if constexpr (STAGE < P::S) {
        constexpr bool last = (STAGE == P::NSTAGE - 1);
        team_stage<N, PT, PT>(v, lds, tw, j, Ns, last);
        if constexpr (!last) {
            team_sync<WAVE>();
            team_load<N, PT>(v, lds, j);
            team_sync<WAVE>();
            team_fft_from<N, PT, STAGE + 1, WAVE>(v, lds, tw, j, Ns * PT);
        }
    } else if constexpr (P::RL > 1) {
        team_stage<N, PT, P::RL>(v, lds, tw, j, Ns, true);
    }
}

template <int N, int PT>
__device__ __forceinline__ void team_fft(cpx *v, cpx *lds, const cpx *tw, int j) {
    team_fft_from<N, PT, 0, false>(v, lds, tw, j, 1);
}

// ---- team FFT with each thread's twiddles in registers ------------------------
// When every stage has the full radix PT (N = PT^S), thread j runs one
// butterfly a stage and its twiddles tw[r * k * N/(Ns*PT)], k = j mod Ns,
// depend only on j and the stage: a caller transforming many frames loads
// them once (team_twiddles) and the stages read no twiddle from memory. The
// arithmetic is team_fft's, value for value.
template <int N, int PT>
struct TeamTw {
    static constexpr int S = TeamFftPlan<N, PT>::S;
    static constexpr bool FULL = TeamFftPlan<N, PT>::RL == 1;   // team_fft_tw applies
    cpx w[S > 1 ? S - 1 : 1][PT - 1];     // stages 1 .. S-1 (stage 0 has no twiddles)
};

template <int N, int PT>
__device__ __forceinline__ void team_twiddles(TeamTw<N, PT> &t, const cpx *tw, int j) {
    int Ns = PT;
#pragma unroll
    for (int st = 1; st < TeamTw<N, PT>::S; ++st) {
        team_stage_twiddles<N, PT>(t.w[st - 1], tw, j, Ns);
        Ns *= PT;
    }
}

template <int N, int PT, int STAGE, bool WAVE>
__device__ __forceinline__ void team_fft_tw_from(cpx *v, cpx *lds, const TeamTw<N, PT> &t, int j, int Ns) {
    constexpr bool last = STAGE == TeamTw<N, PT>::S - 1;
    team_stage_w<N, PT>(v, lds, STAGE > 0 ? t.w[STAGE > 0 ? STAGE - 1 : 0] : nullptr, j, Ns, last);
    if constexpr (!last) {
        team_sync<WAVE>();
        team_load<N, PT>(v, lds, j);
        team_sync<WAVE>();
        team_fft_tw_from<N, PT, STAGE + 1, WAVE>(v, lds, t, j, Ns * PT);
    }
}

template <int N, int PT, bool WAVE = false>
__device__ __forceinline__ void team_fft_tw(cpx *v, cpx *lds, const TeamTw<N, PT> &t, int j) {
    static_assert(TeamTw<N, PT>::FULL, "every stage at the full radix");
    team_fft_tw_from<N, PT, 0, WAVE>(v, lds, t, j, 1);
}
