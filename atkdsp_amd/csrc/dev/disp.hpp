// The display reduction's block routine, shared by the lockstep chain and the
// serving loop: a block transforms a run of frames of one stream (window,
// N-point team FFT, |X|^2/N^2 in float) and keeps each bin's detector value
// in registers (reference src/fft.c:177-191), then writes one partial line in
// display order (fftshifted, src/fft.c:190).
#pragma once
#include <hip/hip_runtime.h>
#include "dev_common.hpp"
#include "fft_core.hpp"
#include "fft_team.hpp"

// NTHR threads a block; TEAMS FFT teams of N/16 threads transform frames side
// by side
template <int N, int NTHR = 256>
struct DispShape {
    static constexpr int PT = 16;
    static constexpr int NT = N / PT;
    static constexpr int TEAMS = NT >= NTHR ? 1 : NTHR / NT;
    static constexpr int TL = lpad_size<N>();
};

// the reference's detector step (src/fft.c:184-189) as a select, so the
// accumulation has no branches: the first frame takes pw, MAX keeps the first
// of equals (pw > acc), MIN likewise (pw < acc), AVG sums
template <int DET>
__device__ __forceinline__ float det_step(float acc, float pw, bool first) {
    float nv;
    if constexpr (DET == 0) nv = pw > acc ? pw : acc;
    else if constexpr (DET == 1) nv = pw < acc ? pw : acc;
    else nv = acc + pw;
    return first ? pw : nv;
}

template <int DET>
__device__ __forceinline__ float det_merge(float v, float u) {
    if constexpr (DET == 0) return u > v ? u : v;
    else if constexpr (DET == 1) return u < v ? u : v;
    else return v + u;
}

__device__ __forceinline__ float det_merge_rt(int det, float v, float u) {
    if (det == 0) return u > v ? u : v;
    if (det == 1) return u < v ? u : v;
    return v + u;
}

// 20 log10(sqrt(P) + 1e-12), spectrum_reduce's dB (src/fft.c:286-292)
__device__ __forceinline__ float disp_db(float v, float scale) {
#pragma clang fp contract(off)
    return 20.0f * log10f(sqrtf(v * scale) + 1e-12f);
}

// Frames [f0, f1) of one stream, a thread's samples j + r*NT of frame f being
// src.load<PT, NT>(f, j, v). Every thread of the NTHR-thread block calls it
// (the team FFT synchronises the block); src.own<PT, NT>(f, j, v) is called
// once for each frame the block transforms (and for no other) with its
// samples, so a source may write or sum them.
// twl holds the N twiddles in LDS (loaded and synchronised by the caller);
// lds holds TEAMS * TL and red TEAMS * N (TEAMS > 1) entries. The line goes to
// out (global or LDS) in display order; with TEAMS > 1 it is complete when
// the call returns, with TEAMS == 1 after the caller's barrier.
template <int N, int DET, int NTHR = 256, class Src>
__device__ __forceinline__ void disp_frames(Src &src, int f0, int f1, const float *win, const cpx *twl, cpx *lds,
                                            float *red, float *out) {
#pragma clang fp contract(off)
    using D = DispShape<N, NTHR>;
    constexpr int PT = D::PT, NT = D::NT, TEAMS = D::TEAMS, TL = D::TL;
    const int tm = threadIdx.x / NT, j = threadIdx.x % NT;
    float w[PT];
#pragma unroll
    for (int r = 0; r < PT; ++r) w[r] = win ? win[j + r * NT] : 1.0f;
    const float inv2 = 1.0f / ((float)N * (float)N);
    float acc[PT];
#pragma unroll
    for (int r = 0; r < PT; ++r) acc[r] = 0.0f;
    bool first = true;
    // this thread's twiddles for every stage, once for all frames (when every
    // stage has the full radix; other sizes read them from LDS per stage)
    constexpr bool TWR = TeamTw<N, PT>::FULL;
    TeamTw<N, PT> ttw;
    if constexpr (TWR) team_twiddles<N, PT>(ttw, twl, j);
    // each team takes every TEAMS-th frame of the range; the next frame's
    // samples are loaded while this frame is transformed
    cpx nx[PT];
    {
        const int f = f0 + tm;
        src.template load<PT, NT>(f < f1 ? f : f0, j, nx);
    }
    for (int fbase = f0; fbase < f1; fbase += TEAMS) {
        const int f = fbase + tm;
        const bool act = f < f1;
        if (act) src.template own<PT, NT>(f, j, nx);
        cpx v[PT];
#pragma unroll
        for (int r = 0; r < PT; ++r) v[r] = act ? cmake(nx[r].x * w[r], nx[r].y * w[r]) : cmake(0.f, 0.f);
        {
            const int fn = f + TEAMS;
            src.template load<PT, NT>(fn < f1 ? fn : f0, j, nx);
        }
        if constexpr (TWR) team_fft_tw<N, PT>(v, lds + tm * TL, ttw, j);
        else team_fft<N, PT>(v, lds + tm * TL, twl, j);
#pragma unroll
        for (int r = 0; r < PT; ++r) {
            const float pw = (v[r].x * v[r].x + v[r].y * v[r].y) * inv2;
            const float nv = det_step<DET>(acc[r], pw, first);
            acc[r] = act ? nv : acc[r];
        }
        first = first && !act;
    }
    // combine the teams in team order, then write the partial line in display
    // order: FFT bin k = j + r*NT sits at display bin (k + N/2) mod N
    if constexpr (TEAMS == 1) {
#pragma unroll
        for (int r = 0; r < PT; ++r) out[j + ((r + PT / 2) % PT) * NT] = first ? 0.0f : acc[r];
    } else {
        // a team with no frame (only at the very end of a range) contributes
        // nothing: mark it with NaN, skipped below
#pragma unroll
        for (int r = 0; r < PT; ++r) red[tm * N + j + r * NT] = first ? __builtin_nanf("") : acc[r];
        __syncthreads();
        for (int k = threadIdx.x; k < N; k += NTHR) {
            float v = red[k];
            for (int t = 1; t < TEAMS; ++t) {
                const float u = red[t * N + k];
                if (u != u) continue;
                if (v != v) { v = u; continue; }
                v = det_merge<DET>(v, u);
            }
            out[(k + N / 2) % N] = (v != v) ? 0.0f : v;
        }
    }
}

// decimate_max (src/display.c:19-37) of an N-bin line to cols columns, by the
// threads of a block
__device__ __forceinline__ void disp_decimate(const float *in, int n, int cols, float *d) {
    for (int c = threadIdx.x; c < cols; c += blockDim.x) {
        if (cols >= n) {
            d[c] = in[c < n ? c : n - 1];
            continue;
        }
        const int per = n / cols;
        const int a0 = c * per, b0 = (c + 1 == cols) ? n : a0 + per;
        float best = in[a0];
        for (int i = a0 + 1; i < b0; ++i)
            if (in[i] > best) best = in[i];
        d[c] = best;
    }
}
