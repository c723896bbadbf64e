// The polyphase filterbank with its fold on the matrix cores (reference
// algorithm: src/channelizer.c:218-303).
//
// Fold index q = r*HOP + u (branch r, hop-row column u) of hop h is an FIR
// over column u of the stream:
//   y_q[h] = sum_{k<K} g_q[k] * X_u[h + OS*k + r],
//   g_q[k] = taps[k*M + q],   X_u[j] = s[j*HOP + u].
// A chunk is eight hops. Column u's window for column half p in {0,1} is
//   W_p[j] = X_u[wb + p*PS + OS*j],  j < K+3,
// (PS = 1 at 2x, 4 at 1x; wb the chunk's window base) and the fold of 16
// consecutive q is one 16-block 4x4 product per step j:
//   A[a][j] = g_q[j - a]          (zero outside [0, K): a banded Toeplitz)
//   B[j][(c, p)] = part c of W_p[j]
//   C[a][(c, p)] = part c of y_q[wb + p*PS + OS*a - r].
// Block b of the instruction is q = 16*tile + b; lane l holds A row l%4 and
// B, C column l%4 of block l/4.
//
// At 2x both branches of a column read the same window: branch 0 gives hops
// wb + {0..7} and branch 1 hops wb - 1 + {0..7}. With wb = i0 + 1 a chunk
// completes hops i0 .. i0+7 (branch 1 now, branch 0 of i0 in the chunk
// before) and leaves branch 0 of hop i0+8 for the next chunk, so the fold
// rows are a ring of nine hops. The first chunk of a run is preceded by a
// branch-0 pass for its hop i0.
//
// FP32 form: v_mfma_f32_4x4x1_16b_f32, one step j at a time. The matrix unit
// rounds each step as one fused multiply-add, so row a accumulates the taps
// g_q[0], g_q[1], ... in ascending order exactly as the FP32 fold
// fmaf(tap, x, acc) does (a zero tap adds an exact zero to a finite sum):
// the output is bit-identical to the VALU fold of pfb_fused.hpp for finite
// input. Non-finite samples are the exception: a zero entry of the band
// times an infinite or NaN sample puts NaN into up to three neighbouring
// hops of the same column half.
//
// Per chunk a block of 512 threads (8 wavefronts) folds all M outputs of its
// hops into LDS, runs the M-point team FFT, rotation and cal there (one team
// per hop, as pfb_fused.hpp) and stores them channel-major. A block streams a
// contiguous run of chunks of one element: each wavefront keeps the windows
// of its columns in registers and slides them by G/OS steps a chunk, so each
// sample is read from memory about once per run; the taps live in LDS, row k
// of M (+pad) floats, and a lane reads its band entry there at every step.
#pragma once
#include <hip/hip_runtime.h>
#include "dev_common.hpp"
#include "fft_core.hpp"
#include "fft_team.hpp"

typedef float mm_f4 __attribute__((ext_vector_type(4)));

template <int M, int OS, int K>
struct MmShape {
    static constexpr int HOP = M / OS;
    static constexpr int NTHR = 512;
    static constexpr int NW = NTHR / 64;
    static constexpr int WPW = HOP / 16 / NW;           // column tiles (windows) per wavefront
    static constexpr int G = 8;                         // hops per chunk
    static constexpr int RING = OS == 2 ? G + 1 : G;    // fold rows
    static constexpr int PS = OS == 2 ? 1 : 4;          // hop offset of column half 1
    static constexpr int NJ = K + 3;                    // band steps
    static constexpr int ADV = G / OS;                  // steps the window moves a chunk
    static constexpr int PT = M >= 256 ? 16 : 8;        // the fused kernel's team FFT radix
    static constexpr int NT = M / PT;                   // FFT team
    static constexpr int TEAMS = NTHR / NT;
    static constexpr int ROW = lpad_size<M>() | 1;      // one hop in LDS
    static constexpr int TS = M + 20;                   // tap row stride (floats)
    // LDS carve (bytes): fold rows | tap rows | twiddles | guard. A band entry
    // outside [0, K) reads up to three tap rows before (fold rows) or after
    // (twiddles, guard) the table, and the value is replaced by zero.
    static constexpr int OFF_TAP = RING * ROW * 8;
    static constexpr int OFF_TW = OFF_TAP + K * TS * 4;
    static constexpr int GUARD = 3 * TS * 4 > M * 8 ? 3 * TS * 4 - M * 8 : 0;
    static constexpr int LDS_BYTES = OFF_TW + M * 8 + GUARD;
    static_assert(OS == 1 || OS == 2, "1x or 2x");
    static_assert(WPW >= 1 && HOP % (16 * NW) == 0, "whole column tiles per wavefront");
    static_assert(ADV <= NJ && NT <= 64, "window shift; one wavefront per FFT team");
    static_assert(OFF_TAP >= 3 * TS * 4, "negative band rows stay inside the fold rows");
    static_assert(LDS_BYTES <= 160 * 1024, "LDS");
};

struct PfbMmArgs {
    const cpx *s;            // element e's stream row at s + e*spitch (hist ++ leftover ++ input)
    long long spitch;
    long long nvalid;        // valid samples in each row
    long long m;             // hops to produce
    long long b0;            // global hop index of hop 0
    const float *taps;       // taps[k*M + q], k < K
    const cpx *tw;           // M twiddles
    const cpx *cal;          // [cal_nel][M] or null
    int cal_nel;
    cpx *out;                // element e, channel k, hop i at out[(e*M + k)*opitch + ocol + i]
    long long opitch, ocol;
    int nrun;                // runs (blocks) per element
    int cpr;                 // chunks per run
    int nchunk;              // chunks per element
    int nblocks;             // real blocks (the grid is padded to a multiple of 8)
};

template <int M, int OS, int K>
__global__ void __launch_bounds__(512, 1) k_pfb_mm(PfbMmArgs a) {
    using S = MmShape<M, OS, K>;
    constexpr int HOP = S::HOP, G = S::G, ROW = S::ROW, TS = S::TS, NJ = S::NJ, WPW = S::WPW, PS = S::PS;
    constexpr int ADV = S::ADV, RING = S::RING;
    __shared__ __attribute__((aligned(16))) unsigned char lds_raw[S::LDS_BYTES];
    cpx *fold = (cpx *)lds_raw;
    float *tl = (float *)(lds_raw + S::OFF_TAP);
    cpx *twl = (cpx *)(lds_raw + S::OFF_TW);

    const int nb8 = (int)gridDim.x / 8;
    const int lb = (int)(blockIdx.x % 8) * nb8 + (int)(blockIdx.x / 8);
    if (lb >= a.nblocks) return;
    const int e = lb / a.nrun, run = lb % a.nrun;
    const int c_beg = run * a.cpr, c_end = min(c_beg + a.cpr, a.nchunk);
    if (c_beg >= c_end) return;
    const float *sf = (const float *)(a.s + (long long)e * a.spitch);
    cpx *oute = a.out + (long long)e * M * a.opitch + a.ocol;
    const int nlast = (int)a.nvalid - 1, m = (int)a.m, opitch = (int)a.opitch;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int b = lane >> 2, cl = lane & 3;            // block, row (A) / column (B, C)
    const int comp = cl & 1, ph = cl >> 1;             // column = (part, half)

    for (int i = tid; i < M; i += S::NTHR) twl[i] = a.tw[i];
    for (int i = tid; i < K * M; i += S::NTHR) tl[(i / M) * TS + (i % M)] = a.taps[i];
    const cpx *cal = (a.cal && e < a.cal_nel) ? a.cal + (long long)e * M : nullptr;
    const bool flip_base = (a.b0 & 1) != 0;

    // column u of each of this lane's window tiles, as the sample offset of
    // its step 0 at window base 0
    int sb[WPW];
#pragma unroll
    for (int t = 0; t < WPW; ++t) sb[t] = ph * PS * HOP + 16 * (w * WPW + t) + b;
    // win[t][j] = part comp of W_ph[j] for tile t; nxt[t] holds the ADV new
    // steps of the next chunk, loaded at the start of this one so that the
    // memory latency hides behind a whole chunk of work
    float win[WPW][NJ];
    float nxt[WPW][ADV];
    auto ldx = [&](int wb, int t, int j) -> float {
        const int idx = min(max((wb + OS * j) * HOP + sb[t], 0), nlast);
        return sf[2 * idx + comp];
    };
    constexpr int WB0 = OS == 2 ? 1 : 0;               // window base - i0
    // the ring slot of hop h (h >= -8)
    auto slot = [&](int h) -> int { return (h + 9 * RING) % RING; };
    const int c_first = OS == 2 ? c_beg - 1 : c_beg;   // a branch-0 pass before the run
#pragma unroll
    for (int t = 0; t < WPW; ++t)
#pragma unroll
        for (int j = 0; j < NJ; ++j) win[t][j] = ldx(c_first * G + WB0, t, j);
    __syncthreads();

    for (int c = c_first; c < c_end; ++c) {
        const int i0 = c * G, wb = i0 + WB0;
        const bool lead = c < c_beg;                   // the branch-0 pass: no FFT, no store
        // ---- fold on the matrix cores ----
        if (c > c_first) {
#pragma unroll
            for (int t = 0; t < WPW; ++t) {
#pragma unroll
                for (int j = 0; j < NJ - ADV; ++j) win[t][j] = win[t][j + ADV];
#pragma unroll
                for (int k = 0; k < ADV; ++k) win[t][NJ - ADV + k] = nxt[t][k];
            }
        }
        if (c + 1 < c_end) {
#pragma unroll
            for (int t = 0; t < WPW; ++t)
#pragma unroll
                for (int k = 0; k < ADV; ++k) nxt[t][k] = ldx(wb + G, t, NJ - ADV + k);
        }
#pragma unroll
        for (int t = 0; t < WPW; ++t) {
            const int u = 16 * (w * WPW + t) + b;
#pragma unroll
            for (int r = 0; r < OS; ++r) {
                if (r > 0 && lead) break;
                const int q = r * HOP + u;
                mm_f4 acc = {0.f, 0.f, 0.f, 0.f};
                const float *ta = tl + q - cl * TS;
#pragma unroll
                for (int j = 0; j < NJ; ++j) {
                    float av = ta[j * TS];
                    if (j < 3 || j >= K) av = (j - cl >= 0 && j - cl < K) ? av : 0.0f;
                    acc = __builtin_amdgcn_mfma_f32_4x4x1f32(av, win[t][j], acc, 0, 0, 0);
                }
                float *fo = (float *)fold + 2 * lpad(q) + comp;
                const int h0 = wb + ph * PS - r;       // hop of row a = 0
#pragma unroll
                for (int v = 0; v < 4; ++v) fo[2 * slot(h0 + OS * v) * ROW] = acc[v];
            }
        }
        if (lead) continue;
        __syncthreads();
        // ---- FFT, rotation, cal: team tm transforms hop i0 + tm ----
        {
            const int tm = tid / S::NT, j = tid % S::NT;
#pragma unroll 1
            for (int p = tm; p < G; p += S::TEAMS) {
                cpx *row = fold + slot(i0 + p) * ROW;
                cpx v[S::PT];
                team_load<M, S::PT>(v, row, j);
                team_sync<true>();
                team_fft_from<M, S::PT, 0, true>(v, row, twl, j, 1);
                const bool flip = (OS == 2) && (((i0 + p) & 1) == (flip_base ? 1 : 0));
#pragma unroll
                for (int r = 0; r < S::PT; ++r) {
                    const int k = j + r * S::NT;
                    cpx y = v[r];
                    if (flip && (k & 1)) y = cmake(-y.x, -y.y);
                    if (cal) y = cmul(y, cal[k]);
                    row[k] = y;
                }
            }
        }
        __syncthreads();
        // ---- channel-major store: 8 lanes a channel, the chunk's 8 hops ----
        {
            const int lp = tid & 7, lc = tid >> 3;
            const bool ok = i0 + lp < m;
            const cpx *src = fold + slot(i0 + lp) * ROW;
#pragma unroll 4
            for (int k0 = 0; k0 < M; k0 += S::NTHR / 8) {
                const int k = k0 + lc;
                if (ok) oute[k * opitch + i0 + lp] = src[k];
            }
        }
        __syncthreads();
    }
}
