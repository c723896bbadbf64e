// The fused polyphase filterbank kernel for power-of-two channel counts
// (reference algorithm: src/channelizer.c:218-303). One launch does, per
// group of G consecutive hops of one element:
//
//   fold   thread = (column u of a hop-row, subset of hops). Hop i's window is
//          rows i .. i+KOS-1 of hop-rows R_j[u] = s[j*HOP + u], and fold index
//          q = r*HOP + u sums rows t = k*OS + r (k ascending) with tap
//          h[t*HOP + u]. Each loaded sample feeds every open hop it belongs to,
//          so a row is read once per group from L2, FMA accumulate in FP32.
//   FFT    the G fold rows go to LDS; teams of M/PT threads run the M-point
//          team FFT (fft_core.hpp) in registers with LDS exchanges.
//   rotate exp(-2 pi i k shift/M), shift = ((b+1)*HOP) mod M (int64 hop
//          counter; at 2x this is (-1)^k on alternate hops, applied before
//          the transform as a circular shift of its input by M/2: the fold
//          writes the two branches of such a hop to each other's LDS slots),
//          then the cal factor.
//   store  bins go back to LDS and leave channel-major, G consecutive hops of
//          a channel, two hops a lane as 16-byte buffer stores from a scalar
//          base per channel pass, out[(e*M + k)*opitch + i].
//
// Taps are zero-padded from K*OS to KOS (an instantiated size): a zero tap
// adds an exact zero, so results equal the unpadded sum for finite input.
//
// Input: cf32 rows (IN = PFB_IN_CF32), or raw receiver rows (IN = FMT_CU8,
// FMT_CI8, FMT_CI16) converted as they are loaded, ((float)b - c') * k with
// c' = c0 + off/k, the unpack's arithmetic (src/convert.c:33-111), so a raw
// row gives the same fold inputs bit for bit as its unpacked cf32 row; window
// samples before zlo (the stream's zero history) are zero.
#pragma once
#include <hip/hip_runtime.h>
#include "dev_common.hpp"
#include "fft_core.hpp"
#include "fft_team.hpp"

template <int M, int OS>
struct PfbShape {
    static constexpr int HOP = M / OS;
    static constexpr int HOPC = HOP >= 256 ? 256 : HOP;     // columns per pass
    static constexpr int SUB = 256 / HOPC;                   // hop subsets
    static constexpr int COLS = HOP >= 256 ? HOP / 256 : 1;  // column passes
    static constexpr int PT = M >= 256 ? 16 : 8;
    static constexpr int NT = M / PT;                        // FFT team
    static constexpr int TEAMS = 256 / NT;
#ifdef ATKGPU_PFB_GP0
    static constexpr int GP0 = ATKGPU_PFB_GP0;
#else
    static constexpr int GP0 = COLS >= 4 ? 4 : 8;
#endif
    // hops a subset folds: at least enough that every FFT team has a hop
    static constexpr int GP = SUB * GP0 >= TEAMS ? GP0 : TEAMS / SUB;
    static constexpr int G = SUB * GP;                       // hops per group
    // an LDS row holds one hop: the FFT's padded layout (lpad) during the
    // transform, natural bin order after it. In the store phase a lane reads
    // the two hops 2p, 2p+1 it stores, one per instruction: half a wavefront
    // reads L = min(G/2, 32) rows 2p x 32/L channels, 8-byte elements at
    // 2p*ROW + k. For G <= 32 a row stride of 32/G (mod 32) elements puts
    // those 32 reads on 32 distinct bank pairs (2p*32/G = p*32/L)
    static constexpr int RMOD = 32 / (G < 32 ? G : 32);
    static constexpr int ROW = lpad_size<M>() + ((RMOD - lpad_size<M>() % 32) % 32 + 32) % 32;
    static_assert(NT <= 256 && G % TEAMS == 0, "shape");
};

// One block of a ragged batch: the fields of the element it works on (its own
// stream, output, taps and counters), its tile (group) in blk0 and the
// element's tile count in nblk.
struct PfbJob {
    const cpx *s;
    cpx *out;              // channel k, hop i at out[k*opitch + i]
    const float *taps;
    long long b0;
    int nvalid, m, opitch, blk0, nblk;   // opitch at most PFB_OPITCH_MAX
    int pad_;
};

enum { PFB_IN_CF32 = -1 };

// The output pitch a launch accepts (hops): the store's lane offsets, up to
// 128 channel rows of a pass, are 32-bit byte offsets.
#define PFB_OPITCH_MAX ((size_t)1 << 20)

typedef unsigned pfb_u2 __attribute__((ext_vector_type(2)));
typedef unsigned pfb_u4 __attribute__((ext_vector_type(4)));

struct PfbFusedArgs {
    const PfbJob *jobs;    // non-null: a ragged batch, one descriptor per block
    int njobs;
    const cpx *s;          // stream rows, element e at s + e*spitch (raw input: the raw rows, spitch samples)
    long long spitch;
    long long nvalid;      // valid samples in each element's stream
    long long m;           // hops to produce
    long long b0;          // global hop index of hop 0
    const float *taps;     // KOS*HOP, zero-padded
    const cpx *tw;         // M twiddles
    const cpx *cal;        // [cal_nel][M] or null
    int cal_nel;
    cpx *out;              // element e, channel k, hop i at out[(e*M + k)*opitch + ocol + i]
    long long opitch, ocol;
    int groups_per_block;
    int nblk_per_el;       // blocks per element
    int nblocks;           // real blocks (grid may be padded to a multiple of 8)
    float c0, kk, off_re, off_im;   // raw input: the format's constants and the fixed offset
    int zlo;               // raw input: window samples before the stream's first (zero history)
};

// Bytes a sample of input IN takes.
template <int IN>
__host__ __device__ constexpr int pfb_in_bps() {
    return IN == PFB_IN_CF32 ? 8 : (IN == FMT_CU8 || IN == FMT_CI8) ? 2 : 4;
}

// One loaded raw sample word to the cf32 value the unpack makes of it.
template <int IN>
__device__ __forceinline__ cpx pfb_raw_conv(unsigned w, float cre, float cim, float kk) {
#pragma clang fp contract(off)
    float a, b;
    if constexpr (IN == FMT_CU8) {
        a = (float)(w & 0xffu);
        b = (float)((w >> 8) & 0xffu);
    } else if constexpr (IN == FMT_CI8) {
        a = (float)(signed char)(w & 0xffu);
        b = (float)(signed char)((w >> 8) & 0xffu);
    } else {
        a = (float)(short)(w & 0xffffu);
        b = (float)(short)(w >> 16);
    }
    return cmake((a - cre) * kk, (b - cim) * kk);
}

// A sample of input IN at byte offset voff + soff of buffer r.
template <int IN>
__device__ __forceinline__ cpx pfb_load(__amdgpu_buffer_rsrc_t r, int voff, int soff, float cre, float cim,
                                        float kk) {
    if constexpr (IN == PFB_IN_CF32) {
        return __builtin_bit_cast(cpx, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
    } else if constexpr (pfb_in_bps<IN>() == 2) {
        return pfb_raw_conv<IN>((unsigned)__builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0), cre, cim, kk);
    } else {
        return pfb_raw_conv<IN>((unsigned)__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0), cre, cim, kk);
    }
}

// The taps a thread folds with: KOS per column, the same for every group.
// Where a thread's columns need at most 64 of them they stay in registers for
// the whole block (HREG); otherwise each pass loads its own.
template <int M, int OS, int KOS, class S = PfbShape<M, OS>>
struct PfbTaps {
    static constexpr bool HREG = S::COLS * KOS <= 64;
    float h[HREG ? S::COLS : 1][HREG ? KOS : 1];
};

// One column pass of the fold: column u of hops first .. first+GP-1, whose
// window rows start at sample base. Per hop the taps run in ascending order
// (k ascending for each branch offset r). Rows and taps are buffer loads: the
// thread's one 32-bit byte offset in a VGPR, a row's constant offset in an
// SGPR, so a load costs no vector arithmetic (a clamped 64-bit global address
// costs four instructions).
//   CLAMP false: every row read lies inside the valid samples, and the whole
//   window is put in flight at once (KOS+GP-1 rows, and the taps unless they
//   are held in hk).
//   CLAMP true (the groups at an element's end): rows past the valid samples
//   only reach hops >= m (not stored) or zero padding taps, so the row index
//   is clamped to the last valid sample, not the value; the window comes in
//   NCH chunks of TH taps.
//   Raw input (IN != PFB_IN_CF32): the CLAMP path also zeroes the rows
//   before zlo, the stream's zero history.
template <int M, int OS, int KOS, bool CLAMP, bool HREG, int IN = PFB_IN_CF32, class S = PfbShape<M, OS>>
__device__ __forceinline__ void pfb_fold_pass(float (&ar)[S::GP][OS], float (&ai)[S::GP][OS], const void *s,
                                              const float *taps, const float *hk, int base, int nlast, int u,
                                              float cre = 0.f, float cim = 0.f, float kk = 1.f, int zlo = 0) {
    constexpr int HOP = S::HOP, GP = S::GP;
    constexpr int NCH = KOS >= 24 ? 4 : 2;
    constexpr int TH = KOS / NCH;
    constexpr int BPS = pfb_in_bps<IN>();
    // buffer descriptors (base, stride 0, size in bytes, the gfx9 format word)
    const __amdgpu_buffer_rsrc_t xr =
        __builtin_amdgcn_make_buffer_rsrc((void *)s, (short)0, (int)((unsigned)(nlast + 1) * (unsigned)BPS),
                                          0x00020000);
    const __amdgpu_buffer_rsrc_t hr =
        __builtin_amdgcn_make_buffer_rsrc((void *)taps, (short)0, (int)(KOS * HOP * 4), 0x00020000);
    const int xoff = base * BPS, hoff = u * 4;
    auto tap = [&](int t) -> float {
        if constexpr (HREG)
            return hk[t];
        else
            return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(hr, hoff, t * HOP * 4, 0));
    };
#pragma unroll
    for (int p = 0; p < GP; ++p)
#pragma unroll
        for (int r = 0; r < OS; ++r) { ar[p][r] = 0.f; ai[p][r] = 0.f; }
    if constexpr (!CLAMP) {
        float h[KOS];
        cpx x[KOS + GP - 1];
#pragma unroll
        for (int t = 0; t < KOS; ++t) h[t] = tap(t);
        // Where a thread folds two or more columns (hop-rows of 512 and more,
        // issue-bound), the window's loads issue at raised wave priority: the
        // SIMD's other wave (another block, mostly in its transform) yields
        // for those few cycles, so the loads' latency starts sooner (1024
        // channels at 2x: 1.83 -> 1.70 ms). The memory-bound narrow shapes
        // gain nothing from it (64 channels: 1.5 % slower) and keep one
        // priority.
        constexpr bool PRIO = S::COLS >= 2;
        if constexpr (PRIO) __builtin_amdgcn_s_setprio(3);
#pragma unroll
        for (int j = 0; j < KOS + GP - 1; ++j) x[j] = pfb_load<IN>(xr, xoff, j * HOP * BPS, cre, cim, kk);
        if constexpr (PRIO) __builtin_amdgcn_s_setprio(0);
#pragma unroll
        for (int half = 0; half < NCH; ++half)
#pragma unroll
            for (int p = 0; p < GP; ++p)
#pragma unroll
                for (int t = half * TH; t < (half + 1) * TH; ++t) {
                    ar[p][t % OS] = fmaf(h[t], x[p + t].x, ar[p][t % OS]);
                    ai[p][t % OS] = fmaf(h[t], x[p + t].y, ai[p][t % OS]);
                }
    } else {
#pragma unroll
        for (int half = 0; half < NCH; ++half) {
            const int t0 = half * TH;
            cpx x[GP + TH - 1];
            float h[TH];
#pragma unroll
            for (int jj = 0; jj < GP + TH - 1; ++jj) {
                const int row = base + (t0 + jj) * HOP;
                x[jj] = pfb_load<IN>(xr, min(row, nlast) * BPS, 0, cre, cim, kk);
                if constexpr (IN != PFB_IN_CF32)
                    if (row < zlo) x[jj] = cmake(0.f, 0.f);
            }
#pragma unroll
            for (int t = 0; t < TH; ++t) h[t] = tap(t0 + t);
#pragma unroll
            for (int p = 0; p < GP; ++p)
#pragma unroll
                for (int t = 0; t < TH; ++t) {
                    ar[p][(t0 + t) % OS] = fmaf(h[t], x[p + t].x, ar[p][(t0 + t) % OS]);
                    ai[p][(t0 + t) % OS] = fmaf(h[t], x[p + t].y, ai[p][(t0 + t) % OS]);
                }
        }
    }
}

template <int M, int OS, int KOS, int IN = PFB_IN_CF32>
__global__ void __launch_bounds__(256, 2) k_pfb_fused(PfbFusedArgs a) {
    using S = PfbShape<M, OS>;
    constexpr int HOP = S::HOP, G = S::G, GP = S::GP, ROW = S::ROW;
    __shared__ cpx lds[G * ROW];
    __shared__ cpx twl[M];

    // XCD-aware order: the hardware deals blocks round-robin to 8 XCDs, so
    // give each XCD a contiguous run of tiles (neighbours share KOS-1 rows)
    const int nb8 = (int)gridDim.x / 8;
    const int lb = (int)(blockIdx.x % 8) * nb8 + (int)(blockIdx.x / 8);
    if (lb >= a.nblocks) return;
    int e, tile, nlast, m, opitch, gpb, ntile;
    const void *s;
    cpx *oute;
    const float *taps;
    long long b0;
    if (a.jobs) {
        // ragged batch: one descriptor per block (its job's fields, its tile
        // in blk0 and the job's tile count in nblk), so a block makes one read
        // of possibly host-resident memory before it starts
        const PfbJob &jb = a.jobs[lb];
        e = 0;
        tile = jb.blk0;
        s = jb.s;
        oute = jb.out;
        taps = jb.taps;
        b0 = jb.b0;
        nlast = jb.nvalid - 1;
        m = jb.m;
        opitch = jb.opitch;
        gpb = 1;
        ntile = jb.nblk;
    } else {
        e = lb / a.nblk_per_el;
        tile = lb % a.nblk_per_el;
        // per-element bases in 64-bit once; every offset below fits 32 bits
        s = (const unsigned char *)a.s + (long long)e * a.spitch * pfb_in_bps<IN>();
        oute = a.out + (long long)e * M * a.opitch + a.ocol;
        taps = a.taps;
        b0 = a.b0;
        nlast = (int)a.nvalid - 1;
        m = (int)a.m;
        opitch = (int)a.opitch;
        gpb = a.groups_per_block;
        ntile = a.nblk_per_el;
    }
    const int tid = threadIdx.x;
    const int sub = tid / S::HOPC;
    const int ucol = tid % S::HOPC;
    for (int i = tid; i < M; i += 256) twl[i] = a.tw[i];
    const cpx *cal = (a.cal && e < a.cal_nel) ? a.cal + (long long)e * M : nullptr;
    const bool flip_base = (b0 & 1) != 0;     // parity of the hop counter at hop 0
    // raw input: the folded constants as the unpack forms them (src/convert.c:47, :57, :68)
    float cre = 0.f, cim = 0.f, kk = 1.f;
    int zlo = 0;
    if constexpr (IN != PFB_IN_CF32) {
        kk = a.kk;
        cre = a.c0 + a.off_re / a.kk;
        cim = a.c0 + a.off_im / a.kk;
        zlo = a.zlo;
    }

    // The taps of the thread's columns, where they fit in registers
    using TP = PfbTaps<M, OS, KOS>;
    TP tk;
    if constexpr (TP::HREG) {
        const __amdgpu_buffer_rsrc_t hr =
            __builtin_amdgcn_make_buffer_rsrc((void *)taps, (short)0, (int)(KOS * HOP * 4), 0x00020000);
#pragma unroll
        for (int c = 0; c < S::COLS; ++c)
#pragma unroll
            for (int t = 0; t < KOS; ++t)
                tk.h[c][t] = __builtin_bit_cast(
                    float, __builtin_amdgcn_raw_buffer_load_b32(hr, (ucol + c * 256) * 4, t * HOP * 4, 0));
    }

    // The FFT rounds stay rolled, and so do the column passes where the taps
    // are loaded per pass: the unrolled fold is long, and the kernel must fit
    // the instruction cache.
    // Groups are dealt to an element's blocks in turn (group tile, then
    // tile + ntile, ...): blocks resident together on an XCD fold adjacent
    // groups, whose windows share KOS-1 rows, so those rows are read from HBM
    // once and from L2 after; the neighbours' output segments also fill the
    // same cache lines.
    for (int grp = 0; grp < gpb; ++grp) {
        const int i0 = (tile + grp * ntile) * G;
        if (i0 >= m) break;
        // ---- fold ----
        // the group reads rows up to (i0 + G + KOS - 1)*HOP - 1 (same for every block thread)
        const bool inside = (long long)(i0 + G + KOS - 1) * HOP <= (long long)nlast + 1 && i0 * HOP >= zlo;
        auto pass = [&](int c, const float *hk) {
            const int u = ucol + c * 256;
            float ar[GP][OS], ai[GP][OS];
            const int base = (i0 + sub * GP) * HOP + u;
            if (inside)
                pfb_fold_pass<M, OS, KOS, false, TP::HREG, IN>(ar, ai, s, taps, hk, base, nlast, u, cre, cim, kk,
                                                               zlo);
            else
                pfb_fold_pass<M, OS, KOS, true, TP::HREG, IN>(ar, ai, s, taps, hk, base, nlast, u, cre, cim, kk,
                                                              zlo);
            // at 2x a hop whose b+1 is odd takes (-1)^k after the transform
            // (reference src/channelizer.c:275-276 with shift = M/2); the
            // transform of its input circularly shifted by M/2 gives the same
            // values (an exactly zero bin keeps +0 where a negation gives -0):
            // branch r's fold goes to branch 1-r's slot
#pragma unroll
            for (int p = 0; p < GP; ++p) {
                const bool sw = OS == 2 && (((i0 + sub * GP + p) & 1) == (flip_base ? 1 : 0));
#pragma unroll
                for (int r = 0; r < OS; ++r)
                    lds[(sub * GP + p) * ROW + lpad((sw ? OS - 1 - r : r) * HOP + u)] = cmake(ar[p][r], ai[p][r]);
            }
        };
        if constexpr (TP::HREG) {
#pragma unroll
            for (int c = 0; c < S::COLS; ++c) pass(c, tk.h[c]);
        } else {
#pragma unroll 1
            for (int c = 0; c < S::COLS; ++c) pass(c, nullptr);
        }
        __syncthreads();
        // ---- FFT and cal: team tm transforms hops tm, tm+TEAMS, ... ----
        const int tm = tid / S::NT, j = tid % S::NT;
        // a team inside one wavefront synchronises its stages wave-locally;
        // the rows it reads were published by the block barrier above
        constexpr bool WV = S::NT <= 64;
#pragma unroll 1
        for (int rnd = 0; rnd < G / S::TEAMS; ++rnd) {
            const int p = tm + rnd * S::TEAMS;
            cpx *row = lds + p * ROW;
            cpx v[S::PT];
            team_load<M, S::PT>(v, row, j);
            team_sync<WV>();
            team_fft_from<M, S::PT, 0, WV>(v, row, twl, j, 1);
#pragma unroll
            for (int r = 0; r < S::PT; ++r) {
                const int k = j + r * S::NT;
                cpx y = v[r];
                if (cal) y = cmul(y, cal[k]);
                row[k] = y;
            }
            if constexpr (!WV) __syncthreads();
        }
        if constexpr (WV) __syncthreads();       // every row done before the store phase
        // ---- channel-major store: lane (channel lc of a pass, hops 2*lp and
        // 2*lp+1), one 16-byte buffer store per pass at a per-block lane
        // offset from a scalar base, so a store costs no vector arithmetic ----
        {
            static_assert(G % 2 == 0 && G <= 64, "a group's hops are whole lane pairs of one pass");
            constexpr int LP2 = G / 2;                   // lanes per channel
            constexpr int CH_PER_PASS = 256 / LP2;
            constexpr int Q = 4;                         // passes whose bins are read together
            static_assert(M % (Q * CH_PER_PASS) == 0, "whole pass sets");
            const int lp = tid % LP2, lc = tid / LP2;
            const int voff = (lc * opitch + 2 * lp) * 8;
            const bool two = i0 + 2 * lp + 1 < m, one = i0 + 2 * lp < m;
#pragma unroll 1
            for (int k0 = 0; k0 < M; k0 += Q * CH_PER_PASS) {
                cpx v0[Q], v1[Q];
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    v0[q] = lds[(2 * lp) * ROW + k0 + q * CH_PER_PASS + lc];
                    v1[q] = lds[(2 * lp + 1) * ROW + k0 + q * CH_PER_PASS + lc];
                }
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    // records cover the lane offsets of this pass; the hop
                    // tests above keep every store inside rows < m
                    const __amdgpu_buffer_rsrc_t orr = __builtin_amdgcn_make_buffer_rsrc(
                        (void *)(oute + (long long)(k0 + q * CH_PER_PASS) * opitch + i0), (short)0, 0x7fffffff,
                        0x00020000);
                    if (two) {
                        const pfb_u4 w = {__builtin_bit_cast(unsigned, v0[q].x), __builtin_bit_cast(unsigned, v0[q].y),
                                          __builtin_bit_cast(unsigned, v1[q].x), __builtin_bit_cast(unsigned, v1[q].y)};
                        __builtin_amdgcn_raw_buffer_store_b128(w, orr, voff, 0, 0);
                    } else if (one) {
                        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(pfb_u2, v0[q]), orr, voff, 0, 0);
                    }
                }
            }
        }
        __syncthreads();
    }
}
