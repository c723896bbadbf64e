// The filterbank's matrix-core form: the polyphase fold on MFMA, straight
// from raw 8-bit or 12-bit samples held exactly in BF16 / FP16, taps split
// into hi + lo terms (or one term), FP32 accumulation. The FFT, rotation,
// cal and channel-major store are the FP32 fused kernel's (reference
// algorithm: src/channelizer.c:218-303).
//
// The fold of one hop-row column u is an FIR over the rows X[j] = s[j*HOP+u]:
//   fold_h[r*HOP + u] = sum_{t = r mod OS} tap_t(u) * X[h + t],
//   tap_t(u) = taps[t*HOP + u].
// It is lifted to a matrix product per column (one v_mfma_f32_16x16x32 per
// k-step and tap term):
//   rows  m = (r, a): branch r, hop offset a in [0, 16/OS)
//   cols  n = (ri, cb): real/imaginary part, hop block cb in [0, 8)
//   k     p: data row i0 + cb*16/OS + p
//   A[m][p] = tap_{p-a}(u) if 0 <= p-a < KOS and p-a = r (mod OS), else 0
//   B[p][n] = X[i0 + cb*16/OS + p].part(ri)
//   C[m][n] = fold of hop i0 + cb*16/OS + a, branch r, part ri.
// A depends on the column only: a lane's 8 values A[m][32k + 8g .. +7] are a
// window of the column's tap vector, so each (column, tap term, branch) keeps
// one zero-padded tap row in LDS for the whole launch and a lane reads its
// window from it. B comes from the group's rows staged in LDS, column-major
// by column u.
//
// Integer samples v (cu8: 2b - 255, ci8: b, ci16q11: b) are exact in the
// element type, so the only rounding before the FP32 accumulation is the
// taps' split. The unpack's affine map is applied after the product:
//   fold = acc * scale - off * tapsum_r(u),
// where scale maps v to the reference's units (src/convert.c:46-68) and
// removes the taps' power-of-two scale. The history before a stream's first
// sample is zero in those units (the reference callocs it), which no raw
// byte encodes for cu8: samples before zlo are staged as 0 and left out of
// the offset's tap sum.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "dev_common.hpp"
#include "fft_core.hpp"
#include "fft_team.hpp"

typedef _Float16 mx_h8 __attribute__((ext_vector_type(8)));
typedef __bf16 mx_b8 __attribute__((ext_vector_type(8)));
typedef float mx_f4 __attribute__((ext_vector_type(4)));

enum { MX_BF16 = 0, MX_F16 = 1 };

template <int M, int OS, int KOS>
struct MxShape {
    static constexpr int HOP = M / OS;
    static constexpr int HA = 16 / OS;                 // hops per row group of a tile
    static constexpr int NCB = OS == 1 ? 4 : 8;        // hop blocks per tile (of 8 columns pairs)
    static constexpr int G = NCB * HA;                 // hops per group (64)
    static constexpr int KS = (HA + KOS - 1 + 31) / 32;   // k-steps of 32 data rows
    static constexpr int NR = (NCB - 1) * HA + 32 * KS;   // rows staged per group
    static constexpr int NRP = NR + 2;                 // column stride: an odd number of words
    static constexpr int PLANE = HOP * NRP + 32;       // imaginary plane offset (halfs): 16 banks on
    static constexpr int NTHR = 512;
    static constexpr int NW = NTHR / 64;
    static constexpr int CPW = HOP / NW;               // columns per wavefront
    static constexpr int PT = M >= 256 ? 16 : 8;
    static constexpr int NT = M / PT;                  // FFT team
    static constexpr int TEAMS = NTHR / NT;
    static constexpr int ROW = lpad_size<M>() | 1;
    static constexpr int LPC = G >= 64 ? 64 : G;       // store lanes per channel
    static constexpr int NRB = NR / 8;                 // blocks of 8 rows
    static constexpr int NIT = (NRB * HOP + NTHR - 1) / NTHR;  // (row block, column) items per thread
    static constexpr int TPL = 32 * KS + 16;           // padded tap row: taps t = -16 .. 32*KS-1
    static_assert(HOP % NW == 0 && G % TEAMS == 0 && ((NRP / 2) & 1) == 1 && NR % 8 == 0 && NIT <= 4,
                  "shape");
};

struct PfbMxArgs {
    const unsigned char *raw;  // element e's raw row at raw + e*spitch*bps; hop i's window at sample i*HOP
    long long spitch;          // samples per element row
    long long nvalid;          // valid samples per element row
    long long m;               // hops to produce
    long long b0;              // global hop index of hop 0
    const uint16_t *tpad;      // [HOP][NTERM][OS][TPL] padded tap rows (tap t at index t + 16)
    const float *tapsum;       // [OS][HOP]: sum of column u's taps of branch r
    const float *taps;         // KOS*HOP FP32 taps (the partial sums near the stream's start)
    int zlo;                   // window samples before the stream's first sample: zero history
    float scale;               // integer sample -> reference units, tap scale removed
    const float2 *offs;        // per-element offsets
    const cpx *tw;             // M twiddles
    const cpx *cal;            // [cal_nel][M] or null
    int cal_nel;
    cpx *out;                  // element e, channel k, hop i at out[(e*M + k)*opitch + i]
    long long opitch;
    int groups_per_block, nblk_per_el, nblocks;
};

// one raw sample as its bits (2 bytes for 8-bit formats, 4 for ci16q11)
template <int FMT>
__device__ __forceinline__ uint32_t mx_load(const unsigned char *p) {
    if constexpr (FMT == FMT_CU8 || FMT == FMT_CI8) return *(const unsigned short *)p;
    else return *(const uint32_t *)p;
}

// the sample's exact integer parts, as floats (cu8: 2b - 255)
template <int FMT>
__device__ __forceinline__ void mx_parts(uint32_t w, float &fr, float &fi) {
    if constexpr (FMT == FMT_CU8) {
        fr = fmaf(2.0f, (float)(w & 0xffu), -255.0f);
        fi = fmaf(2.0f, (float)((w >> 8) & 0xffu), -255.0f);
    } else if constexpr (FMT == FMT_CI8) {
        fr = (float)(int)(signed char)(w & 0xffu);
        fi = (float)(int)(signed char)((w >> 8) & 0xffu);
    } else {
        fr = (float)(int)(short)(w & 0xffffu);
        fi = (float)(int)(short)(w >> 16);
    }
}

// two exact integers (as floats) packed as element-type bits, a in the low
// half. They are exact in the element type, so BF16's bits are the float's
// top half and FP16's round-toward-zero conversion is exact.
template <int ELT>
__device__ __forceinline__ uint32_t mx_pack(float a, float b) {
    if constexpr (ELT == MX_BF16)
        return __builtin_amdgcn_perm(__float_as_uint(b), __float_as_uint(a), 0x07060302u);
    else
        return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(a, b));
}

template <int ELT>
__device__ __forceinline__ mx_f4 mx_mma(uint4 a, uint4 b, mx_f4 c) {
    if constexpr (ELT == MX_BF16)
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(mx_b8, a), __builtin_bit_cast(mx_b8, b),
                                                       c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(mx_h8, a), __builtin_bit_cast(mx_h8, b),
                                                      c, 0, 0, 0);
}

template <int M, int OS, int KOS, int FMT, int ELT, int NTERM>
__global__ void __launch_bounds__(512, 4) k_pfb_mx(PfbMxArgs a) {
    using S = MxShape<M, OS, KOS>;
    constexpr int HOP = S::HOP, G = S::G, ROW = S::ROW, NR = S::NR, NRP = S::NRP, HA = S::HA, KS = S::KS;
    constexpr int CPW = S::CPW, NTHR = S::NTHR, TPL = S::TPL;
    constexpr int BPS = (FMT == FMT_CU8 || FMT == FMT_CI8) ? 2 : 4;
    __shared__ cpx lds[G * ROW];
    __shared__ uint16_t lx[2 * S::PLANE];
    __shared__ uint32_t ltp[HOP * NTERM * OS * TPL / 2];
    __shared__ float lts[OS * HOP];
    __shared__ cpx twl[M];

    // XCD-aware order, as the FP32 kernel: each XCD a contiguous run of tiles
    const int nb8 = (int)gridDim.x / 8;
    const int lb = (int)(blockIdx.x % 8) * nb8 + (int)(blockIdx.x / 8);
    if (lb >= a.nblocks) return;
    const int e = lb / a.nblk_per_el, tile = lb % a.nblk_per_el, ntile = a.nblk_per_el;
    const unsigned char *raw = a.raw + (long long)e * a.spitch * BPS;
    const int nlast = (int)a.nvalid - 1, m = (int)a.m, opitch = (int)a.opitch;
    cpx *oute = a.out + (long long)e * M * a.opitch;
    const float2 off = a.offs[e];
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    for (int i = tid; i < M; i += NTHR) twl[i] = a.tw[i];
    const cpx *cal = (a.cal && e < a.cal_nel) ? a.cal + (long long)e * M : nullptr;
    const bool flip_base = (a.b0 & 1) != 0;
    // Groups are anchored to the absolute hop index: a group starts on a
    // multiple of HA, so a hop's offset in its tile (and with it the places
    // of its taps along the product's k dimension, which fix the order the
    // matrix unit sums them in) depends on the hop alone, not on where the
    // run started. The run's first group starts c0 hops before hop 0; those
    // hops are computed and not stored, and their rows before the run's
    // first window are staged as zeros (they reach no stored hop).
    const int c0 = (int)(a.b0 % HA);
    auto gstart = [&](int gi) { return gi * G - c0; };

    // the padded tap rows and the tap sums of every column
    {
        const uint32_t *src = (const uint32_t *)a.tpad;
        for (int i = tid; i < HOP * NTERM * OS * TPL / 2; i += NTHR) ltp[i] = src[i];
        for (int i = tid; i < OS * HOP; i += NTHR) lts[i] = a.tapsum[i];
    }
    const int n = lane & 15, g = lane >> 4;
    const int ri = n & 1, cb = n >> 1;
    // as an A lane: row n = (branch ar, hop offset aa); its window of a tap
    // row starts at tap 8g - aa (index 8g - aa + 16), odd when aa is odd
    const int ar = n / HA, aa = n % HA;
    const bool aodd = (aa & 1) != 0;
    const bool live = cb < S::NCB;                     // tile columns past the group are idle
    const float offr = ri ? off.y : off.x;

    // A group's rows i0 .. i0+NR-1 are loaded into registers ahead of use and
    // staged as exact BF16/FP16, column-major. A thread takes 8 consecutive
    // rows of one column (a wavefront's loads cover whole rows: coalesced)
    // and writes each part as four 32-bit words. The loads of group g+2 are
    // issued when group g+1 is staged, so they are in flight behind group
    // g+1's fold and group g's FFT and store. Rows past the valid samples
    // reach only hops >= m or zero taps: the address is clamped, the value
    // stays finite.
    uint32_t pre[S::NIT][8];
    int pre_i0 = 0;                                    // first row of the loaded group
    auto load_group = [&](int i0) {
        pre_i0 = i0;
#pragma unroll
        for (int it = 0; it < S::NIT; ++it) {
            const int idx = tid + it * NTHR;
            const int jb = idx / HOP, u = idx % HOP;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int row = i0 + 8 * jb + q;
                const int si = min(row * HOP + u, nlast);
                pre[it][q] = (idx < S::NRB * HOP && row >= 0) ? mx_load<FMT>(raw + (long long)si * BPS) : 0u;
            }
        }
    };
    auto stage = [&]() {
        uint32_t *lx32 = (uint32_t *)lx;
        // only groups at the stream's very start reach back before its first
        // sample, and a run's first group before its first window: every row
        // there (row * HOP + u < 0 <= zlo) is staged as zero
        const bool zcheck = pre_i0 * HOP < a.zlo;
#pragma unroll
        for (int it = 0; it < S::NIT; ++it) {
            const int idx = tid + it * NTHR;
            if (idx < S::NRB * HOP) {
                const int jb = idx / HOP, u = idx % HOP;
                uint32_t wr[4], wi[4];
#pragma unroll
                for (int q = 0; q < 8; q += 2) {
                    float r0, i0v, r1, i1v;
                    mx_parts<FMT>(pre[it][q], r0, i0v);
                    mx_parts<FMT>(pre[it][q + 1], r1, i1v);
                    if (zcheck) {
                        if ((pre_i0 + 8 * jb + q) * HOP + u < a.zlo) { r0 = 0.f; i0v = 0.f; }
                        if ((pre_i0 + 8 * jb + q + 1) * HOP + u < a.zlo) { r1 = 0.f; i1v = 0.f; }
                    }
                    wr[q / 2] = mx_pack<ELT>(r0, r1);
                    wi[q / 2] = mx_pack<ELT>(i0v, i1v);
                }
                const int w = (u * NRP + 8 * jb) >> 1;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    lx32[w + q] = wr[q];
                    lx32[(S::PLANE >> 1) + w + q] = wi[q];
                }
            }
        }
    };
    if (gstart(tile) < m) {
        load_group(gstart(tile));
        stage();
        const int nx = gstart(tile + ntile);
        if (1 < a.groups_per_block && nx < m) load_group(nx);
    }
    __syncthreads();

    for (int grp = 0; grp < a.groups_per_block; ++grp) {
        const int i0 = gstart(tile + grp * ntile);
        if (i0 >= m) break;
        // ---- fold on MFMA, one column at a time ----
        // (a lane's output rows 4g..4g+3 share one branch: orr)
        const int orr = (4 * g) / HA;
        const bool zfold = i0 * HOP < a.zlo;              // group-uniform
#pragma unroll 2
        for (int c = 0; c < CPW; ++c) {
            const int u = wv * CPW + c;
            mx_f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int k = 0; k < KS; ++k) {
                // 8 consecutive rows of column u, part ri: 4-byte aligned
                uint4 bv = make_uint4(0u, 0u, 0u, 0u);
                if (live) {
                    const uint32_t *bp = (const uint32_t *)&lx[ri * S::PLANE + u * NRP + cb * HA + 32 * k + 8 * g];
                    bv = make_uint4(bp[0], bp[1], bp[2], bp[3]);
                }
#pragma unroll
                for (int t = 0; t < NTERM; ++t) {
                    const int w0 = (((u * NTERM + t) * OS + ar) * TPL + 32 * k + 8 * g - aa + 16) >> 1;
                    const uint32_t d0 = ltp[w0], d1 = ltp[w0 + 1], d2 = ltp[w0 + 2], d3 = ltp[w0 + 3];
                    uint4 av;
                    if (aodd) {
                        const uint32_t d4 = ltp[w0 + 4];
                        av = make_uint4(__builtin_amdgcn_alignbit(d1, d0, 16), __builtin_amdgcn_alignbit(d2, d1, 16),
                                        __builtin_amdgcn_alignbit(d3, d2, 16), __builtin_amdgcn_alignbit(d4, d3, 16));
                    } else {
                        av = make_uint4(d0, d1, d2, d3);
                    }
                    acc = mx_mma<ELT>(av, bv, acc);
                }
            }
            // lane holds rows 4g..4g+3 of tile column n
#pragma unroll
            for (int i = 0; live && i < 4; ++i) {
                const int mrow = 4 * g + i;
                const int r = mrow / HA, ao = mrow % HA;
                const int hh = cb * HA + ao;
                // the offset applies to real samples only: a window that
                // reaches back before the stream's start sums fewer taps
                float tsum = lts[orr * HOP + u];
                if (zfold && (i0 + hh) * HOP + u < a.zlo) {
                    tsum = 0.f;
                    for (int t = r; t < KOS; t += OS)
                        if ((i0 + hh + t) * HOP + u >= a.zlo) tsum += a.taps[t * HOP + u];
                }
                const float y = fmaf(acc[i], a.scale, -offr * tsum);
                ((float *)lds)[2 * (hh * ROW + lpad(r * HOP + u)) + ri] = y;
            }
        }
        __syncthreads();
        // the rows are free: stage the next group (overlapping this group's
        // FFT and store), then start the loads of the one after
        if (grp + 1 < a.groups_per_block && gstart(tile + (grp + 1) * ntile) < m) {
            stage();
            const int nx = gstart(tile + (grp + 2) * ntile);
            if (grp + 2 < a.groups_per_block && nx < m) load_group(nx);
        }
        // ---- FFT, rotation, cal (as the FP32 kernel) ----
        const int tm = tid / S::NT, j = tid % S::NT;
        constexpr bool WV = S::NT <= 64;
#pragma unroll 1
        for (int rnd = 0; rnd < G / S::TEAMS; ++rnd) {
            const int p = tm + rnd * S::TEAMS;
            cpx *row = lds + p * ROW;
            cpx v[S::PT];
            team_load<M, S::PT>(v, row, j);
            team_sync<WV>();
            team_fft_from<M, S::PT, 0, WV>(v, row, twl, j, 1);
            const bool flip = (OS == 2) && (((i0 + p) & 1) == (flip_base ? 1 : 0));
#pragma unroll
            for (int r = 0; r < S::PT; ++r) {
                const int k = j + r * S::NT;
                cpx y = v[r];
                if (flip && (k & 1)) y = cmake(-y.x, -y.y);
                if (cal) y = cmul(y, cal[k]);
                row[k] = y;
            }
            if constexpr (!WV) __syncthreads();
        }
        if constexpr (WV) __syncthreads();
        // ---- channel-major store ----
        {
            constexpr int CH_PER_PASS = NTHR / S::LPC;
            const int lp = tid % S::LPC, lc = tid / S::LPC;
#pragma unroll 2
            for (int k0 = 0; k0 < M; k0 += CH_PER_PASS) {
                const int k = k0 + lc;
                if (k < M) {
                    cpx *ok = oute + k * opitch + i0;
#pragma unroll
                    for (int p = lp; p < G; p += S::LPC)
                        if (i0 + p >= 0 && i0 + p < m) ok[p] = lds[p * ROW + k];
                }
            }
        }
        __syncthreads();
    }
}

// The padded tap rows: row (u, term t, branch r) holds at index tt + 16 the
// term of column u's tap tt when 0 <= tt < KOS and tt = r (mod OS), else 0,
// for tt = -16 .. 32*KS-1. Terms: hi = the tap rounded to the element type,
// lo = the rounded remainder; taps scaled by tscale (a power of two).
template <int ELT>
__global__ void k_mx_tpad(const float *taps, int HOP, int KOS, int OS, int TPL, int NTERM, float tscale,
                          uint16_t *tpad) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= HOP * NTERM * OS * TPL) return;
    const int i = idx % TPL, r = (idx / TPL) % OS, t = (idx / (TPL * OS)) % NTERM, u = idx / (TPL * OS * NTERM);
    const int tt = i - 16;
    const float v = (tt >= 0 && tt < KOS && tt % OS == r) ? taps[tt * HOP + u] * tscale : 0.0f;
    float term = v;
    if (t == 1) {
        if constexpr (ELT == MX_BF16) term = v - (float)(__bf16)v;
        else term = v - (float)(_Float16)v;
    }
    if constexpr (ELT == MX_BF16) tpad[idx] = __builtin_bit_cast(uint16_t, (__bf16)term);
    else tpad[idx] = __builtin_bit_cast(uint16_t, (_Float16)term);
}
