/* Identity, threads, slots and the library-wide device pool, plus the
 * stateless entry points: unpack, unpack_dc, window, window_stats, the display
 * helpers, stitch, and W1/W2. Reference: src/atkdsp.c:43-80, src/convert.c,
 * src/fft.c:133-152 and :433-444, src/display.c, src/detect.c:116-134,
 * src/channelizer.c:420-471. */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <sched.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>
#include "internal.h"

/* ---- identity -------------------------------------------------------------- */
#define STR2(x) #x
#define STR(x) STR2(x)

static char g_build_info[256];
static pthread_once_t g_info_once = PTHREAD_ONCE_INIT;

static void make_info(void) {
    const char *dev = "";
    if (atkgpu_init() == 0) dev = atkgpu_device_name();
    snprintf(g_build_info, sizeof g_build_info,
             ATKDSP_VERSION_STRING " abi " STR(ATKDSP_ABI_VERSION) " gcc hip gfx950 %s built "
             __DATE__ " " __TIME__, dev[0] ? dev : "(no device)");
}

int atkdsp_abi_version(void) { return ATKDSP_ABI_VERSION; }
const char *atkdsp_version_string(void) { return ATKDSP_VERSION_STRING; }
const char *atkdsp_build_info(void) {
    pthread_once(&g_info_once, make_info);
    return g_build_info;
}

int atkdsp_bytes_per_sample(int fmt) {
    switch (fmt) {
    case ATKDSP_FMT_CU8: return 2;
    case ATKDSP_FMT_CI8: return 2;
    case ATKDSP_FMT_CI16: return 4;
    case ATKDSP_FMT_CI16Q11: return 4;
    case ATKDSP_FMT_CF32: return 8;
    default: return 0;
    }
}

/* The reference's OpenMP worker count (src/atkdsp.c:62-80). Device results do
 * not depend on it; it is kept so the binding and its tests behave the same. */
static atomic_int g_threads;

int atkdsp_set_threads(int n) {
    atomic_store(&g_threads, n > 0 ? n : 0);
    return atkdsp_get_threads();
}

int atkdsp_get_threads(void) {
    const int t = atomic_load(&g_threads);
    if (t > 0) return t;
    const long c = sysconf(_SC_NPROCESSORS_ONLN);
    return c > 0 ? (int)c : 1;
}

/* ---- slots ---------------------------------------------------------------------- */
int atk_slot_init(atk_slot *s, size_t a, size_t b, size_t c) {
    memset(s, 0, sizeof *s);
    atomic_init(&s->busy, 0);
    if (atkgpu_stream_create(&s->stream)) return ATKDSP_E_NOMEM;
    s->bytes_a = a; s->bytes_b = b; s->bytes_c = c;
    s->d_a = a ? atkgpu_dmalloc(a) : NULL;
    s->d_b = b ? atkgpu_dmalloc(b) : NULL;
    s->d_c = c ? atkgpu_dmalloc(c) : NULL;
    s->d_small = atkgpu_dmalloc(ATK_SMALL_BYTES);
    if ((a && !s->d_a) || (b && !s->d_b) || (c && !s->d_c) || !s->d_small) {
        atk_slot_free(s);
        return ATKDSP_E_NOMEM;
    }
    return ATKDSP_OK;
}

void atk_slot_free(atk_slot *s) {
    if (s->stream) atkgpu_sync(s->stream);
    if (s->stream2) {
        atkgpu_sync(s->stream2);
        atkgpu_stream_destroy(s->stream2);
        s->stream2 = NULL;
    }
    atkgpu_dfree(s->d_a); atkgpu_dfree(s->d_b); atkgpu_dfree(s->d_c); atkgpu_dfree(s->d_small);
    if (s->stream) atkgpu_stream_destroy(s->stream);
    s->d_a = s->d_b = s->d_c = s->d_small = NULL;
    s->stream = NULL;
}

int atk_claim(atk_slot *slots, int n) {
    for (;;) {
        for (int i = 0; i < n; ++i) {
            int expect = 0;
            if (atomic_compare_exchange_strong(&slots[i].busy, &expect, 1)) return i;
        }
        sched_yield();
    }
}

void atk_release(atk_slot *slots, int i) { atomic_store(&slots[i].busy, 0); }

/* ---- the library-wide pool -------------------------------------------------------- */
#define POOL_CHUNK ATKGPU_POOL_CHUNK           /* samples a slot moves per pass */

static atk_pool g_pool;
static atk_pool *g_pool_ok;
static pthread_once_t g_pool_once = PTHREAD_ONCE_INIT;

static void pool_init(void) {
    if (atkgpu_init() != 0) return;
    g_pool.chunk = POOL_CHUNK;
    g_pool.nslot = 4;
    for (int i = 0; i < g_pool.nslot; ++i) {
        /* a: raw / input block, b: converted block, c: pole scratch */
        if (atk_slot_init(&g_pool.slot[i], POOL_CHUNK * 8, POOL_CHUNK * 8,
                          (4 * (POOL_CHUNK / 256 + 1) + 16) * sizeof(double)) ||
            atkgpu_stream_create(&g_pool.slot[i].stream2)) {
            for (int j = 0; j <= i; ++j) atk_slot_free(&g_pool.slot[j]);
            return;
        }
    }
    g_pool_ok = &g_pool;
}

atk_pool *atk_pool_get(void) {
    pthread_once(&g_pool_once, pool_init);
    return g_pool_ok;
}

/* ---- unpack ------------------------------------------------------------------------ */
/* one piece of a pipelined unpack: items [i0, i0 + cnt) of the slot's
 * buffers; the piece is one pass of the mean (its sums at pass i0 / pass) */
typedef struct {
    const atk_slot *s;
    int fmt, bps;
    float cre, cim, k, off_re, off_im;
    double *parts, *sums;
} unpack_piece_ctx;

static int unpack_piece(void *vc, void *st, size_t i0, size_t cnt) {
    const unpack_piece_ctx *c = (const unpack_piece_ctx *)vc;
    return atkgpu_unpack(st, (const char *)c->s->d_a + i0 * (size_t)c->bps, c->fmt, cnt, c->cre, c->cim, c->k,
                         c->off_re, c->off_im, (atkdsp_cf32 *)c->s->d_b + i0, c->parts,
                         c->sums ? c->sums + 2 * (i0 / ATKGPU_MEAN_PASS) : NULL);
}

static ptrdiff_t unpack_run(const void *raw, size_t nbytes, int fmt, atkdsp_cf32 *out, size_t out_cap,
                            float off_re, float off_im, double *mean_re, double *mean_im,
                            float dc_alpha, atkdsp_cf32 *dc_state) {
    if (!raw || !out) return ATKDSP_E_ARG;
    const int bps = atkdsp_bytes_per_sample(fmt);
    if (bps == 0) return ATKDSP_E_ARG;
    size_t n = nbytes / (size_t)bps;
    if (n > out_cap) n = out_cap;
    const int want_mean = mean_re || mean_im;
    const int pole = dc_alpha > 0.0f && dc_state;
    if (n == 0) {
        if (want_mean) {
            if (mean_re) *mean_re = (double)off_re;
            if (mean_im) *mean_im = (double)off_im;
        }
        return 0;
    }
    atk_pool *pool = atk_pool_get();
    if (!pool) return ATKDSP_E_FFT;
    float cre, cim, k;
    atkgpu_unpack_consts(fmt, off_re, off_im, &cre, &cim, &k);
    const int si = atk_claim(pool->slot, pool->nslot);
    atk_slot *s = &pool->slot[si];
    /* small scratch: the block partials, the pole state, the pass sums */
    double *d_parts = (double *)s->d_small;                       /* 2*PARTS doubles */
    float *d_state = (float *)(d_parts + 2 * ATKGPU_MEAN_PARTS);  /* pole state      */
    double *d_sums = d_parts + 2 * ATKGPU_MEAN_PARTS + 2;         /* 2 per pass      */
    double sum_re = 0.0, sum_im = 0.0;
    int rc = 0;
    if (pole) {
        /* the one-pole DC block runs over whole chunks (its scan's chunking
         * fixes its rounding); no mean is asked for with it */
        rc = atkgpu_h2d(s->stream, d_state, dc_state, sizeof(atkdsp_cf32));
        for (size_t a = 0; a < n && !rc; a += pool->chunk) {
            const size_t c = n - a < pool->chunk ? n - a : pool->chunk;
            rc = atk_put(s->stream, s->d_a, (const char *)raw + a * (size_t)bps, c * (size_t)bps);
            if (!rc) rc = atkgpu_unpack(s->stream, s->d_a, fmt, c, cre, cim, k, off_re, off_im, s->d_b, d_parts, NULL);
            if (!rc) rc = atkgpu_dc_pole(s->stream, s->d_b, c, dc_alpha, d_state, (double *)s->d_c);
            if (!rc) rc = atk_get(s->stream, out + a, s->d_b, c * sizeof(atkdsp_cf32));
            if (!rc) rc = atkgpu_sync(s->stream);
        }
        if (!rc) rc = atkgpu_d2h(s->stream, dc_state, d_state, sizeof(atkdsp_cf32));
        if (!rc) rc = atkgpu_sync(s->stream);
    } else {
        /* pipelined: pieces of one mean pass, each copied in, unpacked and
         * copied back while the next goes in */
        unpack_piece_ctx ctx = {s, fmt, bps, cre, cim, k, off_re, off_im, d_parts, want_mean ? d_sums : NULL};
        for (size_t a = 0; a < n && !rc; a += pool->chunk) {
            const size_t c = n - a < pool->chunk ? n - a : pool->chunk;
            const size_t np = (c + ATKGPU_MEAN_PASS - 1) / ATKGPU_MEAN_PASS;
            rc = atk_pipe(s->stream, s->stream2, (const char *)raw + a * (size_t)bps, (size_t)bps, s->d_a, out + a,
                          sizeof(atkdsp_cf32), s->d_b, c, ATKGPU_MEAN_PASS, unpack_piece, &ctx);
            if (!rc && want_mean) {
                double msum[2 * (ATKGPU_POOL_CHUNK / ATKGPU_MEAN_PASS)];
                rc = atkgpu_d2h(s->stream, msum, d_sums, 2 * np * sizeof(double));
                if (!rc) rc = atkgpu_sync(s->stream);
                for (size_t j = 0; j < np && !rc; ++j) {
                    sum_re += msum[2 * j];
                    sum_im += msum[2 * j + 1];
                }
            }
        }
    }
    atk_release(pool->slot, si);
    if (rc) return rc;
    if (want_mean) {
        /* the mean of the samples before `off` came out (src/convert.c:105-108) */
        const double inv = 1.0 / (double)n;
        if (mean_re) *mean_re = sum_re * inv + (double)off_re;
        if (mean_im) *mean_im = sum_im * inv + (double)off_im;
    }
    return (ptrdiff_t)n;
}

ptrdiff_t atkdsp_unpack(const void *raw, size_t nbytes, int fmt, atkdsp_cf32 *out, size_t out_cap,
                        float dc_alpha, atkdsp_cf32 *dc_state) {
    return unpack_run(raw, nbytes, fmt, out, out_cap, 0.0f, 0.0f, NULL, NULL, dc_alpha, dc_state);
}

ptrdiff_t atkdsp_unpack_dc(const void *raw, size_t nbytes, int fmt, atkdsp_cf32 *out, size_t out_cap,
                           float off_re, float off_im, double *mean_re, double *mean_im) {
    return unpack_run(raw, nbytes, fmt, out, out_cap, off_re, off_im, mean_re, mean_im, 0.0f, NULL);
}

/* ---- windows: create-time design in FP64, as src/fft.c:133-152 ------------------------ */
int atkdsp_window(int kind, size_t n, float *out) {
    if (!out || n == 0) return ATKDSP_E_ARG;
    if (kind < ATKDSP_WIN_RECT || kind > ATKDSP_WIN_BLACKMANHARRIS) return ATKDSP_E_ARG;
    if (n == 1) { out[0] = 1.0f; return ATKDSP_OK; }
    const double d = (double)(n - 1);
    for (size_t i = 0; i < n; ++i) {
        const double x = 2.0 * ATK_PI * (double)i / d;
        double w = 1.0;
        if (kind == ATKDSP_WIN_HANN) w = 0.5 - 0.5 * cos(x);
        else if (kind == ATKDSP_WIN_HAMMING) w = 0.54 - 0.46 * cos(x);
        else if (kind == ATKDSP_WIN_BLACKMANHARRIS)
            w = 0.35875 - 0.48829 * cos(x) + 0.14128 * cos(2 * x) - 0.01168 * cos(3 * x);
        out[i] = (float)w;
    }
    return ATKDSP_OK;
}

/* coherent gain and ENBW of a window: a design-time figure, FP64 on the host
 * as src/fft.c:433-444 */
int atkdsp_window_stats(const float *window, size_t n, double *coherent_gain, double *enbw_bins) {
    if (!window || n == 0) return ATKDSP_E_ARG;
    double s = 0.0, s2 = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const double w = (double)window[i];
        s += w;
        s2 += w * w;
    }
    if (coherent_gain) *coherent_gain = s / (double)n;
    if (enbw_bins) *enbw_bins = (s > 0.0) ? (double)n * s2 / (s * s) : 0.0;
    return ATKDSP_OK;
}

/* ---- display helpers and stitch (void: silent on bad arguments) ------------------------ */
void atkdsp_db_to_pixels(const float *db, size_t n, float lo, float hi, const uint32_t *lut256,
                         uint32_t *out) {
    if (!db || !lut256 || !out || n == 0) return;
    atk_pool *pool = atk_pool_get();
    if (!pool) return;
    const int si = atk_claim(pool->slot, pool->nslot);
    atk_slot *s = &pool->slot[si];
    uint32_t *d_lut = (uint32_t *)s->d_small;
    for (size_t a = 0; a < n; a += pool->chunk) {
        const size_t c = n - a < pool->chunk ? n - a : pool->chunk;
        if (atkgpu_h2d(s->stream, d_lut, lut256, 256 * sizeof(uint32_t)) ||
            atk_put(s->stream, s->d_a, db + a, c * sizeof(float)) ||
            atkgpu_db_to_pixels(s->stream, (const float *)s->d_a, c, lo, hi, d_lut, (uint32_t *)s->d_b) ||
            atk_get(s->stream, out + a, s->d_b, c * sizeof(uint32_t)) || atkgpu_sync(s->stream))
            break;
    }
    atk_release(pool->slot, si);
}

void atkdsp_decimate_max(const float *in, size_t n, float *out, size_t m) {
    if (!in || !out || m == 0 || n == 0) return;
    atk_pool *pool = atk_pool_get();
    if (!pool) return;
    /* lines beyond one slot buffer are not a display size; refuse quietly as
     * the reference does for bad arguments */
    if (n * sizeof(float) > pool->slot[0].bytes_a || m * sizeof(float) > pool->slot[0].bytes_b) return;
    const int si = atk_claim(pool->slot, pool->nslot);
    atk_slot *s = &pool->slot[si];
    if (!atkgpu_h2d(s->stream, s->d_a, in, n * sizeof(float)) &&
        !atkgpu_decimate_max(s->stream, (const float *)s->d_a, n, (float *)s->d_b, m) &&
        !atkgpu_d2h(s->stream, out, s->d_b, m * sizeof(float)))
        atkgpu_sync(s->stream);
    atk_release(pool->slot, si);
}

void atkdsp_stitch_max(const float *seg, size_t nseg, double seg_lo_hz, double seg_hz_per_bin,
                       double keep_lo_hz, double keep_hi_hz, double out_lo_hz, double out_hz_per_bin,
                       float *out, size_t nout) {
    if (!seg || !out || nout == 0 || seg_hz_per_bin <= 0.0 || out_hz_per_bin <= 0.0) return;
    atk_pool *pool = atk_pool_get();
    if (!pool) return;
    if (nseg * sizeof(float) > pool->slot[0].bytes_a || nout * sizeof(float) > pool->slot[0].bytes_b) return;
    const int si = atk_claim(pool->slot, pool->nslot);
    atk_slot *s = &pool->slot[si];
    if (!atkgpu_h2d(s->stream, s->d_a, seg, nseg * sizeof(float)) &&
        !atkgpu_h2d(s->stream, s->d_b, out, nout * sizeof(float)) &&
        !atkgpu_stitch_max(s->stream, (const float *)s->d_a, nseg, seg_lo_hz, seg_hz_per_bin, keep_lo_hz,
                           keep_hi_hz, out_lo_hz, out_hz_per_bin, (float *)s->d_b, nout) &&
        !atkgpu_d2h(s->stream, out, s->d_b, nout * sizeof(float)))
        atkgpu_sync(s->stream);
    atk_release(pool->slot, si);
}

/* ---- W1 / W2 over a host [nel][nchan][m] block ----------------------------------------- */
/* channels per pass so that nel*kc*m samples fit one slot buffer */
static unsigned chan_chunk(const atk_pool *pool, unsigned nel, size_t m) {
    const size_t per = (size_t)nel * (m ? m : 1) * sizeof(atkdsp_cf32);
    size_t kc = pool->slot[0].bytes_a / per;
    return kc < 1 ? 0u : (unsigned)(kc > 65535 ? 65535 : kc);
}

void atkdsp_pfb_cov_accumulate(const atkdsp_cf32 *out, unsigned nel, unsigned nchan, size_t m,
                               const uint8_t *gate, atkdsp_cf32 *R, uint32_t *counts) {
    if (!out || !R || !counts || nel < 1 || nchan < 1) return;
    atk_pool *pool = atk_pool_get();
    if (!pool) return;
    const unsigned kc = chan_chunk(pool, nel, m);
    if (!kc) return;
    const int si = atk_claim(pool->slot, pool->nslot);
    atk_slot *s = &pool->slot[si];
    const size_t pairs = (size_t)nel * nel;
    for (unsigned k0 = 0; k0 < nchan; k0 += kc) {
        const unsigned nk = nchan - k0 < kc ? nchan - k0 : kc;
        /* R and counts for the chunk go in buffer b; gate in small scratch */
        const size_t rbytes = (size_t)nk * pairs * sizeof(atkdsp_cf32);
        if (rbytes + nk * sizeof(uint32_t) > s->bytes_b || nk > ATK_SMALL_BYTES) break;
        atkdsp_cf32 *d_R = (atkdsp_cf32 *)s->d_b;
        uint32_t *d_cnt = (uint32_t *)((char *)s->d_b + rbytes);
        uint8_t *d_gate = gate ? (uint8_t *)s->d_small : NULL;
        /* the chunk's samples: nel rows of nk*m, pitch nchan*m */
        int rc = 0;
        for (unsigned e = 0; e < nel && !rc; ++e)
            rc = atkgpu_h2d(s->stream, (atkdsp_cf32 *)s->d_a + (size_t)e * nk * m,
                            out + ((size_t)e * nchan + k0) * m, (size_t)nk * m * sizeof(atkdsp_cf32));
        if (!rc) rc = atkgpu_h2d(s->stream, d_R, R + (size_t)k0 * pairs, rbytes);
        if (!rc) rc = atkgpu_h2d(s->stream, d_cnt, counts + k0, nk * sizeof(uint32_t));
        if (!rc && gate) rc = atkgpu_h2d(s->stream, d_gate, gate + k0, nk);
        if (!rc) rc = atkgpu_pfb_cov(s->stream, s->d_a, nel, nk, m, m, d_gate, d_R, d_cnt);
        if (!rc) rc = atkgpu_d2h(s->stream, R + (size_t)k0 * pairs, d_R, rbytes);
        if (!rc) rc = atkgpu_d2h(s->stream, counts + k0, d_cnt, nk * sizeof(uint32_t));
        if (!rc) rc = atkgpu_sync(s->stream);
        if (rc) break;
    }
    atk_release(pool->slot, si);
}

/* bookkeeping on the caller's host arrays: zero gated channels (src/channelizer.c:442-450) */
void atkdsp_pfb_cov_reset(atkdsp_cf32 *R, uint32_t *counts, unsigned nel, unsigned nchan,
                          const uint8_t *gate) {
    if (!R || !counts) return;
    for (unsigned k = 0; k < nchan; ++k) {
        if (gate && !gate[k]) continue;
        memset(R + (size_t)k * nel * nel, 0, (size_t)nel * nel * sizeof(atkdsp_cf32));
        counts[k] = 0;
    }
}

void atkdsp_pfb_channel_power(const atkdsp_cf32 *out, unsigned nel, unsigned nchan, size_t m,
                              float *pwr, float *mean_pwr) {
    if (!out || nel < 1 || nchan < 1) return;
    atk_pool *pool = atk_pool_get();
    if (!pool) return;
    const unsigned kc = chan_chunk(pool, nel, m);
    if (!kc) return;
    const int si = atk_claim(pool->slot, pool->nslot);
    atk_slot *s = &pool->slot[si];
    for (unsigned k0 = 0; k0 < nchan; k0 += kc) {
        const unsigned nk = nchan - k0 < kc ? nchan - k0 : kc;
        float *d_p = (float *)s->d_b;                      /* nel*nk */
        float *d_mp = d_p + (size_t)nel * nk;              /* nk     */
        int rc = 0;
        for (unsigned e = 0; e < nel && !rc; ++e)
            rc = atkgpu_h2d(s->stream, (atkdsp_cf32 *)s->d_a + (size_t)e * nk * m,
                            out + ((size_t)e * nchan + k0) * m, (size_t)nk * m * sizeof(atkdsp_cf32));
        if (!rc) rc = atkgpu_pfb_power(s->stream, s->d_a, nel, nk, m, m, d_p, d_mp);
        for (unsigned e = 0; e < nel && !rc && pwr; ++e)
            rc = atkgpu_d2h(s->stream, pwr + (size_t)e * nchan + k0, d_p + (size_t)e * nk, nk * sizeof(float));
        if (!rc && mean_pwr) rc = atkgpu_d2h(s->stream, mean_pwr + k0, d_mp, nk * sizeof(float));
        if (!rc) rc = atkgpu_sync(s->stream);
        if (rc) break;
    }
    atk_release(pool->slot, si);
}
