/* The polyphase filterbank handle (reference include/atkdsp.h:3191-3229,
 * src/channelizer.c:79-417), with its streaming history on the device.
 *
 * Per element the device keeps the stream s = hist(ov) ++ leftover(lc) ++ in
 * contiguous in one of two slab buffers; a call uploads leftover and input
 * behind the history, folds/transforms/rotates every whole hop on the device,
 * copies the next history (s[m hop, m hop + ov)) into the other slab buffer,
 * and keeps the < hop leftover samples in a host mirror, so a call that yields
 * no hop never touches the device (the reference stashes such calls,
 * src/channelizer.c:311-317). Inputs longer than a slab run as consecutive
 * sub-calls; every hop depends only on its own window, so the output does not
 * depend on the split (src/channelizer.c:214-217). */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "internal.h"

struct atkdsp_pfb {
    double fs, atten;
    unsigned nchan, os, K, max_nel, hop, ov;
    size_t L;
    double *h;                 /* prototype, FP64 (host) */
    float *d_taps;             /* (float)h, device, zero-padded to kos*hop */
    int kos;                   /* > 0: the fused kernel runs this shape */
    atkgpu_fftplan *plan;      /* nchan-point            */
    void *stream;
    void *stream2;             /* a single element's output comes back on it, range by range */
    size_t slab;               /* input samples per sub-call (multiple of hop) */
    size_t spitch;             /* samples per element in a slab buffer */
    atkdsp_cf32 *d_s[2];       /* slab buffers, max_nel * spitch */
    int cur;
    size_t hops_cap;           /* hops a sub-call can produce */
    atkdsp_cf32 *d_fold;       /* max_nel * hops_cap * nchan (FFT in place) */
    atkdsp_cf32 *d_out;        /* max_nel * nchan * hops_cap */
    atkdsp_cf32 *d_scr;        /* FFT plan scratch */
    atkdsp_cf32 *d_cal;        /* max_nel * nchan */
    unsigned cal_nel;
    atkdsp_cf32 *lo;           /* host mirror of the leftover, max_nel * hop */
    unsigned lc;
    int64_t b;                 /* global hop index */
};

/* modified Bessel I0 by its power series (as src/channelizer.c:33-43) */
static double pfb_i0(double x) {
    double sum = 1.0, term = 1.0;
    const double y = x * 0.5;
    for (int k = 1; k < 500; ++k) {
        term *= (y / (double)k);
        const double t2 = term * term;
        sum += t2;
        if (t2 < sum * 1e-17) break;
    }
    return sum;
}

/* Kaiser beta for a stopband attenuation (Kaiser's empirical formula, as
 * src/design.c:18-22) */
static double kaiser_beta(double a) {
    if (a > 50.0) return 0.1102 * (a - 8.7);
    if (a >= 21.0) return 0.5842 * pow(a - 21.0, 0.4) + 0.07886 * (a - 21.0);
    return 0.0;
}

/* Kaiser-windowed sinc lowpass at fc = 1/(2 nchan), unit DC gain, times nchan
 * (src/channelizer.c:79-99). Same FP64 operations in the same order, so the
 * taps equal the reference's. */
int atkgpu_pfb_prototype(double *h, unsigned nchan, unsigned K, double atten_db) {
    if (!h || nchan < 2 || K < 1) return ATKDSP_E_ARG;
    if (atten_db < 20.0) atten_db = 20.0;
    const size_t L = (size_t)nchan * K;
    const double fc = 1.0 / (2.0 * (double)nchan);
    const double Mid = (double)(L - 1) * 0.5;
    const double beta = kaiser_beta(atten_db);
    const double i0b = pfb_i0(beta);
    double sum = 0.0;
    for (int pass = 0; pass < 2; ++pass) {
        for (size_t n = 0; n < L; ++n) {
            const double t = (double)n - Mid;
            const double x = 2.0 * fc * t;
            const double sinc = (t == 0.0) ? 1.0 : sin(ATK_PI * x) / (ATK_PI * x);
            const double r = (Mid > 0.0) ? t / Mid : 0.0;
            const double w = pfb_i0(beta * sqrt(r * r < 1.0 ? 1.0 - r * r : 0.0)) / i0b;
            const double hv = 2.0 * fc * sinc * w;
            if (pass == 0) sum += hv;
            else h[n] = (hv / sum) * (double)nchan;
        }
        if (sum == 0.0) sum = 1.0;
    }
    return ATKDSP_OK;
}

static void pfb_free(atkdsp_pfb *p) {
    if (!p) return;
    if (p->stream) atkgpu_sync(p->stream);
    atkgpu_dfree(p->d_taps);
    atkgpu_dfree(p->d_s[0]);
    atkgpu_dfree(p->d_s[1]);
    atkgpu_dfree(p->d_fold);
    atkgpu_dfree(p->d_out);
    atkgpu_dfree(p->d_scr);
    atkgpu_dfree(p->d_cal);
    atkgpu_fftplan_destroy(p->plan);
    if (p->stream) atkgpu_stream_destroy(p->stream);
    if (p->stream2) atkgpu_stream_destroy(p->stream2);
    free(p->h);
    free(p->lo);
    free(p);
}

atkdsp_pfb *atkdsp_pfb_create(double fs, unsigned nchan, unsigned oversample, unsigned taps_per_poly,
                              double atten_db, unsigned max_nel) {
    if (!(fs > 0.0) || nchan < 2 || oversample < 1 || nchan % oversample != 0) return NULL;
    if (taps_per_poly < 2 || max_nel < 1) return NULL;
    if (atten_db < 20.0) atten_db = 20.0;
    if (atkgpu_init() != 0) return NULL;
    atkdsp_pfb *p = (atkdsp_pfb *)calloc(1, sizeof *p);
    if (!p) return NULL;
    p->fs = fs;
    p->atten = atten_db;
    p->nchan = nchan;
    p->os = oversample;
    p->K = taps_per_poly;
    p->max_nel = max_nel;
    p->hop = nchan / oversample;
    p->L = (size_t)nchan * taps_per_poly;
    p->ov = (unsigned)(p->L - p->hop);
    /* a sub-call's input: about 2M samples for one element, fewer per element
     * for many, at least 64 hops */
    size_t slab = ((size_t)1 << 23) / max_nel;
    if (slab > ((size_t)1 << 21)) slab = (size_t)1 << 21;
    if (slab < (size_t)64 * p->hop) slab = (size_t)64 * p->hop;
    slab = (slab + p->hop - 1) / p->hop * p->hop;
    p->slab = slab;
    p->spitch = p->ov + p->hop + slab;
    /* padded to 16 hops: the fused kernel's output rows start on 128-byte lines */
    p->hops_cap = (((slab + p->hop) / p->hop + 1) + 15) & ~(size_t)15;
    p->h = (double *)malloc(p->L * sizeof(double));
    p->lo = (atkdsp_cf32 *)calloc((size_t)max_nel * p->hop, sizeof(atkdsp_cf32));
    if (!p->h || !p->lo || atkgpu_stream_create(&p->stream) || atkgpu_stream_create(&p->stream2)) {
        pfb_free(p);
        return NULL;
    }
    p->plan = atkgpu_fftplan_create(nchan);
    const size_t nm = (size_t)max_nel * p->hops_cap * nchan;
    p->kos = atkgpu_pfb_fused_kos(nchan, oversample, taps_per_poly);
    const size_t ntaps = p->kos ? (size_t)p->kos * p->hop : p->L;
    p->d_taps = (float *)atkgpu_dmalloc(ntaps * sizeof(float));
    p->d_s[0] = (atkdsp_cf32 *)atkgpu_dmalloc((size_t)max_nel * p->spitch * sizeof(atkdsp_cf32));
    p->d_s[1] = (atkdsp_cf32 *)atkgpu_dmalloc((size_t)max_nel * p->spitch * sizeof(atkdsp_cf32));
    p->d_fold = (atkdsp_cf32 *)atkgpu_dmalloc(nm * sizeof(atkdsp_cf32));
    p->d_out = (atkdsp_cf32 *)atkgpu_dmalloc(nm * sizeof(atkdsp_cf32));
    const size_t scr = p->plan ? atkgpu_fftplan_scratch(p->plan, (size_t)max_nel * p->hops_cap) : 0;
    p->d_scr = (atkdsp_cf32 *)atkgpu_dmalloc((scr ? scr : 1) * sizeof(atkdsp_cf32));
    p->d_cal = (atkdsp_cf32 *)atkgpu_dmalloc((size_t)max_nel * nchan * sizeof(atkdsp_cf32));
    if (!p->plan || !p->d_taps || !p->d_s[0] || !p->d_s[1] || !p->d_fold || !p->d_out || !p->d_scr ||
        !p->d_cal) {
        pfb_free(p);
        return NULL;
    }
    atkgpu_pfb_prototype(p->h, nchan, taps_per_poly, p->atten);
    float *hf = (float *)calloc(ntaps, sizeof(float));
    if (!hf) { pfb_free(p); return NULL; }
    for (size_t q = 0; q < p->L; ++q) hf[q] = (float)p->h[q];
    int rc = atkgpu_h2d(p->stream, p->d_taps, hf, ntaps * sizeof(float));
    /* zero history (the reference callocs it) */
    if (!rc) rc = atkgpu_fill_u32(p->stream, p->d_s[0], 0, (size_t)max_nel * p->spitch * 2);
    if (!rc) rc = atkgpu_sync(p->stream);
    free(hf);
    if (rc) { pfb_free(p); return NULL; }
    p->cur = 0;
    p->lc = 0;
    p->b = 0;
    p->cal_nel = 0;
    return p;
}

void atkdsp_pfb_destroy(atkdsp_pfb *p) { pfb_free(p); }

void atkdsp_pfb_reset(atkdsp_pfb *p) {
    if (!p) return;
    atkgpu_fill_u32(p->stream, p->d_s[p->cur], 0, (size_t)p->max_nel * p->spitch * 2);
    atkgpu_sync(p->stream);
    memset(p->lo, 0, (size_t)p->max_nel * p->hop * sizeof(atkdsp_cf32));
    p->lc = 0;
    p->b = 0;
}

unsigned atkdsp_pfb_nchan(const atkdsp_pfb *p) { return p ? p->nchan : 0; }

double atkdsp_pfb_out_rate(const atkdsp_pfb *p) { return p ? (double)p->os * p->fs / (double)p->nchan : 0.0; }

size_t atkdsp_pfb_out_hops(const atkdsp_pfb *p, size_t n) {
    if (!p) return 0;
    return ((size_t)p->lc + n) / p->hop;
}

double atkdsp_pfb_group_delay(const atkdsp_pfb *p) { return p ? (double)(p->L - 1) * 0.5 : 0.0; }

int64_t atkdsp_pfb_block_origin(const atkdsp_pfb *p, int64_t hop_index) {
    return p ? hop_index * (int64_t)p->hop - (int64_t)p->ov : 0;
}

/* A config call: the table goes to the device buffer made at create. */
int atkdsp_pfb_set_cal(atkdsp_pfb *p, const atkdsp_cf32 *table, unsigned nel) {
    if (!p) return ATKDSP_E_ARG;
    if (!table) { p->cal_nel = 0; return ATKDSP_OK; }
    if (nel < 1 || nel > p->max_nel) return ATKDSP_E_ARG;
    int rc = atkgpu_h2d(p->stream, p->d_cal, table, (size_t)nel * p->nchan * sizeof(atkdsp_cf32));
    if (!rc) rc = atkgpu_sync(p->stream);
    if (rc) return ATKDSP_E_FFT;
    p->cal_nel = nel;
    return ATKDSP_OK;
}

int atkdsp_pfb_describe(const atkdsp_pfb *p, char *buf, size_t cap) {
    if (!p || !buf || cap == 0) return ATKDSP_E_ARG;
    return snprintf(buf, cap,
                    "PFB %.6g MSPS: %u chan x%u oversample, %.6g kHz spacing, %.6g kHz out, "
                    "%zu-tap proto (%u/poly), %.0f dB, <=%u el; hip gfx950 %s, fp32 fold and fft",
                    p->fs / 1e6, p->nchan, p->os, p->fs / p->nchan / 1e3, atkdsp_pfb_out_rate(p) / 1e3, p->L,
                    p->K, p->atten, p->max_nel, p->kos ? "fused kernel" : "general kernels");
}

/* One sub-call over ins[e][pos, pos+c): writes its m_s hops at column i_done
 * of the caller's [nel][nchan][m] output. */
/* One element's sub-call through the fused kernel in ranges of R hops, each
 * range's output (nchan rows of at most R hops, one staging piece) coming
 * back while the next range's input goes in (atk_pipe_ranges). Every hop
 * depends only on its own window, so the ranges give the one-launch output. */
typedef struct {
    atkdsp_pfb *p;
    atkdsp_cf32 *S;            /* the slab: window coordinate 0 */
    size_t ms, R, R16, avail;
    atkdsp_cf32 *out;          /* the call's first output hop of this sub-call */
    size_t m;                  /* output row pitch, hops */
} pfb_pipe_ctx;

static int pfb_pipe_desc(void *ctx, size_t r, atk_range *g) {
    const pfb_pipe_ctx *x = (const pfb_pipe_ctx *)ctx;
    const size_t h0 = r * x->R, h1 = h0 + x->R < x->ms ? h0 + x->R : x->ms;
    const size_t hop = x->p->hop, lc = x->p->lc;
    /* the range's last window ends at sample h1*hop + ov of the slab; the
     * call's input starts at ov + lc */
    g->need = (h1 * hop > lc ? h1 * hop - lc : 0) * sizeof(atkdsp_cf32);
    g->src = x->p->d_fold + (r & 1) * (size_t)x->p->nchan * x->R;
    g->dst = (char *)(x->out + h0);
    g->dpitch = x->m * sizeof(atkdsp_cf32);
    g->width = (h1 - h0) * sizeof(atkdsp_cf32);
    g->rows = x->p->nchan;
    return 0;
}

static int pfb_pipe_run(void *ctx, void *s, size_t r) {
    const pfb_pipe_ctx *x = (const pfb_pipe_ctx *)ctx;
    atkdsp_pfb *p = x->p;
    const size_t h0 = r * x->R, h = (h0 + x->R < x->ms ? h0 + x->R : x->ms) - h0;
    atkdsp_cf32 *o = p->d_out + (r & 1) * (size_t)p->nchan * x->R16;
    int rc = atkgpu_pfb_fused(s, x->S + h0 * p->hop, p->spitch, (size_t)p->ov + x->avail - h0 * p->hop, 1, h,
                              (long long)p->b + (long long)h0, p->nchan, p->os, (unsigned)p->kos, p->d_taps,
                              atkgpu_fftplan_twiddles(p->plan), p->cal_nel ? (const void *)p->d_cal : NULL,
                              p->cal_nel, o, x->R16, 0);
    if (!rc)
        rc = atkgpu_d2d_2d(s, p->d_fold + (r & 1) * (size_t)p->nchan * x->R, h * sizeof(atkdsp_cf32), o,
                           x->R16 * sizeof(atkdsp_cf32), h * sizeof(atkdsp_cf32), p->nchan);
    return rc;
}

static int pfb_sub(atkdsp_pfb *p, const atkdsp_cf32 *const *ins, unsigned nel, size_t pos, size_t c,
                   atkdsp_cf32 *out, size_t m, size_t i_done, size_t *made) {
    const unsigned M = p->nchan, hop = p->hop, ov = p->ov;
    const size_t avail = (size_t)p->lc + c;
    const size_t ms = avail / hop;
    *made = ms;
    if (ms == 0) {
        for (unsigned e = 0; e < nel; ++e)
            if (c) memcpy(p->lo + (size_t)e * hop + p->lc, ins[e] + pos, c * sizeof(atkdsp_cf32));
        p->lc = (unsigned)avail;
        return 0;
    }
    atkdsp_cf32 *S = p->d_s[p->cur];
    /* device output pitch: whole 128-byte lines for the fused kernel */
    const size_t mp = p->kos ? (ms + 15) & ~(size_t)15 : ms;
    int rc = 0;
    /* one element with more hops than a staging piece of output holds: in ranges */
    const size_t R = (ATK_PIPE_PIECE / ((size_t)M * sizeof(atkdsp_cf32))) & ~(size_t)15;
    if (p->kos && nel == 1 && R >= 16 && ms > R && 2 * R <= (size_t)p->max_nel * p->hops_cap) {
        if (p->lc) rc = atkgpu_h2d(p->stream, S + ov, p->lo, p->lc * sizeof(atkdsp_cf32));
        pfb_pipe_ctx x = {p, S, ms, R, R, avail, out + i_done, m};
        if (!rc) rc = atk_pipe_ranges(p->stream, p->stream2, S + ov + p->lc, ins[0] + pos,
                                      c * sizeof(atkdsp_cf32), (ms + R - 1) / R, pfb_pipe_desc, pfb_pipe_run, &x);
        goto history;
    }
    for (unsigned e = 0; e < nel && !rc; ++e) {
        atkdsp_cf32 *se = S + (size_t)e * p->spitch;
        if (p->lc) rc = atkgpu_h2d(p->stream, se + ov, p->lo + (size_t)e * hop, p->lc * sizeof(atkdsp_cf32));
        if (!rc && c) rc = atk_put(p->stream, se + ov + p->lc, ins[e] + pos, c * sizeof(atkdsp_cf32));
    }
    if (p->kos) {
        if (!rc) rc = atkgpu_pfb_fused(p->stream, S, p->spitch, (size_t)ov + avail, nel, ms, (long long)p->b, M,
                                       p->os, (unsigned)p->kos, p->d_taps, atkgpu_fftplan_twiddles(p->plan),
                                       p->cal_nel ? (const void *)p->d_cal : NULL, p->cal_nel, p->d_out, mp, 0);
    } else {
        if (!rc) rc = atkgpu_pfb_fold(p->stream, S, p->spitch, nel, ms, M, hop, p->K, p->d_taps, p->d_fold);
        if (!rc) rc = atkgpu_fft_batch(p->stream, p->plan, p->d_fold, M, NULL, (size_t)nel * ms, 0, p->d_fold,
                                       p->d_scr);
        if (!rc) rc = atkgpu_pfb_rotate_store(p->stream, p->d_fold, nel, ms, M, hop, (long long)p->b,
                                              atkgpu_fftplan_twiddles(p->plan),
                                              p->cal_nel ? (const void *)p->d_cal : NULL, p->cal_nel, p->d_out);
    }
    /* the rows packed on the device first (at HBM speed): one packed
     * device-to-host copy instead of nel * M short rows */
    const atkdsp_cf32 *packed = p->d_out;
    if (!rc && mp != ms) {
        rc = atkgpu_d2d_2d(p->stream, p->d_fold, ms * sizeof(atkdsp_cf32), p->d_out, mp * sizeof(atkdsp_cf32),
                           ms * sizeof(atkdsp_cf32), (size_t)nel * M);
        packed = p->d_fold;
    }
    if (!rc) rc = atk_get_rows(p->stream, out + i_done, m * sizeof(atkdsp_cf32), packed, ms * sizeof(atkdsp_cf32),
                               (size_t)nel * M);
history:;
    /* next history: s[ms hop, ms hop + ov) into the other slab buffer */
    atkdsp_cf32 *T = p->d_s[p->cur ^ 1];
    for (unsigned e = 0; e < nel && !rc; ++e)
        rc = atkgpu_d2d(p->stream, T + (size_t)e * p->spitch, S + (size_t)e * p->spitch + ms * hop,
                        (size_t)ov * sizeof(atkdsp_cf32));
    if (rc) return ATKDSP_E_FFT;
    /* next leftover: s[ms hop + ov, ov + avail), all inside this sub-call's input */
    const size_t rem = avail - ms * (size_t)hop;
    const size_t from = ms * (size_t)hop - p->lc;            /* index into the chunk */
    for (unsigned e = 0; e < nel; ++e)
        if (rem) memcpy(p->lo + (size_t)e * hop, ins[e] + pos + from, rem * sizeof(atkdsp_cf32));
    p->cur ^= 1;
    p->lc = (unsigned)rem;
    p->b += (int64_t)ms;
    return 0;
}

ptrdiff_t atkdsp_pfb_process_n(atkdsp_pfb *p, const atkdsp_cf32 *const *ins, unsigned nel, size_t n,
                               atkdsp_cf32 *out, size_t cap) {
    if (!p || !ins || nel < 1 || nel > p->max_nel) return ATKDSP_E_ARG;
    const size_t avail = (size_t)p->lc + n;
    const size_t m = avail / p->hop;
    if (m == 0) {
        for (unsigned e = 0; e < nel; ++e)
            if (n) memcpy(p->lo + (size_t)e * p->hop + p->lc, ins[e], n * sizeof(atkdsp_cf32));
        p->lc = (unsigned)avail;
        return 0;
    }
    if (!out || cap < (size_t)nel * p->nchan * m) return ATKDSP_E_CAP;
    for (unsigned e = 0; e < nel; ++e)
        if (!ins[e]) return ATKDSP_E_ARG;
    size_t pos = 0, done = 0;
    while (pos < n) {
        const size_t c = n - pos < p->slab ? n - pos : p->slab;
        size_t made = 0;
        const int rc = pfb_sub(p, ins, nel, pos, c, out, m, done, &made);
        if (rc) return rc;
        done += made;
        pos += c;
    }
    if (atkgpu_sync(p->stream)) return ATKDSP_E_FFT;
    return (ptrdiff_t)m;
}

ptrdiff_t atkdsp_pfb_process(atkdsp_pfb *p, const atkdsp_cf32 *in, size_t n, atkdsp_cf32 *out, size_t cap) {
    const atkdsp_cf32 *ins[1] = {in};
    return atkdsp_pfb_process_n(p, ins, 1, n, out, cap);
}
