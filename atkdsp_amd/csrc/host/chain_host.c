/* The device-resident multi-stream chain (include/atkgpu.h). Every stream has
 * the same shape and each run takes one block of n samples per stream, so one
 * launch per stage covers all streams. Stream state (kept prefix, filterbank
 * hop counter and leftover count, display tail, DC offset, trace) lives on
 * the device; the host only tracks the shared counters. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "internal.h"
#include "../../../include/atkgpu.h"

int atkgpu_pfb_prototype(double *h, unsigned nchan, unsigned K, double atten_db);

struct atkgpu_chain {
    atkgpu_chain_cfg cfg;
    unsigned S;
    unsigned M, os, K, hop, ov;
    int kos;
    size_t keep, spitch, bps;
    void *stream;
    void *ev[4];
    int timing;                /* record the timing events in each run */
    int timed;                 /* the last run recorded them */
    atkdsp_cf32 *d_rows[2];
    int cur;
    unsigned char *d_raw;
    float *d_taps, *d_win, *d_off, *d_partial, *d_line, *d_dec, *d_trace;
    double *d_upart, *d_mean;
    atkgpu_fftplan *plan_m, *plan_n;
    atkdsp_cf32 *d_out;
    size_t m_cap;
    int uparts, nfb_max;
    /* shared counters (every stream sees the same block sizes) */
    unsigned lc;
    long long b;
    size_t dlen;
    size_t last_m;
    long long last_frames;
    int ran;
};

int atkgpu_version(void) { return 1; }

static void chain_free(atkgpu_chain *c) {
    if (!c) return;
    if (c->stream) atkgpu_sync(c->stream);
    atkgpu_dfree(c->d_rows[0]); atkgpu_dfree(c->d_rows[1]); atkgpu_dfree(c->d_raw);
    atkgpu_dfree(c->d_taps); atkgpu_dfree(c->d_win); atkgpu_dfree(c->d_off); atkgpu_dfree(c->d_partial);
    atkgpu_dfree(c->d_line); atkgpu_dfree(c->d_dec); atkgpu_dfree(c->d_trace); atkgpu_dfree(c->d_upart);
    atkgpu_dfree(c->d_mean); atkgpu_dfree(c->d_out);
    atkgpu_fftplan_destroy(c->plan_m); atkgpu_fftplan_destroy(c->plan_n);
    for (int i = 0; i < 4; ++i) atkgpu_event_destroy(c->ev[i]);
    if (c->stream) atkgpu_stream_destroy(c->stream);
    free(c);
}

static int fill_nan(void *stream, float *p, size_t n) { return atkgpu_fill_u32(stream, p, 0x7fc00000u, n); }

atkgpu_chain *atkgpu_chain_create(const atkgpu_chain_cfg *cfg, unsigned S) {
    if (!cfg || S < 1) return NULL;
    const int bps = atkdsp_bytes_per_sample(cfg->fmt);
    if (!bps || !(cfg->fs > 0) || cfg->block_max < 1) return NULL;
    if (cfg->nfft != 1024 && cfg->nfft != 2048 && cfg->nfft != 4096) return NULL;
    if (cfg->fft_hop == 0 || cfg->fft_hop > cfg->nfft || cfg->line_cols < 1) return NULL;
    if (cfg->detector < 0 || cfg->detector > 2) return NULL;
    const int kos = atkgpu_pfb_fused_kos(cfg->nchan, cfg->oversample, cfg->taps_per_poly);
    if (!kos || cfg->taps_per_poly < 2) return NULL;            /* the fused shapes only */
    if (atkgpu_init() != 0) return NULL;
    atkgpu_chain *c = (atkgpu_chain *)calloc(1, sizeof *c);
    if (!c) return NULL;
    c->timing = 1;
    c->cfg = *cfg;
    c->S = S;
    c->M = cfg->nchan;
    c->os = cfg->oversample;
    c->K = cfg->taps_per_poly;
    c->hop = c->M / c->os;
    c->ov = c->M * c->K - c->hop;
    c->kos = kos;
    c->bps = (size_t)bps;
    size_t keep = (size_t)c->ov + c->hop;
    if (keep < cfg->nfft) keep = cfg->nfft;
    c->keep = (keep + 1) & ~(size_t)1;
    c->spitch = (c->keep + cfg->block_max + 63) & ~(size_t)63;
    /* output rows are padded to whole 128-byte lines: the filterbank's
     * per-channel segments then fill whole 32-byte sectors */
    c->m_cap = (((cfg->block_max + c->hop) / c->hop + 1) + 15) & ~(size_t)15;
    c->uparts = 256;
    c->nfb_max = 64;
    if (atkgpu_stream_create(&c->stream)) { chain_free(c); return NULL; }
    for (int i = 0; i < 4; ++i)
        if (atkgpu_event_create(&c->ev[i])) { chain_free(c); return NULL; }
    const size_t rows = (size_t)S * c->spitch;
    c->d_rows[0] = (atkdsp_cf32 *)atkgpu_dmalloc(rows * sizeof(atkdsp_cf32));
    c->d_rows[1] = (atkdsp_cf32 *)atkgpu_dmalloc(rows * sizeof(atkdsp_cf32));
    c->d_raw = (unsigned char *)atkgpu_dmalloc((size_t)S * cfg->block_max * c->bps);
    c->d_taps = (float *)atkgpu_dmalloc((size_t)kos * c->hop * sizeof(float));
    c->d_win = (float *)atkgpu_dmalloc(cfg->nfft * sizeof(float));
    c->d_off = (float *)atkgpu_dmalloc((size_t)S * 2 * sizeof(float));
    c->d_partial = (float *)atkgpu_dmalloc((size_t)S * c->nfb_max * cfg->nfft * sizeof(float));
    c->d_line = (float *)atkgpu_dmalloc((size_t)S * cfg->nfft * sizeof(float));
    c->d_dec = (float *)atkgpu_dmalloc((size_t)S * cfg->line_cols * sizeof(float));
    c->d_trace = cfg->trace_bins ? (float *)atkgpu_dmalloc((size_t)S * cfg->trace_bins * sizeof(float)) : NULL;
    c->d_upart = (double *)atkgpu_dmalloc((size_t)S * c->uparts * 2 * sizeof(double));
    c->d_mean = (double *)atkgpu_dmalloc((size_t)S * 2 * sizeof(double));
    c->d_out = (atkdsp_cf32 *)atkgpu_dmalloc((size_t)S * c->M * c->m_cap * sizeof(atkdsp_cf32));
    c->plan_m = atkgpu_fftplan_create(c->M);
    c->plan_n = atkgpu_fftplan_create(cfg->nfft);
    if (!c->d_rows[0] || !c->d_rows[1] || !c->d_raw || !c->d_taps || !c->d_win || !c->d_off ||
        !c->d_partial || !c->d_line || !c->d_dec || (cfg->trace_bins && !c->d_trace) || !c->d_upart ||
        !c->d_mean || !c->d_out || !c->plan_m || !c->plan_n) {
        chain_free(c);
        return NULL;
    }
    /* taps: the reference prototype, zero-padded to kos rows of hop */
    const size_t L = (size_t)c->M * c->K;
    double *h = (double *)malloc(L * sizeof(double));
    float *hf = (float *)calloc((size_t)kos * c->hop, sizeof(float));
    float *w = (float *)malloc(cfg->nfft * sizeof(float));
    int rc = (!h || !hf || !w) ? ATKDSP_E_NOMEM : 0;
    if (!rc) rc = atkgpu_pfb_prototype(h, c->M, c->K, cfg->atten_db);
    if (!rc) {
        for (size_t q = 0; q < L; ++q) hf[q] = (float)h[q];
        rc = atkdsp_window(cfg->window, cfg->nfft, w);
    }
    if (!rc) rc = atkgpu_h2d(c->stream, c->d_taps, hf, (size_t)kos * c->hop * sizeof(float));
    if (!rc) rc = atkgpu_h2d(c->stream, c->d_win, w, cfg->nfft * sizeof(float));
    if (!rc) rc = atkgpu_sync(c->stream);
    free(h); free(hf); free(w);
    if (rc) { chain_free(c); return NULL; }
    atkgpu_chain_reset(c);
    return c;
}

void atkgpu_chain_destroy(atkgpu_chain *c) { chain_free(c); }

void atkgpu_chain_reset(atkgpu_chain *c) {
    if (!c) return;
    atkgpu_fill_u32(c->stream, c->d_rows[0], 0, (size_t)c->S * c->spitch * 2);
    atkgpu_fill_u32(c->stream, c->d_off, 0, (size_t)c->S * 2);
    if (c->d_trace) fill_nan(c->stream, c->d_trace, (size_t)c->S * c->cfg.trace_bins);
    atkgpu_sync(c->stream);
    c->cur = 0;
    c->lc = 0;
    c->b = 0;
    c->dlen = 0;
    c->last_m = 0;
    c->last_frames = 0;
    c->ran = 0;
}

void *atkgpu_chain_input(atkgpu_chain *c, unsigned s) {
    if (!c || s >= c->S) return NULL;
    return c->d_raw + (size_t)s * c->cfg.block_max * c->bps;
}

int atkgpu_chain_upload(atkgpu_chain *c, unsigned s, const void *raw, size_t nbytes) {
    if (!c || s >= c->S || !raw || nbytes > c->cfg.block_max * c->bps) return ATKDSP_E_ARG;
    return atk_put(c->stream, atkgpu_chain_input(c, s), raw, nbytes) ? ATKDSP_E_FFT : ATKDSP_OK;
}

ptrdiff_t atkgpu_chain_run(atkgpu_chain *c, size_t n) {
    if (!c || n < 1 || n > c->cfg.block_max) return ATKDSP_E_ARG;
    const atkgpu_chain_cfg *g = &c->cfg;
    atkdsp_cf32 *R = c->d_rows[c->cur], *T = c->d_rows[c->cur ^ 1];
    const size_t keep = c->keep;
    const size_t avail = (size_t)c->lc + n;
    const size_t m = avail / c->hop;
    const size_t d = c->dlen + n;
    const long long frames = d >= g->nfft ? (long long)((d - g->nfft) / g->fft_hop + 1) : 0;
    /* about 512 display blocks in all: the 4096-point reduction holds two
     * blocks per CU, so on 256 CUs every block is resident at once and each
     * transforms a long run of frames with its twiddles in registers
     * (measured 0.762 ms a step against 0.792 ms with 2048 blocks) */
    int nfb = (int)((512 + c->S - 1) / c->S);
    if (nfb > c->nfb_max) nfb = c->nfb_max;
    if (nfb < 1) nfb = 1;
    const int tm = c->timing;
    int rc = tm ? atkgpu_event_record(c->ev[0], c->stream) : 0;
    /* the display reads the raw blocks and unpacks them into the rows on the
     * way (with the block mean); a block that completes no frame is
     * unpacked on its own */
    if (!rc && frames)
        rc = atkgpu_chain_display_unpack(c->stream, R + (keep - c->dlen), c->spitch, c->d_raw, g->block_max * c->bps,
                                         g->fmt, n, c->dlen, c->d_off, c->S, (size_t)frames, g->fft_hop, g->nfft, nfb,
                                         c->d_win, atkgpu_fftplan_twiddles(c->plan_n), g->detector, c->d_partial,
                                         c->d_line, c->d_upart, c->d_mean, g->dc_mode == 1);
    else if (!rc)
        rc = atkgpu_chain_unpack(c->stream, c->d_raw, g->block_max * c->bps, g->fmt, n, R + keep, c->spitch,
                                 c->d_off, c->d_upart, c->uparts, c->d_mean, g->dc_mode == 1, c->S);
    if (!rc && tm) rc = atkgpu_event_record(c->ev[1], c->stream);
    if (!rc && m)
        rc = atkgpu_pfb_fused(c->stream, R + (keep - c->ov - c->lc), c->spitch, (size_t)c->ov + avail, c->S, m,
                              c->b, c->M, c->os, (unsigned)c->kos, c->d_taps, atkgpu_fftplan_twiddles(c->plan_m),
                              NULL, 0, c->d_out, (m + 15) & ~(size_t)15, 0);
    if (!rc && tm) rc = atkgpu_event_record(c->ev[2], c->stream);
    if (!rc && frames) {
        const double seg_step = g->fs / (double)g->line_cols;
        rc = atkgpu_chain_post(c->stream, c->d_line, c->S, g->nfft, g->line_cols, c->d_dec, c->d_trace,
                               g->trace_bins, g->fc0_hz - g->fs / 2, g->fc_step_hz, seg_step, g->trace_lo_hz,
                               g->trace_hz_per_bin);
    }
    if (!rc) rc = atkgpu_chain_keep(c->stream, R + n, c->spitch, T, keep, c->S);
    if (!rc && tm) rc = atkgpu_event_record(c->ev[3], c->stream);
    if (rc) return ATKDSP_E_FFT;
    c->cur ^= 1;
    c->lc = (unsigned)(avail - m * c->hop);
    c->b += (long long)m;
    c->dlen = d - (size_t)frames * g->fft_hop;
    c->last_m = m;
    c->last_frames = frames;
    c->ran = 1;
    c->timed = tm;
    return (ptrdiff_t)m;
}

int atkgpu_chain_sync(atkgpu_chain *c) {
    if (!c) return ATKDSP_E_ARG;
    return atkgpu_sync(c->stream) ? ATKDSP_E_FFT : ATKDSP_OK;
}

int atkgpu_chain_set_timing(atkgpu_chain *c, int on) {
    if (!c) return ATKDSP_E_ARG;
    c->timing = on != 0;
    return ATKDSP_OK;
}

int atkgpu_chain_timing(atkgpu_chain *c, float *ms_run, float *ms_pfb, float *ms_disp) {
    if (!c || !c->ran || !c->timed) return ATKDSP_E_ARG;
    /* ev0 | display with the unpack, block mean | ev1 | filterbank | ev2 | decimate, stitch, keep | ev3 */
    float a = 0, b = 0, d = 0;
    if (atkgpu_event_elapsed_ms(c->ev[0], c->ev[3], &a) || atkgpu_event_elapsed_ms(c->ev[1], c->ev[2], &b) ||
        atkgpu_event_elapsed_ms(c->ev[0], c->ev[1], &d))
        return ATKDSP_E_FFT;
    if (ms_run) *ms_run = a;
    if (ms_pfb) *ms_pfb = b;
    if (ms_disp) *ms_disp = d;
    return ATKDSP_OK;
}

ptrdiff_t atkgpu_chain_read_pfb(atkgpu_chain *c, unsigned s, atkdsp_cf32 *out, size_t cap) {
    if (!c || s >= c->S || !out) return ATKDSP_E_ARG;
    const size_t cnt = (size_t)c->M * c->last_m, mp = (c->last_m + 15) & ~(size_t)15;
    if (cap < cnt) return ATKDSP_E_CAP;
    int rc = cnt ? atkgpu_d2h_2d(c->stream, out, c->last_m * sizeof(atkdsp_cf32), c->d_out + (size_t)s * c->M * mp,
                                 mp * sizeof(atkdsp_cf32), c->last_m * sizeof(atkdsp_cf32), c->M)
                 : 0;
    if (!rc) rc = atkgpu_sync(c->stream);
    return rc ? ATKDSP_E_FFT : (ptrdiff_t)c->last_m;
}

int atkgpu_chain_read_line(atkgpu_chain *c, unsigned s, float *line_db, float *cols, float *trace) {
    if (!c || s >= c->S) return ATKDSP_E_ARG;
    const atkgpu_chain_cfg *g = &c->cfg;
    int rc = 0;
    if (line_db) rc = atkgpu_d2h(c->stream, line_db, c->d_line + (size_t)s * g->nfft, g->nfft * sizeof(float));
    if (!rc && cols) rc = atkgpu_d2h(c->stream, cols, c->d_dec + (size_t)s * g->line_cols, g->line_cols * sizeof(float));
    if (!rc && trace && c->d_trace)
        rc = atkgpu_d2h(c->stream, trace, c->d_trace + (size_t)s * g->trace_bins, g->trace_bins * sizeof(float));
    if (!rc) rc = atkgpu_sync(c->stream);
    return rc ? ATKDSP_E_FFT : ATKDSP_OK;
}

int atkgpu_chain_read_mean(atkgpu_chain *c, unsigned s, double *mean2) {
    if (!c || s >= c->S || !mean2) return ATKDSP_E_ARG;
    int rc = atkgpu_d2h(c->stream, mean2, c->d_mean + 2 * (size_t)s, 2 * sizeof(double));
    if (!rc) rc = atkgpu_sync(c->stream);
    return rc ? ATKDSP_E_FFT : ATKDSP_OK;
}

long long atkgpu_chain_frames(const atkgpu_chain *c) { return c ? c->last_frames : 0; }
