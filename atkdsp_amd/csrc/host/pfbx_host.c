/* The filterbank in matrix-core form over S streams of one shape
 * (include/atkgpu.h atkgpu_pfbx_*), beside the plain FP32 form on the same
 * plumbing so the two are measured alike.
 *
 *   FP32 form   raw block -> ragged unpack (fixed offset, bit-exact with
 *               unpack_dc) into cf32 rows -> fused FP32 filterbank -> the
 *               kept prefix to the other row.
 *   MX forms    raw block straight into raw rows (the kept raw history sits
 *               in front of it) -> the MFMA filterbank -> the kept raw
 *               prefix to the other row. No cf32 copy of the input exists.
 *   FP32 raw    the MX forms' raw rows -> the fused FP32 filterbank, which
 *               converts each sample as it loads it (bit-identical to the
 *               FP32 form) -> the kept raw prefix to the other row.
 *
 * Streaming state is the reference's (src/channelizer.c:305-408): a row holds
 * keep = ov + hop samples of history in front of each new block; hop i's
 * window starts keep - ov - lc samples into the row; lc and the int64 hop
 * counter are shared by all streams (every run gives each stream n samples). */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "internal.h"
#include "../../../include/atkgpu.h"

int atkgpu_pfb_prototype(double *h, unsigned nchan, unsigned K, double atten_db);

struct atkgpu_pfbx {
    atkgpu_pfbx_cfg cfg;
    unsigned S, M, os, K, hop, ov;
    int kos, bps, mx, rawrows;      /* rawrows: raw rows (MX forms, FP32 raw), else cf32 rows */
    size_t keep, spitch, m_cap;
    void *stream;
    void *ev[5];
    unsigned char *d_raw;      /* FP32 form: [S][block_max*bps] upload blocks */
    unsigned char *d_rows[2];  /* FP32 form: cf32 rows; MX and FP32 raw forms: raw rows; [S][spitch] samples */
    int cur;
    float *d_taps, *d_tapsum;
    void *d_atab, *d_offs;
    float scale;
    atkgpu_fftplan *plan;
    atkdsp_cf32 *d_out;        /* [S][M][m_cap] */
    atkdsp_cf32 *d_R;          /* coherent: [M][S][S] */
    uint32_t *d_cnt;           /* coherent: [M] */
    float *d_pwr, *d_mean;     /* coherent: [S][M], [M] */
    void *desc_h[2], *desc_d[2];
    size_t desc_n[2];          /* block size each descriptor set was built for (0: none) */
    unsigned lc;
    long long b;
    size_t last_m;
    int ran;
};

static size_t round16(size_t x) { return (x + 15) & ~(size_t)15; }

void atkgpu_pfbx_destroy(atkgpu_pfbx *v) {
    if (!v) return;
    if (v->stream) atkgpu_sync(v->stream);
    atkgpu_dfree(v->d_raw);
    atkgpu_dfree(v->d_rows[0]);
    atkgpu_dfree(v->d_rows[1]);
    atkgpu_dfree(v->d_taps);
    atkgpu_dfree(v->d_tapsum);
    atkgpu_dfree(v->d_atab);
    atkgpu_dfree(v->d_offs);
    atkgpu_dfree(v->d_out);
    for (int i = 0; i < 2; ++i) atkgpu_hfree(v->desc_h[i]);
    atkgpu_fftplan_destroy(v->plan);
    atkgpu_dfree(v->d_R);
    atkgpu_dfree(v->d_cnt);
    atkgpu_dfree(v->d_pwr);
    atkgpu_dfree(v->d_mean);
    for (int i = 0; i < 5; ++i) atkgpu_event_destroy(v->ev[i]);
    if (v->stream) atkgpu_stream_destroy(v->stream);
    free(v);
}

int atkgpu_pfbx_cov_reset(atkgpu_pfbx *v) {
    if (!v || !v->cfg.coherent) return ATKDSP_E_ARG;
    int rc = atkgpu_pfb_cov_reset(v->stream, v->d_R, v->d_cnt, v->S, v->M, NULL);
    if (!rc) rc = atkgpu_sync(v->stream);
    return rc ? ATKDSP_E_FFT : ATKDSP_OK;
}

int atkgpu_pfbx_reset(atkgpu_pfbx *v) {
    if (!v) return ATKDSP_E_ARG;
    v->lc = 0;
    v->b = 0;
    v->cur = 0;
    v->last_m = 0;
    v->ran = 0;
    /* history starts at zero (the reference callocs it). Raw rows cannot
     * encode cu8's zero (127.5 counts): the matrix-core kernel is told how
     * many window samples precede the stream's start and treats them as 0 */
    int rc = 0;
    const size_t row_bytes = (size_t)v->S * v->spitch * (v->rawrows ? (size_t)v->bps : sizeof(atkdsp_cf32));
    for (int i = 0; i < 2 && !rc; ++i) rc = atkgpu_fill_u32(v->stream, v->d_rows[i], 0u, row_bytes / 4);
    if (!rc && v->cfg.coherent) rc = atkgpu_pfb_cov_reset(v->stream, v->d_R, v->d_cnt, v->S, v->M, NULL);
    if (!rc) rc = atkgpu_sync(v->stream);
    return rc ? ATKDSP_E_FFT : ATKDSP_OK;
}

atkgpu_pfbx *atkgpu_pfbx_create(const atkgpu_pfbx_cfg *cfg, unsigned S) {
    if (!cfg || S < 1 || !(cfg->fs > 0) || cfg->block_max < 8 || (cfg->block_max & 7)) return NULL;
    const int bps = atkdsp_bytes_per_sample(cfg->fmt);
    const int kos = atkgpu_pfb_fused_kos(cfg->nchan, cfg->oversample, cfg->taps_per_poly);
    if (!bps || !kos || cfg->taps_per_poly < 2 || cfg->form < ATKGPU_FORM_FP32 || cfg->form > ATKGPU_FORM_FP32RAW)
        return NULL;
    const int mx = cfg->form != ATKGPU_FORM_FP32 && cfg->form != ATKGPU_FORM_FP32MM &&
                   cfg->form != ATKGPU_FORM_FP32RAW;
    const int rawrows = mx || cfg->form == ATKGPU_FORM_FP32RAW;
    if (cfg->form == ATKGPU_FORM_FP32RAW &&
        !atkgpu_pfb_fused_raw_supported(cfg->nchan, cfg->oversample, (unsigned)kos, cfg->fmt))
        return NULL;
    if (cfg->form == ATKGPU_FORM_FP32MM && !atkgpu_pfb_mm_k(cfg->nchan, cfg->oversample, (unsigned)kos)) return NULL;
    if (mx && !atkgpu_pfb_mx_supported(cfg->nchan, cfg->oversample, (unsigned)kos, cfg->fmt, cfg->form)) return NULL;
    if (atkgpu_init() != 0) return NULL;
    atkgpu_pfbx *v = (atkgpu_pfbx *)calloc(1, sizeof *v);
    if (!v) return NULL;
    v->cfg = *cfg;
    v->S = S;
    v->M = cfg->nchan;
    v->os = cfg->oversample;
    v->K = cfg->taps_per_poly;
    v->hop = v->M / v->os;
    v->ov = v->M * v->K - v->hop;
    v->kos = kos;
    v->bps = bps;
    v->mx = mx;
    v->rawrows = rawrows;
    v->keep = (size_t)v->ov + v->hop;                  /* M*K: a multiple of 64 */
    v->spitch = v->keep + cfg->block_max;
    v->m_cap = round16((cfg->block_max + v->hop) / v->hop + 1);
    const size_t ntap = (size_t)kos * v->hop;
    int ok = !atkgpu_stream_create(&v->stream);
    for (int i = 0; ok && i < 5; ++i) ok = !atkgpu_event_create(&v->ev[i]);
    if (ok && cfg->coherent) {
        v->d_R = (atkdsp_cf32 *)atkgpu_dmalloc((size_t)v->M * S * S * sizeof(atkdsp_cf32));
        v->d_cnt = (uint32_t *)atkgpu_dmalloc((size_t)v->M * sizeof(uint32_t));
        v->d_pwr = (float *)atkgpu_dmalloc((size_t)S * v->M * sizeof(float));
        v->d_mean = (float *)atkgpu_dmalloc((size_t)v->M * sizeof(float));
        ok = v->d_R && v->d_cnt && v->d_pwr && v->d_mean;
    }
    const size_t row_bytes = (size_t)S * v->spitch * (rawrows ? (size_t)bps : sizeof(atkdsp_cf32));
    if (ok) {
        v->d_rows[0] = (unsigned char *)atkgpu_dmalloc(row_bytes);
        v->d_rows[1] = (unsigned char *)atkgpu_dmalloc(row_bytes);
        if (!rawrows) v->d_raw = (unsigned char *)atkgpu_dmalloc((size_t)S * cfg->block_max * bps);
        v->d_taps = (float *)atkgpu_dmalloc(ntap * sizeof(float));
        v->d_out = (atkdsp_cf32 *)atkgpu_dmalloc((size_t)S * v->M * v->m_cap * sizeof(atkdsp_cf32));
        v->plan = atkgpu_fftplan_create(v->M);
        ok = v->d_rows[0] && v->d_rows[1] && (rawrows || v->d_raw) && v->d_taps && v->d_out && v->plan;
    }
    const size_t dbytes = (size_t)S * (sizeof(atkgpu_unpack_job) + sizeof(atkgpu_copy_job)) + 256;
    for (int i = 0; ok && i < 2; ++i) ok = (v->desc_h[i] = atkgpu_hmalloc_mapped(dbytes, &v->desc_d[i])) != NULL;
    /* taps: the FP64 prototype rounded once to FP32, zero-padded to kos rows */
    double *h = ok ? (double *)malloc((size_t)v->M * v->K * sizeof(double)) : NULL;
    float *hf = ok ? (float *)calloc(ntap, sizeof(float)) : NULL;
    ok = ok && h && hf && atkgpu_pfb_prototype(h, v->M, v->K, cfg->atten_db) == 0;
    if (ok) {
        for (size_t q = 0; q < (size_t)v->M * v->K; ++q) hf[q] = (float)h[q];
        ok = !atkgpu_h2d(v->stream, v->d_taps, hf, ntap * sizeof(float));
    }
    if (ok && mx) {
        const int f16 = cfg->form == ATKGPU_FORM_FP16X2 || cfg->form == ATKGPU_FORM_FP16;
        /* FP16 taps are scaled by a power of two so the largest is near 2^14:
         * the low term of the split then stays clear of FP16's subnormals */
        float amax = 0.f;
        for (size_t q = 0; q < ntap; ++q) amax = fmaxf(amax, fabsf(hf[q]));
        float tscale = 1.0f;
        if (f16 && amax > 0.f) {
            int e2 = (int)floor(log2(16384.0 / amax));
            if (e2 > 24) e2 = 24;
            if (e2 < -24) e2 = -24;
            tscale = ldexpf(1.0f, e2);
        }
        /* sample units: cu8 v = 2b - 255 (x = v * kk/2 - off), ci8 v = b, ci16q11 v = b */
        const float kk = cfg->fmt == ATKDSP_FMT_CU8 ? 1.0f / 127.5f : cfg->fmt == ATKDSP_FMT_CI8 ? 1.0f / 128.0f
                                                                                                : 1.0f / 2048.0f;
        v->scale = (cfg->fmt == ATKDSP_FMT_CU8 ? kk * 0.5f : kk) / tscale;
        float *ts = (float *)calloc((size_t)v->os * v->hop, sizeof(float));
        float *offs = (float *)malloc((size_t)S * 2 * sizeof(float));
        ok = ts && offs;
        if (ok) {
            for (unsigned r = 0; r < v->os; ++r)
                for (unsigned u = 0; u < v->hop; ++u) {
                    double acc = 0.0;
                    for (int t = (int)r; t < kos; t += (int)v->os) acc += hf[(size_t)t * v->hop + u];
                    ts[(size_t)r * v->hop + u] = (float)acc;
                }
            for (unsigned s = 0; s < S; ++s) { offs[2 * s] = cfg->off_re; offs[2 * s + 1] = cfg->off_im; }
            const size_t na = atkgpu_pfb_mx_tpad_count(v->M, v->os, (unsigned)kos, cfg->form);
            v->d_atab = atkgpu_dmalloc(na * 2);
            v->d_tapsum = (float *)atkgpu_dmalloc((size_t)v->os * v->hop * sizeof(float));
            v->d_offs = atkgpu_dmalloc((size_t)S * 2 * sizeof(float));
            ok = v->d_atab && v->d_tapsum && v->d_offs;
            if (ok) ok = !atkgpu_h2d(v->stream, v->d_tapsum, ts, (size_t)v->os * v->hop * sizeof(float));
            if (ok) ok = !atkgpu_h2d(v->stream, v->d_offs, offs, (size_t)S * 2 * sizeof(float));
            if (ok) ok = !atkgpu_pfb_mx_tpad(v->stream, v->d_taps, v->M, v->os, (unsigned)kos, cfg->form, tscale,
                                             v->d_atab);
            if (ok) ok = !atkgpu_sync(v->stream);
        }
        free(ts);
        free(offs);
    }
    free(h);
    free(hf);
    if (!ok || atkgpu_pfbx_reset(v) != ATKDSP_OK) {
        atkgpu_pfbx_destroy(v);
        return NULL;
    }
    return v;
}

void *atkgpu_pfbx_input(atkgpu_pfbx *v, unsigned s) {
    if (!v || s >= v->S) return NULL;
    if (!v->rawrows) return v->d_raw + (size_t)s * v->cfg.block_max * v->bps;
    return v->d_rows[v->cur] + ((size_t)s * v->spitch + v->keep) * v->bps;
}

int atkgpu_pfbx_upload(atkgpu_pfbx *v, unsigned s, const void *raw, size_t nbytes) {
    if (!v || s >= v->S || !raw || nbytes > v->cfg.block_max * (size_t)v->bps) return ATKDSP_E_ARG;
    return atk_put(v->stream, atkgpu_pfbx_input(v, s), raw, nbytes) ? ATKDSP_E_FFT : ATKDSP_OK;
}

/* the descriptor set for (cur, n): unpack jobs (FP32 form) then copy jobs */
static int build_desc(atkgpu_pfbx *v, int cur, size_t n) {
    if (v->desc_n[cur] == n) return 0;
    if (v->desc_n[cur] && atkgpu_sync(v->stream)) return ATKDSP_E_FFT;    /* the set may still be in use */
    atkgpu_unpack_job *uj = (atkgpu_unpack_job *)v->desc_h[cur];
    atkgpu_copy_job *cj = (atkgpu_copy_job *)(uj + v->S);
    const size_t es = v->rawrows ? (size_t)v->bps : sizeof(atkdsp_cf32);
    unsigned char *R = v->d_rows[cur], *T = v->d_rows[cur ^ 1];
    int ublk = 0, cblk = 0;
    for (unsigned s = 0; s < v->S; ++s) {
        if (!v->rawrows) {
            atkgpu_unpack_job *j = &uj[s];
            j->raw = v->d_raw + (size_t)s * v->cfg.block_max * v->bps;
            j->dst = R + ((size_t)s * v->spitch + v->keep) * es;
            j->off_re = v->cfg.off_re;
            j->off_im = v->cfg.off_im;
            j->n = (int)n;
            j->blk0 = ublk;
            j->nblk = (int)((n + ATKGPU_RAGGED_SPB - 1) / ATKGPU_RAGGED_SPB);
            j->fmt = v->cfg.fmt;
            ublk += j->nblk;
        }
        atkgpu_copy_job *c = &cj[s];
        c->src = R + ((size_t)s * v->spitch + n) * es;
        c->dst = T + (size_t)s * v->spitch * es;
        c->blk0 = cblk;
        if (v->rawrows) {
            c->n = (int)(v->keep * es);                   /* bytes */
            c->nblk = (c->n + 4095) / 4096;
        } else {
            c->n = (int)v->keep;                          /* cf32 samples */
            c->nblk = (c->n + ATKGPU_RAGGED_SPB - 1) / ATKGPU_RAGGED_SPB;
        }
        cblk += c->nblk;
    }
    v->desc_n[cur] = n;
    return 0;
}

ptrdiff_t atkgpu_pfbx_run(atkgpu_pfbx *v, size_t n) {
    if (!v || n > v->cfg.block_max) return ATKDSP_E_ARG;
    const size_t avail = (size_t)v->lc + n;
    const size_t m = avail / v->hop;
    int rc = n ? build_desc(v, v->cur, n) : 0;
    if (rc) return rc;
    const int cur = v->cur;
    atkgpu_unpack_job *uj = (atkgpu_unpack_job *)v->desc_h[cur];
    atkgpu_copy_job *cj = (atkgpu_copy_job *)(uj + v->S);
    char *dbase = (char *)v->desc_d[cur], *hbase = (char *)v->desc_h[cur];
#define DEV(p) ((void *)(dbase + ((char *)(p) - hbase)))
    const size_t es = v->rawrows ? (size_t)v->bps : sizeof(atkdsp_cf32);
    const size_t mp = round16(m);
    unsigned char *R = v->d_rows[cur];
    rc = atkgpu_event_record(v->ev[0], v->stream);
    if (!rc && n && !v->rawrows) {
        int ublk = 0;
        for (unsigned s = 0; s < v->S; ++s) ublk += uj[s].nblk;
        rc = atkgpu_unpack_jobs(v->stream, DEV(uj), (int)v->S, ublk, 0, 0);
    }
    if (!rc) rc = atkgpu_event_record(v->ev[1], v->stream);
    if (!rc && m) {
        const void *base = R + (v->keep - v->ov - v->lc) * es;
        const void *tw = atkgpu_fftplan_twiddles(v->plan);
        /* hop b's window starts at stream sample b*hop - ov (block_origin) */
        const long long zl = (long long)v->ov - v->b * (long long)v->hop;
        const int zlo = zl > 0 ? (int)zl : 0;
        if (v->mx)
            rc = atkgpu_pfb_mx(v->stream, base, v->spitch, (size_t)v->ov + avail, v->S, m, v->b, v->M, v->os,
                               (unsigned)v->kos, v->cfg.fmt, v->cfg.form, v->d_atab, v->d_tapsum, v->d_taps, zlo,
                               v->scale, v->d_offs, tw, NULL, 0, v->d_out, mp);
        else if (v->cfg.form == ATKGPU_FORM_FP32RAW)
            rc = atkgpu_pfb_fused_raw(v->stream, base, v->spitch, (size_t)v->ov + avail, v->S, m, v->b, v->M, v->os,
                                      (unsigned)v->kos, v->cfg.fmt, v->cfg.off_re, v->cfg.off_im, zlo, v->d_taps,
                                      tw, v->d_out, mp);
        else if (v->cfg.form == ATKGPU_FORM_FP32MM)
            rc = atkgpu_pfb_mm(v->stream, base, v->spitch, (size_t)v->ov + avail, v->S, m, v->b, v->M, v->os,
                               (unsigned)v->kos, v->d_taps, tw, NULL, 0, v->d_out, mp, 0);
        else
            rc = atkgpu_pfb_fused(v->stream, base, v->spitch, (size_t)v->ov + avail, v->S, m, v->b, v->M, v->os,
                                  (unsigned)v->kos, v->d_taps, tw, NULL, 0, v->d_out, mp, 0);
    }
    if (!rc) rc = atkgpu_event_record(v->ev[2], v->stream);
    if (!rc && m && v->cfg.coherent) {
        rc = atkgpu_pfb_cov(v->stream, v->d_out, v->S, v->M, m, mp, NULL, v->d_R, v->d_cnt);
        if (!rc) rc = atkgpu_pfb_power(v->stream, v->d_out, v->S, v->M, m, mp, v->d_pwr, v->d_mean);
    }
    if (!rc) rc = atkgpu_event_record(v->ev[3], v->stream);
    if (!rc && n) {
        int cblk = 0;
        for (unsigned s = 0; s < v->S; ++s) cblk += cj[s].nblk;
        rc = v->rawrows ? atkgpu_copy_bytes_jobs(v->stream, DEV(cj), (int)v->S, cblk)
                   : atkgpu_copy_jobs(v->stream, DEV(cj), (int)v->S, cblk);
    }
    if (!rc) rc = atkgpu_event_record(v->ev[4], v->stream);
#undef DEV
    if (rc) return ATKDSP_E_FFT;
    v->lc = (unsigned)(avail - m * v->hop);
    v->b += (long long)m;
    if (n) v->cur ^= 1;
    v->last_m = m;
    v->ran = 1;
    return (ptrdiff_t)m;
}

int atkgpu_pfbx_sync(atkgpu_pfbx *v) {
    if (!v) return ATKDSP_E_ARG;
    return atkgpu_sync(v->stream) ? ATKDSP_E_FFT : ATKDSP_OK;
}

int atkgpu_pfbx_timing(atkgpu_pfbx *v, float *ms_run, float *ms_unpack, float *ms_pfb, float *ms_w) {
    if (!v || !v->ran) return ATKDSP_E_ARG;
    float a = 0, b = 0, c = 0, d = 0;
    if (atkgpu_event_elapsed_ms(v->ev[0], v->ev[4], &a) || atkgpu_event_elapsed_ms(v->ev[0], v->ev[1], &b) ||
        atkgpu_event_elapsed_ms(v->ev[1], v->ev[2], &c) || atkgpu_event_elapsed_ms(v->ev[2], v->ev[3], &d))
        return ATKDSP_E_FFT;
    if (ms_run) *ms_run = a;
    if (ms_unpack) *ms_unpack = b;
    if (ms_pfb) *ms_pfb = c;
    if (ms_w) *ms_w = d;
    return ATKDSP_OK;
}

int atkgpu_pfbx_read_cov(atkgpu_pfbx *v, atkdsp_cf32 *R, uint32_t *counts) {
    if (!v || !v->cfg.coherent) return ATKDSP_E_ARG;
    int rc = 0;
    if (R) rc = atkgpu_d2h(v->stream, R, v->d_R, (size_t)v->M * v->S * v->S * sizeof(atkdsp_cf32));
    if (!rc && counts) rc = atkgpu_d2h(v->stream, counts, v->d_cnt, (size_t)v->M * sizeof(uint32_t));
    if (!rc) rc = atkgpu_sync(v->stream);
    return rc ? ATKDSP_E_FFT : ATKDSP_OK;
}

int atkgpu_pfbx_read_power(atkgpu_pfbx *v, float *pwr, float *mean_pwr) {
    if (!v || !v->cfg.coherent || !v->last_m) return ATKDSP_E_ARG;
    int rc = 0;
    if (pwr) rc = atkgpu_d2h(v->stream, pwr, v->d_pwr, (size_t)v->S * v->M * sizeof(float));
    if (!rc && mean_pwr) rc = atkgpu_d2h(v->stream, mean_pwr, v->d_mean, (size_t)v->M * sizeof(float));
    if (!rc) rc = atkgpu_sync(v->stream);
    return rc ? ATKDSP_E_FFT : ATKDSP_OK;
}

ptrdiff_t atkgpu_pfbx_read(atkgpu_pfbx *v, unsigned s, atkdsp_cf32 *out, size_t cap) {
    if (!v || s >= v->S || !out) return ATKDSP_E_ARG;
    const size_t m = v->last_m, mp = round16(m);
    if (cap < (size_t)v->M * m) return ATKDSP_E_CAP;
    int rc = m ? atkgpu_d2h_2d(v->stream, out, m * sizeof(atkdsp_cf32), v->d_out + (size_t)s * v->M * mp,
                               mp * sizeof(atkdsp_cf32), m * sizeof(atkdsp_cf32), v->M)
               : 0;
    if (!rc) rc = atkgpu_sync(v->stream);
    return rc ? ATKDSP_E_FFT : (ptrdiff_t)m;
}
