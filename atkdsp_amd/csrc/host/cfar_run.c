/* CFAR over spectrum frames and the burst tracker (reference
 * include/atkdsp.h:1733-1925, src/cfar.c:270-819): the atkdsp_cfar handle.
 *
 * Per call, frames go to the device in slabs; the device turns dB into
 * power, averages the last k_frames frames and runs the CA / OS / GO / SO test
 * of every bin against its reference cells (csrc/dev/cfar.hip), and returns
 * each frame's average, threshold, noise estimate, hits and power. The
 * tracker then runs on the host, frame by frame, as the header defines it:
 *   blobs        hit bins closer than merge_gap non-hit bins, min_bins wide;
 *   association  each blob to the track it overlaps most within merge_gap+1
 *                bins (ties to the older), a track may take several blobs; a
 *                blob near no track starts a pending one (beyond max_tracks
 *                it is dropped and counted);
 *   life         OPEN after open_m hit frames in the last open_n, forgotten
 *                when pending with none; CLOSE after close_k misses in a
 *                row; UPDATE when an open track's span grows; one event per
 *                track per frame, tracks visited in id order;
 *   times        with k_frames > 1 the first and last hit frames are refined
 *                by a single-frame band test at ATKDSP_CFAR_EDGE_PFA
 *                (T0_COARSE / T1_COARSE when it fails).
 * The tables of reference cells (left nearest first, right nearest first,
 * then a side's shortfall borrowed from beyond the other side) are built on
 * the host when the mask changes. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "internal.h"

_Static_assert(sizeof(atkdsp_cfar_event) == 64, "event record: 64 bytes, as the reference's");
_Static_assert(sizeof(atkdsp_cfar_counters) == 88, "counters: 88 bytes, as the reference's");

#define SLAB_VALUES ((size_t)1 << 22)   /* frame values per device slab */

typedef struct {
    uint32_t id;
    int open;                 /* 0 pending, 1 open */
    int fresh;                /* no frame processed since it started */
    uint64_t window;          /* hit frames, newest in bit 0 */
    int misses;
    int32_t n_hits, cur_lo, cur_hi, lo, hi;   /* this frame's span, the union of spans */
    double sum_be, sum_e;     /* excess-weighted bin sum and excess over the track's life */
    double peak_x, peak_n;    /* the strongest averaged bin and its noise */
    double band;              /* best frame's hit-bin power over noise */
    uint32_t flags;
    double t0;
    int64_t last_hit, last_raw;
    double t_hit_end, dt_hit, t_raw_end;
    /* this frame's observation */
    int seen;
    int32_t obs_lo, obs_hi;
    double obs_x, obs_n;
} cfar_track;

typedef struct {
    int32_t lo, hi;
    double sum_be, sum_e, sum_x, sum_n, peak_x, peak_n;
    int track;
} cfar_blob;

struct atkdsp_cfar {
    atkdsp_cfar_config cfg;
    size_t nbins;
    int N;
    double n_eff, K, alpha, alpha_ca, et;
    uint8_t *mask, *ok, *sided, *reg;   /* reg: the table is the plain window */
    int32_t *cells;
    /* device */
    void *stream;
    int32_t *d_cells;
    uint8_t *d_ok, *d_sided, *d_reg, *d_hits;
    float *d_db;
    double *d_raw, *d_hist, *d_X, *d_thr, *d_noise;
    size_t slab;
    /* host copies of a slab's results (pinned): the average, noise and hits
     * of every frame, the power of every frame when k_frames > 1, the
     * threshold of the slab's last frame */
    double *X, *thr, *noise, *raw;
    uint8_t *hits;
    /* with one frame averaged the tracker reads X and noise at hit bins only:
     * their indices go up and the values come back gathered, up to gcap */
    unsigned *idx, *d_idx;
    double *gath, *d_gath;
    size_t gcap;
    size_t n_ok, n_untestable;   /* bins tested / unmasked but lacking cells, per frame */
    /* the last k frames' power and times, ring by frame number */
    double *ring, *ring_t, *ring_dt;
    int64_t seen, frame;
    /* the last tested frame, for atkdsp_cfar_last */
    double *last_X, *last_thr;
    uint8_t *last_hits;
    int has_last;
    double *beta;             /* per band width, NaN until needed */
    cfar_track *tracks;
    int ntracks;
    uint32_t next_id;
    cfar_blob *blobs;
    atkdsp_cfar_counters st;
};


static int config_ok(size_t nbins, const atkdsp_cfar_config *c) {
    const long N = 2L * c->n_ref;
    if (c->mode < ATKDSP_CFAR_CA || c->mode > ATKDSP_CFAR_SO) return 0;
    if (c->n_ref < 1 || c->n_ref > 256 || c->n_guard < 0 || c->n_guard > 256) return 0;
    if (nbins < (size_t)(N + 2L * c->n_guard + 1) || nbins > ((size_t)1 << 24) ||
        (double)nbins * (double)N > (double)((size_t)1 << 26))
        return 0;
    if (c->k_frames < 1 || c->k_frames > 64 || !(c->looks >= 0.25 && c->looks <= 1e4)) return 0;
    if (!(c->pfa >= 1e-15 && c->pfa <= 0.5)) return 0;
    if (c->mode == ATKDSP_CFAR_OS && (c->os_k < 0 || c->os_k > N)) return 0;
    if (c->open_n < 1 || c->open_n > 64 || c->open_m < 1 || c->open_m > c->open_n) return 0;
    if (c->close_k < 1 || c->close_k > 1000 || c->merge_gap < 0 || (size_t)c->merge_gap > nbins) return 0;
    if (c->min_bins < 1 || (size_t)c->min_bins > nbins || c->max_tracks < 1 || c->max_tracks > 65536) return 0;
    for (int m = 0; m < ATKDSP_CFAR_RHO_MAX; ++m)
        if (!(c->rho2[m] >= 0.0 && c->rho2[m] < 1.0)) return 0;
    return 1;
}

/* each bin's reference cells; ok = it has N of them, sided = none borrowed */
static void build_tables(atkdsp_cfar *c) {
    const long nb = (long)c->nbins, nr = c->cfg.n_ref, ng = c->cfg.n_guard;
    for (long i = 0; i < nb; ++i) {
        c->ok[i] = 0;
        c->sided[i] = 0;
        c->reg[i] = 0;
        if (c->mask[i]) continue;
        int32_t *out = c->cells + (size_t)i * (size_t)c->N;
        int32_t left[256], right[256], borrowed[512];
        long nl = 0, nrt = 0, nbw = 0;
        long j = i - ng - 1;
        for (; j >= 0 && nl < nr; --j)
            if (!c->mask[j]) left[nl++] = (int32_t)j;
        const long left_end = j;
        j = i + ng + 1;
        for (; j < nb && nrt < nr; ++j)
            if (!c->mask[j]) right[nrt++] = (int32_t)j;
        const long right_end = j;
        /* a short left side borrows beyond the right side, then a short right
         * side beyond the left */
        long need = nr - nl;
        for (j = right_end; need > 0 && j < nb; ++j)
            if (!c->mask[j]) { borrowed[nbw++] = (int32_t)j; --need; }
        need = nr - nrt;
        for (j = left_end; need > 0 && j >= 0; --j)
            if (!c->mask[j]) { borrowed[nbw++] = (int32_t)j; --need; }
        if (nl + nrt + nbw < c->N) continue;
        long o = 0;
        for (long q = 0; q < nl; ++q) out[o++] = left[q];
        for (long q = 0; q < nrt; ++q) out[o++] = right[q];
        for (long q = 0; q < nbw; ++q) out[o++] = borrowed[q];
        c->ok[i] = 1;
        c->sided[i] = nbw == 0;
        int plain = nbw == 0;
        for (long q = 0; q < nr && plain; ++q)
            plain = out[q] == (int32_t)(i - ng - 1 - q) && out[nr + q] == (int32_t)(i + ng + 1 + q);
        c->reg[i] = (uint8_t)plain;
    }
    c->n_ok = c->n_untestable = 0;
    for (long i = 0; i < nb; ++i) {
        c->n_ok += c->ok[i];
        c->n_untestable += !c->ok[i] && !c->mask[i];
    }
}

static int upload_tables(atkdsp_cfar *c) {
    return atkgpu_h2d(c->stream, c->d_cells, c->cells, c->nbins * (size_t)c->N * sizeof(int32_t)) ||
           atkgpu_h2d(c->stream, c->d_ok, c->ok, c->nbins) || atkgpu_h2d(c->stream, c->d_sided, c->sided, c->nbins) ||
           atkgpu_h2d(c->stream, c->d_reg, c->reg, c->nbins) ||
           atkgpu_sync(c->stream);
}

void atkdsp_cfar_destroy(atkdsp_cfar *c) {
    if (!c) return;
    if (c->stream) atkgpu_sync(c->stream);
    atkgpu_dfree(c->d_cells); atkgpu_dfree(c->d_ok); atkgpu_dfree(c->d_sided); atkgpu_dfree(c->d_reg);
    atkgpu_dfree(c->d_hits);
    atkgpu_dfree(c->d_db); atkgpu_dfree(c->d_raw); atkgpu_dfree(c->d_hist); atkgpu_dfree(c->d_X);
    atkgpu_dfree(c->d_thr); atkgpu_dfree(c->d_noise); atkgpu_dfree(c->d_idx); atkgpu_dfree(c->d_gath);
    atkgpu_hfree(c->idx); atkgpu_hfree(c->gath);
    if (c->stream) atkgpu_stream_destroy(c->stream);
    free(c->mask); free(c->ok); free(c->sided); free(c->reg); free(c->cells);
    atkgpu_hfree(c->X); atkgpu_hfree(c->thr); atkgpu_hfree(c->noise); atkgpu_hfree(c->raw); atkgpu_hfree(c->hits);
    free(c->ring); free(c->ring_t); free(c->ring_dt);
    free(c->last_X); free(c->last_thr); free(c->last_hits);
    free(c->beta); free(c->tracks); free(c->blobs);
    free(c);
}

atkdsp_cfar *atkdsp_cfar_create(size_t nbins, const atkdsp_cfar_config *cfg) {
    if (!cfg || !config_ok(nbins, cfg)) return NULL;
    if (atkgpu_init() != 0) return NULL;
    atkdsp_cfar *c = (atkdsp_cfar *)calloc(1, sizeof *c);
    if (!c) return NULL;
    c->cfg = *cfg;
    c->nbins = nbins;
    c->N = 2 * cfg->n_ref;
    if (c->cfg.mode == ATKDSP_CFAR_OS && c->cfg.os_k == 0) c->cfg.os_k = (3 * c->N + 3) / 4;
    c->K = c->cfg.looks * (double)c->cfg.k_frames;
    c->n_eff = atkdsp_cfar_n_eff(c->cfg.n_ref, c->cfg.rho2);
    if (atkdsp_cfar_alpha(c->cfg.mode, c->cfg.n_ref, c->cfg.pfa, c->K, c->cfg.os_k, c->cfg.rho2, &c->alpha,
                          &c->et) != ATKDSP_OK) {
        free(c);
        return NULL;
    }
    if (c->cfg.mode == ATKDSP_CFAR_GO || c->cfg.mode == ATKDSP_CFAR_SO) {
        if (atkdsp_cfar_alpha(ATKDSP_CFAR_CA, c->cfg.n_ref, c->cfg.pfa, c->K, 0, c->cfg.rho2, &c->alpha_ca, NULL) !=
            ATKDSP_OK) {
            free(c);
            return NULL;
        }
    } else {
        c->alpha_ca = c->alpha;
    }
    const size_t k = (size_t)c->cfg.k_frames, nb = nbins;
    size_t slab = SLAB_VALUES / nb;
    if (slab < 1) slab = 1;
    if (slab > 256) slab = 256;
    c->slab = slab;
    c->mask = (uint8_t *)calloc(nb, 1);
    c->ok = (uint8_t *)calloc(nb, 1);
    c->sided = (uint8_t *)calloc(nb, 1);
    c->reg = (uint8_t *)calloc(nb, 1);
    c->cells = (int32_t *)calloc(nb * (size_t)c->N, sizeof(int32_t));
    c->X = (double *)atkgpu_hmalloc(slab * nb * sizeof(double));
    c->thr = (double *)atkgpu_hmalloc(nb * sizeof(double));
    c->noise = (double *)atkgpu_hmalloc(slab * nb * sizeof(double));
    c->raw = k > 1 ? (double *)atkgpu_hmalloc(slab * nb * sizeof(double)) : NULL;
    c->hits = (uint8_t *)atkgpu_hmalloc(slab * nb + 8);
    c->gcap = k == 1 ? (slab * nb + 3) / 4 : 0;
    if (c->gcap) {
        c->idx = (unsigned *)atkgpu_hmalloc(c->gcap * sizeof(unsigned));
        c->gath = (double *)atkgpu_hmalloc(2 * c->gcap * sizeof(double));
    }
    c->ring = (double *)calloc(k * nb, sizeof(double));
    c->ring_t = (double *)calloc(k, sizeof(double));
    c->ring_dt = (double *)calloc(k, sizeof(double));
    c->last_X = (double *)malloc(nb * sizeof(double));
    c->last_thr = (double *)malloc(nb * sizeof(double));
    c->last_hits = (uint8_t *)malloc(nb);
    c->beta = (double *)malloc((nb + 1) * sizeof(double));
    c->tracks = (cfar_track *)calloc((size_t)c->cfg.max_tracks, sizeof(cfar_track));
    c->blobs = (cfar_blob *)calloc(nb, sizeof(cfar_blob));
    int ok = c->mask && c->ok && c->sided && c->reg && c->cells && c->X && c->thr && c->noise && (c->raw || k == 1) && c->hits &&
             c->ring && c->ring_t && c->ring_dt && c->last_X && c->last_thr && c->last_hits && c->beta &&
             c->tracks && c->blobs && !atkgpu_stream_create(&c->stream);
    if (ok) {
        c->d_cells = (int32_t *)atkgpu_dmalloc(nb * (size_t)c->N * sizeof(int32_t));
        c->d_ok = (uint8_t *)atkgpu_dmalloc(nb);
        c->d_sided = (uint8_t *)atkgpu_dmalloc(nb);
        c->d_reg = (uint8_t *)atkgpu_dmalloc(nb);
        c->d_hits = (uint8_t *)atkgpu_dmalloc(slab * nb);
        c->d_db = (float *)atkgpu_dmalloc(slab * nb * sizeof(float));
        c->d_raw = (double *)atkgpu_dmalloc((k - 1 + slab) * nb * sizeof(double));
        c->d_hist = (double *)atkgpu_dmalloc((k > 1 ? k - 1 : 1) * nb * sizeof(double));
        c->d_X = k > 1 ? (double *)atkgpu_dmalloc(slab * nb * sizeof(double)) : NULL;
        c->d_thr = (double *)atkgpu_dmalloc(slab * nb * sizeof(double));
        c->d_noise = (double *)atkgpu_dmalloc(slab * nb * sizeof(double));
        if (c->gcap) {
            c->d_idx = (unsigned *)atkgpu_dmalloc(c->gcap * sizeof(unsigned));
            c->d_gath = (double *)atkgpu_dmalloc(2 * c->gcap * sizeof(double));
        }
        ok = c->d_cells && c->d_ok && c->d_sided && c->d_reg && c->d_hits && c->d_db && c->d_raw && c->d_hist && (c->d_X || k == 1) &&
             c->d_thr && c->d_noise && (!c->gcap || (c->idx && c->gath && c->d_idx && c->d_gath));
    }
    if (ok && k > 1) ok = !atkgpu_fill_u32(c->stream, c->d_hist, 0, (k - 1) * nb * 2) && !atkgpu_sync(c->stream);
    if (!ok) { atkdsp_cfar_destroy(c); return NULL; }
    for (size_t i = 0; i <= nb; ++i) c->beta[i] = NAN;
    c->next_id = 1;
    build_tables(c);
    if (upload_tables(c)) { atkdsp_cfar_destroy(c); return NULL; }
    return c;
}

void atkdsp_cfar_reset(atkdsp_cfar *c) {
    if (!c) return;
    c->ntracks = 0;
    c->seen = 0;
    c->frame = 0;
    memset(c->ring, 0, (size_t)c->cfg.k_frames * c->nbins * sizeof(double));
    if (c->cfg.k_frames > 1) {
        atkgpu_fill_u32(c->stream, c->d_hist, 0, (size_t)(c->cfg.k_frames - 1) * c->nbins * 2);
        atkgpu_sync(c->stream);
    }
    c->has_last = 0;
    c->st.resets++;
}

int atkdsp_cfar_set_mask(atkdsp_cfar *c, const uint8_t *mask) {
    if (!c) return ATKDSP_E_ARG;
    for (size_t i = 0; i < c->nbins; ++i) c->mask[i] = mask ? (mask[i] != 0) : 0;
    build_tables(c);
    return upload_tables(c) ? ATKDSP_E_FFT : ATKDSP_OK;
}

size_t atkdsp_cfar_events_max(const atkdsp_cfar *c, size_t nframes) {
    return c ? nframes * (size_t)c->cfg.max_tracks : 0;
}

/* the single-frame band test's factor: P(Gamma(s)/s > beta) = EDGE_PFA for the
 * equivalent independent bins among nb contiguous ones */
static double edge_beta(atkdsp_cfar *c, int nb) {
    if (!isnan(c->beta[nb])) return c->beta[nb];
    double pairs = 0.0;
    for (int m = 1; m <= ATKDSP_CFAR_RHO_MAX && m < nb; ++m) pairs += (double)(nb - m) * c->cfg.rho2[m - 1];
    const double s = (double)nb / (1.0 + 2.0 * pairs / (double)nb) * c->cfg.looks;
    const double target = log(ATKDSP_CFAR_EDGE_PFA);
    double a = log(1e-6), b = log(1e6);
    for (int i = 0; i < 100; ++i) {
        const double mid = 0.5 * (a + b);
        double lp, lq;
        atk_cfar_lgam_pq(s, s * exp(mid), &lp, &lq);
        if (lq > target) a = mid; else b = mid;
    }
    c->beta[nb] = exp(b);
    return c->beta[nb];
}

/* does the band [lo, hi] of one frame's power pass the single-frame test
 * against this frame's noise estimate */
static int band_pass(atkdsp_cfar *c, const double *pw, const double *noise, int32_t lo, int32_t hi) {
    double sp = 0.0, sn = 0.0;
    int nb = 0;
    for (int32_t b = lo; b <= hi; ++b)
        if (c->ok[b]) { sp += pw[b]; sn += noise[b]; ++nb; }
    return nb > 0 && sp > edge_beta(c, nb) * sn;
}

static uint32_t track_flags(const atkdsp_cfar *c, const cfar_track *t) {
    uint32_t f = t->flags;
    for (int32_t b = t->lo; b <= t->hi; ++b)
        if (!c->sided[b] || !c->ok[b]) { f |= ATKDSP_CFAR_EDGE; break; }
    const long m = c->cfg.merge_gap + 1;
    long a = (long)t->lo - m, z = (long)t->hi + m + 1;
    if (a < 0) a = 0;
    if (z > (long)c->nbins) z = (long)c->nbins;
    for (long b = a; b < z; ++b)
        if (c->mask[b]) { f |= ATKDSP_CFAR_MASK_ADJ; break; }
    return f;
}

static void emit(const atkdsp_cfar *c, const cfar_track *t, int type, double t1, uint32_t extra,
                 atkdsp_cfar_event *e) {
    memset(e, 0, sizeof *e);
    e->type = type;
    e->id = t->id;
    e->flags = track_flags(c, t) | extra;
    e->lo_bin = t->lo;
    e->hi_bin = t->hi;
    e->n_hits = t->n_hits;
    e->t0 = t->t0;
    e->t1 = t1;
    e->fc_bin = t->sum_e > 0.0 ? t->sum_be / t->sum_e : 0.5 * ((double)t->lo + (double)t->hi);
    e->level_db = t->peak_x > 0.0 ? (float)(10.0 * log10(t->peak_x)) : -INFINITY;
    e->noise_db = t->peak_n > 0.0 ? (float)(10.0 * log10(t->peak_n)) : -INFINITY;
    e->snr_db = t->peak_n > 0.0 ? e->level_db - e->noise_db : INFINITY;
    e->band_snr_db = t->band > 0.0 ? (float)(10.0 * log10(t->band)) : -INFINITY;
}

static void emit_close(atkdsp_cfar *c, cfar_track *t, uint32_t extra, atkdsp_cfar_event *e) {
    const int k = c->cfg.k_frames;
    double t1;
    if (k == 1) {
        t1 = t->t_hit_end;
    } else if (t->last_raw >= t->last_hit - (k - 1)) {
        t1 = t->t_raw_end;
    } else {
        t1 = t->t_hit_end - 0.5 * (double)(k - 1) * t->dt_hit;
        t->flags |= ATKDSP_CFAR_T1_COARSE;
    }
    c->st.closed++;
    emit(c, t, ATKDSP_CFAR_CLOSE, t1, extra, e);
}

/* one tested frame: X, noise, hits and pw (this frame's power; NULL with one
 * frame averaged, where no band test runs) are the device's results, X and
 * noise valid at least at the hit bins (the only ones read here); returns the
 * events written */
static size_t track_frame(atkdsp_cfar *c, const double *X, const double *noise, const uint8_t *hits,
                          const double *pw, double tf, double dt, atkdsp_cfar_event *ev) {
    const size_t nb = c->nbins;
    const int k = c->cfg.k_frames;
    const int64_t F = c->frame - 1;                          /* this frame's number */
    /* blobs */
    const long gap = c->cfg.merge_gap;
    size_t nblob = 0;
    for (long i = 0; i < (long)nb;) {
        if (!hits[i]) {
            ++i;
            while (i + 8 <= (long)nb) {                         /* whole zero words at a time */
                uint64_t w;
                memcpy(&w, hits + i, 8);
                if (w) break;
                i += 8;
            }
            continue;
        }
        long hi = i, prev = i;
        for (long j = i + 1; j < (long)nb && j - prev - 1 <= gap; ++j)
            if (hits[j]) { hi = j; prev = j; }
        const long lo = i;
        i = hi + 1;
        if (hi - lo + 1 < c->cfg.min_bins) continue;
        cfar_blob *o = &c->blobs[nblob++];
        memset(o, 0, sizeof *o);
        o->lo = (int32_t)lo;
        o->hi = (int32_t)hi;
        long pk = -1;
        for (long b = lo; b <= hi; ++b) {
            if (!hits[b]) continue;
            const double ex = X[b] - noise[b] > 0.0 ? X[b] - noise[b] : 0.0;
            o->sum_be += (double)b * ex;
            o->sum_e += ex;
            o->sum_x += X[b];
            o->sum_n += noise[b];
            if (pk < 0 || X[b] > X[pk]) pk = b;
        }
        o->peak_x = X[pk];
        o->peak_n = noise[pk];
        o->track = -1;
    }
    /* each blob to the existing track it overlaps most (ties: the older) */
    const int existing = c->ntracks;
    const long reach = gap + 1;
    for (size_t q = 0; q < nblob; ++q) {
        cfar_blob *o = &c->blobs[q];
        long best_ov = 0;
        for (int t = 0; t < existing; ++t) {
            const cfar_track *tr = &c->tracks[t];
            if ((long)o->lo > (long)tr->cur_hi + reach || (long)o->hi < (long)tr->cur_lo - reach) continue;
            const long ov = (long)(o->hi < tr->cur_hi ? o->hi : tr->cur_hi) -
                            (long)(o->lo > tr->cur_lo ? o->lo : tr->cur_lo) + 1;
            if (o->track < 0 || ov > best_ov) { o->track = t; best_ov = ov; }
        }
    }
    /* blobs near no track start pending ones */
    for (size_t q = 0; q < nblob; ++q) {
        cfar_blob *o = &c->blobs[q];
        if (o->track >= 0) continue;
        if (c->ntracks >= c->cfg.max_tracks) { c->st.dropped++; continue; }
        cfar_track *tr = &c->tracks[c->ntracks];
        memset(tr, 0, sizeof *tr);
        tr->id = c->next_id;
        c->next_id = c->next_id + 1 ? c->next_id + 1 : 1;
        tr->cur_lo = tr->lo = o->lo;
        tr->cur_hi = tr->hi = o->hi;
        tr->peak_x = -1.0;
        tr->last_hit = -1;
        tr->last_raw = -1;
        tr->dt_hit = dt;
        tr->fresh = 1;
        o->track = c->ntracks++;
    }
    /* each track's observation: its blobs in blob order */
    for (size_t q = 0; q < nblob; ++q) {
        const cfar_blob *o = &c->blobs[q];
        if (o->track < 0) continue;
        cfar_track *tr = &c->tracks[o->track];
        if (!tr->seen) {
            tr->seen = 1;
            tr->obs_lo = o->lo;
            tr->obs_hi = o->hi;
            tr->obs_x = 0.0;
            tr->obs_n = 0.0;
        } else {
            if (o->lo < tr->obs_lo) tr->obs_lo = o->lo;
            if (o->hi > tr->obs_hi) tr->obs_hi = o->hi;
        }
        tr->sum_be += o->sum_be;
        tr->sum_e += o->sum_e;
        tr->obs_x += o->sum_x;
        tr->obs_n += o->sum_n;
        if (o->peak_x > tr->peak_x) { tr->peak_x = o->peak_x; tr->peak_n = o->peak_n; }
    }
    /* every track, in id order: its window, its state, its event */
    const uint64_t wmask = c->cfg.open_n >= 64 ? ~(uint64_t)0 : (((uint64_t)1 << c->cfg.open_n) - 1);
    size_t nev = 0;
    int kept = 0;
    for (int t = 0; t < c->ntracks; ++t) {
        cfar_track *tr = &c->tracks[t];
        int grew = 0;
        if (tr->seen) {
            const int32_t lo = tr->obs_lo, hi = tr->obs_hi;
            tr->window = ((tr->window << 1) | 1u) & wmask;
            tr->misses = 0;
            tr->n_hits++;
            tr->cur_lo = lo;
            tr->cur_hi = hi;
            if (!tr->fresh) grew = lo < tr->lo || hi > tr->hi;
            if (lo < tr->lo) tr->lo = lo;
            if (hi > tr->hi) tr->hi = hi;
            if (tr->obs_n > 0.0 && tr->obs_x / tr->obs_n > tr->band) tr->band = tr->obs_x / tr->obs_n;
            tr->last_hit = F;
            tr->t_hit_end = tf + dt;
            tr->dt_hit = dt;
            if (tr->fresh) {
                tr->fresh = 0;
                if (k == 1) {
                    tr->t0 = tf;
                } else {
                    /* the earliest raw frame of the averaging window whose band passes */
                    int found = 0;
                    for (int r = 0; r < k && !found; ++r) {
                        const size_t s = (size_t)((F + 1 + r) % k);
                        if (band_pass(c, c->ring + s * nb, noise, lo, hi)) { tr->t0 = c->ring_t[s]; found = 1; }
                    }
                    if (!found) {
                        tr->t0 = tf - 0.5 * (double)(k - 1) * dt;
                        tr->flags |= ATKDSP_CFAR_T0_COARSE;
                    }
                }
            }
        } else {
            tr->window = (tr->window << 1) & wmask;
            tr->misses++;
        }
        tr->seen = 0;
        if (k > 1 && band_pass(c, pw, noise, tr->cur_lo, tr->cur_hi)) {
            tr->last_raw = F;
            tr->t_raw_end = tf + dt;
        }
        if (!tr->open) {
            if (__builtin_popcountll(tr->window) >= c->cfg.open_m) {
                tr->open = 1;
                c->st.opened++;
                emit(c, tr, ATKDSP_CFAR_OPEN, NAN, 0, &ev[nev++]);
            } else if (tr->window == 0) {
                continue;                                     /* forgotten */
            }
        } else if (tr->misses >= c->cfg.close_k) {
            emit_close(c, tr, 0, &ev[nev++]);
            continue;
        } else if (grew) {
            c->st.updates++;
            emit(c, tr, ATKDSP_CFAR_UPDATE, NAN, 0, &ev[nev++]);
        }
        if (kept != t) c->tracks[kept] = *tr;
        ++kept;
    }
    c->ntracks = kept;
    return nev;
}

/* nf frames already in d_db, the first of them frame f0 of the call: the
 * device test, the results back, the tracker; returns the events written or
 * ATKDSP_E_FFT */
static ptrdiff_t run_slab(atkdsp_cfar *c, size_t nf, size_t f0, double frame_t0, double frame_dt,
                          atkdsp_cfar_event *ev) {
    const size_t nb = c->nbins, k = (size_t)c->cfg.k_frames;
    /* the raw rows: k-1 frames of history, then this slab; with one frame
     * averaged, the raw rows are the averages */
    const double *dX = k > 1 ? c->d_X : c->d_raw;
    if ((k > 1 && atkgpu_d2d(c->stream, c->d_raw, c->d_hist, (k - 1) * nb * sizeof(double))) ||
        atkgpu_cfar_raw(c->stream, c->d_db, nf * nb, c->d_raw + (k - 1) * nb) ||
        atkgpu_cfar_test(c->stream, c->d_raw, (int)nb, (int)nf, (int)k, c->cfg.mode, c->N, c->cfg.os_k,
                         c->cfg.n_guard, c->alpha, c->alpha_ca, c->et, c->d_cells, c->d_ok, c->d_sided, c->d_reg,
                         c->d_X, c->d_thr, c->d_noise, c->d_hits) ||
        (k > 1 && atkgpu_d2d(c->stream, c->d_hist, c->d_raw + nf * nb, (k - 1) * nb * sizeof(double))) ||
        atkgpu_d2h(c->stream, c->thr, c->d_thr + (nf - 1) * nb, nb * sizeof(double)) ||
        atkgpu_d2h(c->stream, c->hits, c->d_hits, nf * nb) ||
        (k > 1 && atkgpu_d2h(c->stream, c->raw, c->d_raw + (k - 1) * nb, nf * nb * sizeof(double))))
        return ATKDSP_E_FFT;
    size_t nh = 0;
    if (c->gcap) {
        /* one frame averaged: the last frame's X whole (for atkdsp_cfar_last),
         * then X and noise at the hit bins only, gathered on the device */
        if (atkgpu_d2h(c->stream, c->X + (nf - 1) * nb, dX + (nf - 1) * nb, nb * sizeof(double)) ||
            atkgpu_sync(c->stream))
            return ATKDSP_E_FFT;
        const size_t tot = nf * nb;
        for (size_t i = 0; i < tot && nh <= c->gcap;) {
            if (i + 8 <= tot) {                                /* whole zero words at a time */
                uint64_t w;
                memcpy(&w, c->hits + i, 8);
                if (!w) { i += 8; continue; }
            }
            if (c->hits[i]) {
                if (nh < c->gcap) c->idx[nh] = (unsigned)i;
                ++nh;
            }
            ++i;
        }
    }
    if (c->gcap && nh <= c->gcap) {
        if (nh && (atkgpu_h2d(c->stream, c->d_idx, c->idx, nh * sizeof(unsigned)) ||
                   atkgpu_cfar_gather(c->stream, dX, c->d_noise, c->d_idx, nh, c->d_gath) ||
                   atkgpu_d2h(c->stream, c->gath, c->d_gath, 2 * nh * sizeof(double)) || atkgpu_sync(c->stream)))
            return ATKDSP_E_FFT;
        for (size_t q = 0; q < nh; ++q) {
            c->X[c->idx[q]] = c->gath[2 * q];
            c->noise[c->idx[q]] = c->gath[2 * q + 1];
        }
    } else if (atkgpu_d2h(c->stream, c->X, dX, nf * nb * sizeof(double)) ||
               atkgpu_d2h(c->stream, c->noise, c->d_noise, nf * nb * sizeof(double)) || atkgpu_sync(c->stream)) {
        return ATKDSP_E_FFT;
    }
    size_t nev = 0;
    for (size_t j = 0; j < nf; ++j) {
        const double tf = frame_t0 + (double)(f0 + j) * frame_dt;
        const size_t slot = (size_t)(c->frame % (int64_t)k);
        const double *pw = k > 1 ? c->raw + j * nb : NULL;
        if (k > 1) memcpy(c->ring + slot * nb, pw, nb * sizeof(double));
        c->ring_t[slot] = tf;
        c->ring_dt[slot] = frame_dt;
        c->st.frames++;
        c->seen++;
        c->frame++;
        if (c->seen < (int64_t)k) { c->st.warmup++; continue; }
        const double *X = c->X + j * nb, *noise = c->noise + j * nb;
        const uint8_t *hits = c->hits + j * nb;
        uint64_t nhit = 0;
        size_t i = 0;
        for (; i + 8 <= nb; i += 8) {                          /* hits are 0 / 1 bytes */
            uint64_t w;
            memcpy(&w, hits + i, 8);
            nhit += (uint64_t)__builtin_popcountll(w);
        }
        for (; i < nb; ++i) nhit += hits[i];
        c->st.cells_tested += c->n_ok;
        c->st.cell_hits += nhit;
        c->st.untestable += c->n_untestable;
        if (j == nf - 1) {                                     /* the call's last tested frame so far */
            memcpy(c->last_X, X, nb * sizeof(double));
            memcpy(c->last_thr, c->thr, nb * sizeof(double));
            memcpy(c->last_hits, hits, nb);
            c->has_last = 1;
        }
        nev += track_frame(c, X, noise, hits, pw, tf, frame_dt, ev + nev);
    }
    return (ptrdiff_t)nev;
}

ptrdiff_t atkdsp_cfar_process(atkdsp_cfar *c, const float *frames_db, size_t nframes, double frame_t0,
                              double frame_dt, atkdsp_cfar_event *ev, size_t cap) {
    if (!c || (!frames_db && nframes) || (!ev && nframes)) return ATKDSP_E_ARG;
    if (cap < atkdsp_cfar_events_max(c, nframes)) return ATKDSP_E_CAP;
    const size_t nb = c->nbins;
    size_t nev = 0;
    for (size_t f0 = 0; f0 < nframes; f0 += c->slab) {
        const size_t nf = nframes - f0 < c->slab ? nframes - f0 : c->slab;
        if (atkgpu_h2d(c->stream, c->d_db, frames_db + f0 * nb, nf * nb * sizeof(float))) return ATKDSP_E_FFT;
        const ptrdiff_t got = run_slab(c, nf, f0, frame_t0, frame_dt, ev + nev);
        if (got < 0) return got;
        nev += (size_t)got;
    }
    return (ptrdiff_t)nev;
}

size_t atk_cfar_nbins(const atkdsp_cfar *c) { return c->nbins; }
size_t atk_cfar_slab(const atkdsp_cfar *c) { return c->slab; }

int atk_cfar_stage_rows(atkdsp_cfar *c, const float *const *rows, size_t n) {
    if (!c || !rows || n > c->slab) return ATKDSP_E_ARG;
    for (size_t j = 0; j < n; ++j)
        if (atkgpu_d2d(c->stream, c->d_db + j * c->nbins, rows[j], c->nbins * sizeof(float))) return ATKDSP_E_FFT;
    return atkgpu_sync(c->stream) ? ATKDSP_E_FFT : ATKDSP_OK;
}

ptrdiff_t atk_cfar_process_staged(atkdsp_cfar *c, size_t n, double frame_t0, double frame_dt,
                                  atkdsp_cfar_event *ev, size_t cap) {
    if (!c || n > c->slab || (!ev && n)) return ATKDSP_E_ARG;
    if (cap < atkdsp_cfar_events_max(c, n)) return ATKDSP_E_CAP;
    return n ? run_slab(c, n, 0, frame_t0, frame_dt, ev) : 0;
}

ptrdiff_t atkdsp_cfar_flush(atkdsp_cfar *c, atkdsp_cfar_event *ev, size_t cap) {
    if (!c) return ATKDSP_E_ARG;
    size_t open = 0;
    for (int t = 0; t < c->ntracks; ++t) open += c->tracks[t].open == 1;
    if (open && (!ev || cap < open)) return ATKDSP_E_CAP;
    size_t nev = 0;
    for (int t = 0; t < c->ntracks; ++t)
        if (c->tracks[t].open) emit_close(c, &c->tracks[t], ATKDSP_CFAR_FLUSHED, &ev[nev++]);
    c->ntracks = 0;
    return (ptrdiff_t)nev;
}

int atkdsp_cfar_last(const atkdsp_cfar *c, float *avg_db, float *thr_db, uint8_t *hits) {
    if (!c) return ATKDSP_E_ARG;
    if (!c->has_last) return ATKDSP_E_EMPTY;
    for (size_t i = 0; i < c->nbins; ++i) {
        if (avg_db) avg_db[i] = (float)(10.0 * log10(c->last_X[i]));
        if (thr_db) thr_db[i] = (float)(10.0 * log10(c->last_thr[i]));
        if (hits) hits[i] = c->last_hits[i];
    }
    return ATKDSP_OK;
}

int atkdsp_cfar_info(const atkdsp_cfar *c, double *alpha, double *alpha_ca, double *et, atkdsp_cfar_config *cfg) {
    if (!c) return ATKDSP_E_ARG;
    if (alpha) *alpha = c->alpha;
    if (alpha_ca) *alpha_ca = c->alpha_ca;
    if (et) *et = c->et;
    if (cfg) *cfg = c->cfg;
    return ATKDSP_OK;
}

int atkdsp_cfar_describe(const atkdsp_cfar *c, char *buf, size_t cap) {
    if (!c || !buf || !cap) return ATKDSP_E_ARG;
    static const char *names[4] = {"CA", "OS", "GO", "SO"};
    char rank[48] = "", eff[64] = "";
    if (c->cfg.mode == ATKDSP_CFAR_OS) snprintf(rank, sizeof rank, " (rank %d of %d)", c->cfg.os_k, c->N);
    if (c->n_eff != (double)c->N) snprintf(eff, sizeof eff, " (%.1f effective, window-correlated)", c->n_eff);
    size_t masked = 0;
    for (size_t i = 0; i < c->nbins; ++i) masked += c->mask[i] != 0;
    snprintf(buf, cap,
             "CFAR %s%s: %d+%d reference cells%s, %d guard each side, Pfa %.3g per cell per frame, K = %g look(s) "
             "x %d frame(s), alpha %.4g (%.2f dB); open %d of %d, close after %d misses, merge gap %d, min %d "
             "bin(s); %zu bins, %zu masked, tracks %d/%d",
             names[c->cfg.mode], rank, c->cfg.n_ref, c->cfg.n_ref, eff, c->cfg.n_guard, c->cfg.pfa, c->cfg.looks,
             c->cfg.k_frames, c->alpha, 10.0 * log10(c->alpha), c->cfg.open_m, c->cfg.open_n, c->cfg.close_k,
             c->cfg.merge_gap, c->cfg.min_bins, c->nbins, masked, c->ntracks, c->cfg.max_tracks);
    return ATKDSP_OK;
}

int atkdsp_cfar_stats(const atkdsp_cfar *c, atkdsp_cfar_counters *st) {
    if (!c || !st) return ATKDSP_E_ARG;
    *st = c->st;
    st->tracks = (uint64_t)c->ntracks;
    return ATKDSP_OK;
}
