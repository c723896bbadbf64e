/* CFAR over spectrum frames and the burst tracker (reference
 * include/atkdsp.h:1733-1925, src/cfar.c:270-819): the atkdsp_cfar handle.
 *
 * Per call, frames go to the device in slabs; the device turns dB into
 * power, averages the last k_frames frames and runs the CA / OS / GO / SO test
 * of every bin against its reference cells (csrc/dev/cfar.hip), and returns
 * each frame's average, threshold, noise estimate, hits and power. The
 * tracker then follows, frame by frame, on the host. Its rules are the
 * header's; its steps here are
 *   runs       the frame's hit bins cut into runs wherever more than
 *              merge_gap non-hit bins separate two hits; runs narrower than
 *              min_bins are ignored; each run's excess-weighted bin sums, its
 *              power and noise sums and its strongest bin are taken in bin order;
 *   assign     each run to the track (among those alive when the frame came)
 *              whose latest span it overlaps most, within merge_gap + 1 bins;
 *              equal overlaps go to the older track;
 *   spawn      a pending track for every run left over, in run order, up to
 *              max_tracks (the rest are counted as dropped);
 *   observe    a track's runs of the frame merged into one observation;
 *   step       every track in id order: its hit history over open_n frames,
 *              OPEN at open_m hits, forgotten when pending with none left,
 *              CLOSE after close_k misses in a row, UPDATE when an open
 *              track's span grows; its first and last times, refined with
 *              k_frames > 1 by the single-frame band test at
 *              ATKDSP_CFAR_EDGE_PFA (T0_COARSE / T1_COARSE when it fails).
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

enum { TRK_PENDING = 0, TRK_OPEN = 1 };

/* a tracked emitter */
typedef struct {
    uint32_t id;
    int state;                /* TRK_PENDING or TRK_OPEN */
    int unseen;               /* created this frame, no observation applied yet */
    uint64_t hist;            /* hit frames over the last open_n, newest in bit 0 */
    int miss_run;             /* frames in a row without a hit */
    int32_t hits_total;
    int32_t now_lo, now_hi;   /* the latest observation's span */
    int32_t lo, hi;           /* the union of every observed span */
    double wsum, esum;        /* sum of bin * excess and of excess, over the track's life */
    double pk_x, pk_n;        /* the strongest averaged bin seen, and its noise */
    double best_band;         /* the best frame's power / noise over its hit bins */
    uint32_t flags;
    double t_start;
    int64_t f_hit, f_raw;     /* the last frame with a hit / whose single frame passed the band test */
    double end_hit, dt_hit, end_raw;
} cfar_trk;

/* one frame's observation of a track, the merge of its runs */
typedef struct {
    int any;
    int32_t lo, hi;
    double xsum, nsum;
} cfar_obs;

/* a run of hit bins in one frame */
typedef struct {
    int32_t lo, hi;
    double wsum, esum, xsum, nsum, pk_x, pk_n;
    int owner;                /* the track it belongs to, -1 before assignment */
} cfar_run;

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
    cfar_trk *trk;            /* live tracks in id order */
    cfar_obs *obs;            /* this frame's observation of trk[i] */
    int ntrk;
    uint32_t next_id;
    cfar_run *runs;
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
    free(c->beta); free(c->trk); free(c->obs); free(c->runs);
    free(c);
}

/* What a configuration implies: the OS rank when left to the default (the
 * ceiling of 3N/4), K = looks x frames averaged, the effective cell count,
 * and the thresholds (GO and SO fall back to CA's where a side borrowed). */
static int derive(atkdsp_cfar *c) {
    atkdsp_cfar_config *g = &c->cfg;
    c->N = 2 * g->n_ref;
    if (g->mode == ATKDSP_CFAR_OS && !g->os_k) g->os_k = (3 * c->N + 3) / 4;
    c->K = g->looks * (double)g->k_frames;
    c->n_eff = atkdsp_cfar_n_eff(g->n_ref, g->rho2);
    if (atkdsp_cfar_alpha(g->mode, g->n_ref, g->pfa, c->K, g->os_k, g->rho2, &c->alpha, &c->et) != ATKDSP_OK)
        return 0;
    c->alpha_ca = c->alpha;
    if (g->mode == ATKDSP_CFAR_GO || g->mode == ATKDSP_CFAR_SO)
        return atkdsp_cfar_alpha(ATKDSP_CFAR_CA, g->n_ref, g->pfa, c->K, 0, g->rho2, &c->alpha_ca, NULL) == ATKDSP_OK;
    return 1;
}

atkdsp_cfar *atkdsp_cfar_create(size_t nbins, const atkdsp_cfar_config *cfg) {
    if (!cfg || !config_ok(nbins, cfg)) return NULL;
    if (atkgpu_init() != 0) return NULL;
    atkdsp_cfar *c = (atkdsp_cfar *)calloc(1, sizeof *c);
    if (!c) return NULL;
    c->cfg = *cfg;
    c->nbins = nbins;
    if (!derive(c)) {
        free(c);
        return NULL;
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
    c->trk = (cfar_trk *)calloc((size_t)c->cfg.max_tracks, sizeof(cfar_trk));
    c->obs = (cfar_obs *)calloc((size_t)c->cfg.max_tracks, sizeof(cfar_obs));
    c->runs = (cfar_run *)calloc(nb, sizeof(cfar_run));
    int ok = c->mask && c->ok && c->sided && c->reg && c->cells && c->X && c->thr && c->noise && (c->raw || k == 1) && c->hits &&
             c->ring && c->ring_t && c->ring_dt && c->last_X && c->last_thr && c->last_hits && c->beta &&
             c->trk && c->obs && c->runs && !atkgpu_stream_create(&c->stream);
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
    c->ntrk = 0;
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

/* the event flags of a track: its own (coarse times) plus EDGE when a bin of
 * its span was tested with borrowed cells or not at all, MASK_ADJ when a
 * masked bin lies within merge_gap + 1 of the span */
static uint32_t span_flags(const atkdsp_cfar *c, const cfar_trk *t) {
    uint32_t f = t->flags;
    int edge = 0, near_mask = 0;
    for (int32_t b = t->lo; b <= t->hi && !edge; ++b) edge = !c->ok[b] || !c->sided[b];
    const long reach = (long)c->cfg.merge_gap + 1;
    const long a = (long)t->lo - reach > 0 ? (long)t->lo - reach : 0;
    const long z = (long)t->hi + reach < (long)c->nbins - 1 ? (long)t->hi + reach : (long)c->nbins - 1;
    for (long b = a; b <= z && !near_mask; ++b) near_mask = c->mask[b] != 0;
    if (edge) f |= ATKDSP_CFAR_EDGE;
    if (near_mask) f |= ATKDSP_CFAR_MASK_ADJ;
    return f;
}

static double to_db(double p) { return 10.0 * log10(p); }

static void make_event(const atkdsp_cfar *c, const cfar_trk *t, int type, double t1, uint32_t extra,
                       atkdsp_cfar_event *e) {
    memset(e, 0, sizeof *e);
    e->type = type;
    e->id = t->id;
    e->flags = span_flags(c, t) | extra;
    e->lo_bin = t->lo;
    e->hi_bin = t->hi;
    e->n_hits = t->hits_total;
    e->t0 = t->t_start;
    e->t1 = t1;
    /* the excess-weighted centre, or the span's middle when nothing exceeded */
    e->fc_bin = t->esum > 0.0 ? t->wsum / t->esum : 0.5 * ((double)t->lo + (double)t->hi);
    const float level = t->pk_x > 0.0 ? (float)to_db(t->pk_x) : -INFINITY;
    const float noise = t->pk_n > 0.0 ? (float)to_db(t->pk_n) : -INFINITY;
    e->level_db = level;
    e->noise_db = noise;
    e->snr_db = t->pk_n > 0.0 ? level - noise : INFINITY;
    e->band_snr_db = t->best_band > 0.0 ? (float)to_db(t->best_band) : -INFINITY;
}

/* CLOSE: the end time is the last hit frame's end; averaged over k frames it
 * is the last raw frame that passed the band test when that frame belongs to
 * the last hit's averaging window, else half the window before the hit end
 * (T1_COARSE) */
static void close_event(atkdsp_cfar *c, cfar_trk *t, uint32_t extra, atkdsp_cfar_event *e) {
    const int k = c->cfg.k_frames;
    double t1 = t->end_hit;
    if (k > 1) {
        if (t->f_raw >= t->f_hit - (k - 1)) {
            t1 = t->end_raw;
        } else {
            t1 = t->end_hit - 0.5 * (double)(k - 1) * t->dt_hit;
            t->flags |= ATKDSP_CFAR_T1_COARSE;
        }
    }
    c->st.closed++;
    make_event(c, t, ATKDSP_CFAR_CLOSE, t1, extra, e);
}

/* Close the run [lo, last] of one frame: its statistics over its hit bins in
 * bin order; runs narrower than min_bins are not kept. */
static void add_run(atkdsp_cfar *c, size_t *n, long lo, long last, const uint8_t *hits, const double *X,
                    const double *noise) {
    if (last - lo + 1 < c->cfg.min_bins) return;
    cfar_run *r = &c->runs[(*n)++];
    r->lo = (int32_t)lo;
    r->hi = (int32_t)last;
    r->wsum = r->esum = r->xsum = r->nsum = 0.0;
    r->owner = -1;
    long best = -1;
    for (long b = lo; b <= last; ++b) {
        if (!hits[b]) continue;
        const double ex = X[b] > noise[b] ? X[b] - noise[b] : 0.0;
        r->wsum += (double)b * ex;
        r->esum += ex;
        r->xsum += X[b];
        r->nsum += noise[b];
        if (best < 0 || X[b] > X[best]) best = b;
    }
    r->pk_x = X[best];
    r->pk_n = noise[best];
}

/* Runs of hit bins in one frame, in c->runs: one pass over the hit bins; a
 * hit more than merge_gap non-hit bins after the previous one starts a new
 * run. Returns the number kept. */
static size_t frame_runs(atkdsp_cfar *c, const uint8_t *hits, const double *X, const double *noise) {
    const long nb = (long)c->nbins, gap = c->cfg.merge_gap;
    size_t n = 0;
    long lo = -1, last = -1;
    long i = 0;
    while (i < nb) {
        if (i + 8 <= nb) {                       /* a word of non-hit bins at a time */
            uint64_t w;
            memcpy(&w, hits + i, 8);
            if (!w) {
                i += 8;
                continue;
            }
        }
        if (hits[i]) {
            if (lo >= 0 && i - last - 1 > gap) {
                add_run(c, &n, lo, last, hits, X, noise);
                lo = -1;
            }
            if (lo < 0) lo = i;
            last = i;
        }
        ++i;
    }
    if (lo >= 0) add_run(c, &n, lo, last, hits, X, noise);
    return n;
}

/* the overlap of a run with a track's latest span (may be zero or negative
 * for a run that is merely within reach) */
static long overlap(const cfar_run *r, const cfar_trk *t) {
    const long hi = r->hi < t->now_hi ? r->hi : t->now_hi;
    const long lo = r->lo > t->now_lo ? r->lo : t->now_lo;
    return hi - lo + 1;
}

/* Each run to the live track it overlaps most (within reach), the older one
 * on equal overlap; then a new pending track for each run left over. */
static void assign_runs(atkdsp_cfar *c, size_t nrun, double dt) {
    const int alive = c->ntrk;
    const long reach = (long)c->cfg.merge_gap + 1;
    for (size_t q = 0; q < nrun; ++q) {
        cfar_run *r = &c->runs[q];
        long best = 0;
        for (int t = 0; t < alive; ++t) {
            const cfar_trk *tr = &c->trk[t];
            if ((long)r->lo - (long)tr->now_hi > reach || (long)tr->now_lo - (long)r->hi > reach) continue;
            const long ov = overlap(r, tr);
            if (r->owner < 0 || ov > best) {
                r->owner = t;
                best = ov;
            }
        }
    }
    for (size_t q = 0; q < nrun; ++q) {
        cfar_run *r = &c->runs[q];
        if (r->owner >= 0) continue;
        if (c->ntrk == c->cfg.max_tracks) {
            c->st.dropped++;
            continue;
        }
        cfar_trk *tr = &c->trk[c->ntrk];
        memset(tr, 0, sizeof *tr);
        tr->id = c->next_id;
        c->next_id = c->next_id == UINT32_MAX ? 1 : c->next_id + 1;
        tr->lo = tr->now_lo = r->lo;
        tr->hi = tr->now_hi = r->hi;
        tr->pk_x = -1.0;
        tr->f_hit = tr->f_raw = -1;
        tr->dt_hit = dt;
        tr->unseen = 1;
        r->owner = c->ntrk++;
    }
}

/* A track's runs of this frame, in run order, into its observation and its
 * life-long sums. */
static void observe(atkdsp_cfar *c, size_t nrun) {
    memset(c->obs, 0, (size_t)c->ntrk * sizeof *c->obs);
    for (size_t q = 0; q < nrun; ++q) {
        const cfar_run *r = &c->runs[q];
        if (r->owner < 0) continue;
        cfar_obs *o = &c->obs[r->owner];
        cfar_trk *tr = &c->trk[r->owner];
        if (!o->any) {
            o->any = 1;
            o->lo = r->lo;
            o->hi = r->hi;
        } else {
            o->lo = r->lo < o->lo ? r->lo : o->lo;
            o->hi = r->hi > o->hi ? r->hi : o->hi;
        }
        o->xsum += r->xsum;
        o->nsum += r->nsum;
        tr->wsum += r->wsum;
        tr->esum += r->esum;
        if (r->pk_x > tr->pk_x) {
            tr->pk_x = r->pk_x;
            tr->pk_n = r->pk_n;
        }
    }
}

/* The start time of a track seen for the first time in frame F: that frame's
 * start with one frame averaged; otherwise the oldest raw frame of the
 * averaging window whose band passes, or half the window before (T0_COARSE). */
static double start_time(atkdsp_cfar *c, cfar_trk *tr, const double *noise, int64_t F, double tf, double dt) {
    const int k = c->cfg.k_frames;
    if (k == 1) return tf;
    for (int age = k - 1; age >= 0; --age) {
        const size_t slot = (size_t)(((F - age) % k + k) % k);
        if (band_pass(c, c->ring + slot * c->nbins, noise, tr->now_lo, tr->now_hi)) return c->ring_t[slot];
    }
    tr->flags |= ATKDSP_CFAR_T0_COARSE;
    return tf - 0.5 * (double)(k - 1) * dt;
}

/* One frame of one track; returns 1 to keep it, 0 when it ends. An event, if
 * any, goes to *e and *emitted is set. */
static int step_track(atkdsp_cfar *c, cfar_trk *tr, const cfar_obs *o, const double *noise, const double *pw,
                      int64_t F, double tf, double dt, atkdsp_cfar_event *e, int *emitted) {
    const uint64_t keep_bits = c->cfg.open_n >= 64 ? ~(uint64_t)0 : (((uint64_t)1 << c->cfg.open_n) - 1);
    int grew = 0;
    *emitted = 0;
    tr->hist = (tr->hist << 1) & keep_bits;
    if (o->any) {
        tr->hist |= 1u;
        tr->miss_run = 0;
        tr->hits_total++;
        tr->now_lo = o->lo;
        tr->now_hi = o->hi;
        grew = !tr->unseen && (o->lo < tr->lo || o->hi > tr->hi);
        tr->lo = o->lo < tr->lo ? o->lo : tr->lo;
        tr->hi = o->hi > tr->hi ? o->hi : tr->hi;
        if (o->nsum > 0.0 && o->xsum / o->nsum > tr->best_band) tr->best_band = o->xsum / o->nsum;
        tr->f_hit = F;
        tr->end_hit = tf + dt;
        tr->dt_hit = dt;
        if (tr->unseen) {
            tr->unseen = 0;
            tr->t_start = start_time(c, tr, noise, F, tf, dt);
        }
    } else {
        tr->miss_run++;
    }
    if (pw && band_pass(c, pw, noise, tr->now_lo, tr->now_hi)) {
        tr->f_raw = F;
        tr->end_raw = tf + dt;
    }
    if (tr->state == TRK_PENDING) {
        if (__builtin_popcountll(tr->hist) >= c->cfg.open_m) {
            tr->state = TRK_OPEN;
            c->st.opened++;
            make_event(c, tr, ATKDSP_CFAR_OPEN, NAN, 0, e);
            *emitted = 1;
            return 1;
        }
        return tr->hist != 0;                     /* a pending track with no hit left is forgotten */
    }
    if (tr->miss_run >= c->cfg.close_k) {
        close_event(c, tr, 0, e);
        *emitted = 1;
        return 0;
    }
    if (grew) {
        c->st.updates++;
        make_event(c, tr, ATKDSP_CFAR_UPDATE, NAN, 0, e);
        *emitted = 1;
    }
    return 1;
}

/* one tested frame: X, noise, hits and pw (this frame's power; NULL with one
 * frame averaged, where no band test runs) are the device's results, X and
 * noise valid at least at the hit bins (the only ones read here); returns the
 * events written */
static size_t track_frame(atkdsp_cfar *c, const double *X, const double *noise, const uint8_t *hits,
                          const double *pw, double tf, double dt, atkdsp_cfar_event *ev) {
    const int64_t F = c->frame - 1;                          /* this frame's number */
    const size_t nrun = frame_runs(c, hits, X, noise);
    assign_runs(c, nrun, dt);
    observe(c, nrun);
    size_t nev = 0;
    int live = 0;
    for (int t = 0; t < c->ntrk; ++t) {
        int emitted;
        const int keep = step_track(c, &c->trk[t], &c->obs[t], noise, pw, F, tf, dt, &ev[nev], &emitted);
        nev += (size_t)emitted;
        if (keep) {
            if (live != t) c->trk[live] = c->trk[t];
            ++live;
        }
    }
    c->ntrk = live;
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
        if (atk_put(c->stream, c->d_db, frames_db + f0 * nb, nf * nb * sizeof(float))) return ATKDSP_E_FFT;
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
    size_t nopen = 0;
    for (int t = 0; t < c->ntrk; ++t) nopen += c->trk[t].state == TRK_OPEN;
    if (nopen && (!ev || cap < nopen)) return ATKDSP_E_CAP;
    size_t nev = 0;
    for (int t = 0; t < c->ntrk; ++t)
        if (c->trk[t].state == TRK_OPEN) close_event(c, &c->trk[t], ATKDSP_CFAR_FLUSHED, &ev[nev++]);
    c->ntrk = 0;
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
    const double v[3] = {c->alpha, c->alpha_ca, c->et};
    double *dst[3] = {alpha, alpha_ca, et};
    for (int i = 0; i < 3; ++i)
        if (dst[i]) *dst[i] = v[i];
    if (cfg) *cfg = c->cfg;                     /* with the OS rank resolved */
    return ATKDSP_OK;
}

/* A one-line summary: the statistic, the cells, the false-alarm design, the
 * tracker's rules and the handle's state. */
int atkdsp_cfar_describe(const atkdsp_cfar *c, char *buf, size_t cap) {
    if (!c || !buf || !cap) return ATKDSP_E_ARG;
    static const char *const stat_name[4] = {"cell-averaging", "ordered-statistic", "greatest-of", "smallest-of"};
    const atkdsp_cfar_config *g = &c->cfg;
    size_t masked = 0;
    for (size_t i = 0; i < c->nbins; ++i) masked += c->mask[i] != 0;
    char rank[40] = "";
    if (g->mode == ATKDSP_CFAR_OS) snprintf(rank, sizeof rank, ", rank %d", g->os_k);
    snprintf(buf, cap,
             "cfar[%s%s] bins=%zu masked=%zu ref=2x%d guard=%d n_eff=%.2f pfa=%.3g looks=%g frames=%d "
             "alpha=%.6g (%+.2f dB) open=%d/%d close=%d gap=%d min=%d tracks=%d/%d",
             stat_name[g->mode], rank, c->nbins, masked, g->n_ref, g->n_guard, c->n_eff, g->pfa, g->looks,
             g->k_frames, c->alpha, 10.0 * log10(c->alpha), g->open_m, g->open_n, g->close_k, g->merge_gap,
             g->min_bins, c->ntrk, g->max_tracks);
    return ATKDSP_OK;
}

int atkdsp_cfar_stats(const atkdsp_cfar *c, atkdsp_cfar_counters *st) {
    if (!c || !st) return ATKDSP_E_ARG;
    atkdsp_cfar_counters now = c->st;
    now.tracks = (uint64_t)c->ntrk;             /* live tracks, pending or open */
    *st = now;
    return ATKDSP_OK;
}
