/* FFT plans, power spectrum, the 100 % POI reduce and per-bin stats
 * (reference include/atkdsp.h:379-442, src/fft.c:16-431).
 *
 * A plan owns NSLOT device slots (stream, input block, transform block, line
 * planes) made at create. Each call claims one with a CAS, so a plan may be
 * shared by caller threads as the reference allows (include/atkdsp.h:382-383);
 * nothing is allocated per call. Long inputs run in chunks of frames; the
 * per-bin accumulation runs over frames in ascending order in every chunk. */
#include <stdlib.h>
#include <string.h>
#include "internal.h"

#define NSLOT 4
#define CHUNK_SAMPLES ((size_t)1 << 21)

struct atkdsp_fft {
    size_t n;
    atkgpu_fftplan *plan;
    size_t fchunk;             /* frames per chunk at hop == n (>= 1) */
    size_t in_cap;             /* samples the input buffer holds      */
    atk_slot slot[NSLOT];
};

/* slot buffers: a = input samples, b = transforms, c = line planes
 * (double sum, double sum2, float line, max, min, window, 4 dB lines) and then
 * the frame groups' partials */
static size_t planes_head(size_t n) {
    return ((n * (2 * sizeof(double) + 8 * sizeof(float))) + 255) & ~(size_t)255;
}
static size_t planes_bytes(size_t n) { return planes_head(n) + atkgpu_spec_part_entries(n) * sizeof(double); }

atkdsp_fft *atkdsp_fft_create(size_t n) {
    if (n < 2) return NULL;
    if (atkgpu_init() != 0) return NULL;
    atkdsp_fft *p = (atkdsp_fft *)calloc(1, sizeof *p);
    if (!p) return NULL;
    p->n = n;
    p->plan = atkgpu_fftplan_create(n);
    if (!p->plan) { free(p); return NULL; }
    p->fchunk = CHUNK_SAMPLES / n;
    if (p->fchunk < 1) p->fchunk = 1;
    p->in_cap = p->fchunk * n;
    const size_t scr = atkgpu_fftplan_scratch(p->plan, p->fchunk);
    for (int i = 0; i < NSLOT; ++i) {
        if (atk_slot_init(&p->slot[i], p->in_cap * sizeof(atkdsp_cf32),
                          (p->fchunk * n + scr) * sizeof(atkdsp_cf32), planes_bytes(n))) {
            for (int j = 0; j < i; ++j) atk_slot_free(&p->slot[j]);
            atkgpu_fftplan_destroy(p->plan);
            free(p);
            return NULL;
        }
    }
    return p;
}

void atkdsp_fft_destroy(atkdsp_fft *p) {
    if (!p) return;
    for (int i = 0; i < NSLOT; ++i) atk_slot_free(&p->slot[i]);
    atkgpu_fftplan_destroy(p->plan);
    free(p);
}

size_t atkdsp_fft_length(const atkdsp_fft *p) { return p ? p->n : 0; }

typedef struct {
    float *line, *mx, *mn, *win, *db0, *db1, *db2, *db3;
    double *su, *s2;
    void *part;
} planes;

static planes slot_planes(const atk_slot *s, size_t n) {
    planes q;
    double *d = (double *)s->d_c;
    q.su = d;
    q.s2 = d + n;
    float *f = (float *)(d + 2 * n);
    q.line = f;
    q.mx = f + n;
    q.mn = f + 2 * n;
    q.win = f + 3 * n;
    q.db0 = f + 4 * n;
    q.db1 = f + 5 * n;
    q.db2 = f + 6 * n;
    q.db3 = f + 7 * n;
    q.part = (char *)s->d_c + planes_head(n);
    return q;
}

static void *xbuf(const atkdsp_fft *p, const atk_slot *s) { return s->d_b; }
static void *scrbuf(const atkdsp_fft *p, const atk_slot *s) {
    return (atkdsp_cf32 *)s->d_b + p->fchunk * p->n;
}

int atkdsp_fft_exec(const atkdsp_fft *cp, const atkdsp_cf32 *in, atkdsp_cf32 *out, int inverse) {
    if (!cp || !in || !out) return ATKDSP_E_ARG;
    atkdsp_fft *p = (atkdsp_fft *)cp;    /* slots are scratch, not plan state */
    const int si = atk_claim(p->slot, NSLOT);
    atk_slot *s = &p->slot[si];
    const size_t n = p->n;
    int rc = atk_put(s->stream, s->d_a, in, n * sizeof(atkdsp_cf32));
    if (!rc) rc = atkgpu_fft_batch(s->stream, p->plan, s->d_a, n, NULL, 1, inverse, xbuf(p, s), scrbuf(p, s));
    if (!rc) rc = atk_get(s->stream, out, xbuf(p, s), n * sizeof(atkdsp_cf32));
    if (!rc) rc = atkgpu_sync(s->stream);
    atk_release(p->slot, si);
    return rc ? ATKDSP_E_FFT : ATKDSP_OK;
}

int atkdsp_power_db(const atkdsp_fft *cp, const atkdsp_cf32 *in, const float *window, float *out_db) {
    if (!cp || !in || !out_db) return ATKDSP_E_ARG;
    atkdsp_fft *p = (atkdsp_fft *)cp;
    const int si = atk_claim(p->slot, NSLOT);
    atk_slot *s = &p->slot[si];
    const size_t n = p->n;
    planes q = slot_planes(s, n);
    int rc = atk_put(s->stream, s->d_a, in, n * sizeof(atkdsp_cf32));
    if (!rc && window) rc = atkgpu_h2d(s->stream, q.win, window, n * sizeof(float));
    if (!rc) rc = atkgpu_fft_batch(s->stream, p->plan, s->d_a, n, window ? q.win : NULL, 1, 0, xbuf(p, s),
                                   scrbuf(p, s));
    if (!rc) rc = atkgpu_power_db(s->stream, xbuf(p, s), n, q.db0);
    if (!rc) rc = atkgpu_d2h(s->stream, out_db, q.db0, n * sizeof(float));
    if (!rc) rc = atkgpu_sync(s->stream);
    atk_release(p->slot, si);
    return rc ? ATKDSP_E_FFT : ATKDSP_OK;
}

/* Runs every frame of `in` through the plan chunk by chunk, calling the
 * accumulate step per chunk; returns 0 or an error. */
static int frames_run(atkdsp_fft *p, atk_slot *s, const atkdsp_cf32 *in, size_t hop, long long frames,
                      const float *d_win, int stats, int detector, planes q) {
    const size_t n = p->n;
    /* frames per chunk so that (F-1)*hop + n samples fit the input buffer */
    size_t F = (p->in_cap - n) / hop + 1;
    if (F > p->fchunk) F = p->fchunk;
    int rc = 0;
    for (long long f0 = 0; f0 < frames && !rc; f0 += (long long)F) {
        const size_t nf = (size_t)(frames - f0) < F ? (size_t)(frames - f0) : F;
        const size_t span = (nf - 1) * hop + n;
        rc = atk_put(s->stream, s->d_a, in + (size_t)f0 * hop, span * sizeof(atkdsp_cf32));
        if (!rc) rc = atkgpu_fft_batch(s->stream, p->plan, s->d_a, hop, d_win, nf, 0, xbuf(p, s), scrbuf(p, s));
        if (!rc) {
            if (stats) rc = atkgpu_spec_stats_accum(s->stream, xbuf(p, s), n, nf, f0 == 0, q.mx, q.mn, q.su, q.s2,
                                                    q.part);
            else rc = atkgpu_spec_accum(s->stream, xbuf(p, s), n, nf, detector, f0 == 0, q.line, q.part);
        }
        /* the next chunk's copy reuses the input buffer: the stream orders it */
    }
    return rc;
}

ptrdiff_t atkdsp_spectrum_reduce(const atkdsp_fft *cp, const atkdsp_cf32 *in, size_t n_samples, size_t hop,
                                 const float *window, int detector, float *out_line) {
    if (!cp || !in || !out_line) return ATKDSP_E_ARG;
    if (detector < ATKDSP_DET_MAX || detector > ATKDSP_DET_AVG) return ATKDSP_E_ARG;
    atkdsp_fft *p = (atkdsp_fft *)cp;
    const size_t n = p->n;
    if (hop == 0 || hop > n) hop = n;
    if (n_samples < n) return 0;
    const long long frames = (long long)((n_samples - n) / hop + 1);
    const int si = atk_claim(p->slot, NSLOT);
    atk_slot *s = &p->slot[si];
    planes q = slot_planes(s, n);
    int rc = 0;
    if (window) rc = atkgpu_h2d(s->stream, q.win, window, n * sizeof(float));
    if (!rc) rc = frames_run(p, s, in, hop, frames, window ? q.win : NULL, 0, detector, q);
    if (!rc) rc = atkgpu_spec_reduce_db(s->stream, q.line, n, detector, frames, q.db0);
    if (!rc) rc = atkgpu_d2h(s->stream, out_line, q.db0, n * sizeof(float));
    if (!rc) rc = atkgpu_sync(s->stream);
    atk_release(p->slot, si);
    return rc ? ATKDSP_E_FFT : (ptrdiff_t)frames;
}

ptrdiff_t atkdsp_spectrum_stats(const atkdsp_fft *cp, const atkdsp_cf32 *in, size_t n_samples, size_t hop,
                                const float *window, float *max_db, float *avg_db, float *min_db, float *sk) {
    if (!cp || !in) return ATKDSP_E_ARG;
    if (!max_db && !avg_db && !min_db && !sk) return ATKDSP_E_ARG;
    atkdsp_fft *p = (atkdsp_fft *)cp;
    const size_t n = p->n;
    if (hop == 0 || hop > n) hop = n;
    if (n_samples < n) return 0;
    const long long frames = (long long)((n_samples - n) / hop + 1);
    const int si = atk_claim(p->slot, NSLOT);
    atk_slot *s = &p->slot[si];
    planes q = slot_planes(s, n);
    int rc = 0;
    if (window) rc = atkgpu_h2d(s->stream, q.win, window, n * sizeof(float));
    if (!rc) rc = frames_run(p, s, in, hop, frames, window ? q.win : NULL, 1, 0, q);
    if (!rc) rc = atkgpu_spec_stats_db(s->stream, n, frames, q.mx, q.mn, q.su, q.s2, max_db ? q.db0 : NULL,
                                       avg_db ? q.db1 : NULL, min_db ? q.db2 : NULL, sk ? q.db3 : NULL);
    if (!rc && max_db) rc = atkgpu_d2h(s->stream, max_db, q.db0, n * sizeof(float));
    if (!rc && avg_db) rc = atkgpu_d2h(s->stream, avg_db, q.db1, n * sizeof(float));
    if (!rc && min_db) rc = atkgpu_d2h(s->stream, min_db, q.db2, n * sizeof(float));
    if (!rc && sk) rc = atkgpu_d2h(s->stream, sk, q.db3, n * sizeof(float));
    if (!rc) rc = atkgpu_sync(s->stream);
    atk_release(p->slot, si);
    return rc ? ATKDSP_E_FFT : (ptrdiff_t)frames;
}
