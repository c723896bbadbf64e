/* Detection on a spectrum line (reference include/atkdsp.h:465-486,
 * src/detect.c:4-114): the drop-in median and detect_channels. Both run the
 * device kernels of csrc/dev/detect.hip on a pool slot: the line goes up,
 * the records (or the median) come back. The records arrive in bin order and
 * are put strongest first here with the reference's own comparator and the C
 * library's qsort, so ties between equal SNRs come out as the reference's do. */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "internal.h"

_Static_assert(sizeof(atkgpu_chan) == sizeof(atkdsp_channel) &&
                   offsetof(atkgpu_chan, snr_db) == offsetof(atkdsp_channel, snr_db),
               "device channel records are atkdsp_channel");

#define DETECT_MAX_BINS ((size_t)1 << 21)
#define MEDIAN_MAX ((size_t)1 << 23)

float atkdsp_median(const float *x, size_t n, float *scratch) {
    if (!x || !scratch || n == 0) return -120.0f;
    if (n > MEDIAN_MAX) return NAN;
    atk_pool *pool = atk_pool_get();
    if (!pool || n * sizeof(float) > pool->slot[0].bytes_a) return NAN;
    const int si = atk_claim(pool->slot, pool->nslot);
    atk_slot *s = &pool->slot[si];
    float m = NAN;
    float *d_m = (float *)s->d_small;
    if (atkgpu_h2d(s->stream, s->d_a, x, n * sizeof(float)) ||
        atkgpu_median(s->stream, (const float *)s->d_a, n, d_m) ||
        atkgpu_d2h(s->stream, &m, d_m, sizeof(float)) || atkgpu_sync(s->stream))
        m = NAN;
    atk_release(pool->slot, si);
    return m;
}

/* strongest first: the reference's comparator (src/detect.c:52-56) */
static int snr_desc(const void *a, const void *b) {
    const float x = ((const atkdsp_channel *)a)->snr_db;
    const float y = ((const atkdsp_channel *)b)->snr_db;
    return (x < y) - (x > y);
}

void atk_sort_channels(atkdsp_channel *c, size_t n) {
    if (n > 1) qsort(c, n, sizeof *c, snr_desc);
}

ptrdiff_t atkdsp_detect_channels(const float *line, size_t n, float floor_db, float threshold_db, int gap_bins,
                                 int min_run, float *smooth_scratch, atkdsp_channel *out, size_t out_cap) {
    if (!line || !smooth_scratch || (!out && out_cap)) return ATKDSP_E_ARG;
    if (n == 0) return 0;
    if (floor_db != floor_db) return ATKDSP_E_ARG;     /* the caller computes the floor (atkdsp_median) */
    if (n > DETECT_MAX_BINS) return ATKDSP_E_ARG;
    if (gap_bins < 2) gap_bins = 2;
    if (min_run < 2) min_run = 2;
    atk_pool *pool = atk_pool_get();
    if (!pool) return ATKDSP_E_FFT;
    /* the line at the start of buffer a, its records after it (runs need two
     * ones and, between runs, at least two zeros: at most n / 4 + 1 of them);
     * the smoothed line and mask in buffer b; the count in the small buffer */
    const size_t line_bytes = (n * sizeof(float) + 255) & ~(size_t)255;
    const size_t rec_room = (pool->slot[0].bytes_a - line_bytes) / sizeof(atkgpu_chan);
    const size_t cap = out_cap < rec_room ? out_cap : rec_room;
    if (n * sizeof(float) > pool->slot[0].bytes_b || n / 4 + 1 > rec_room) return ATKDSP_E_ARG;
    const int si = atk_claim(pool->slot, pool->nslot);
    atk_slot *s = &pool->slot[si];
    atkgpu_chan *d_rec = (atkgpu_chan *)((char *)s->d_a + line_bytes);
    unsigned long long *d_found = (unsigned long long *)s->d_small;
    unsigned long long found = 0;
    int rc = ATKDSP_OK;
    if (atkgpu_h2d(s->stream, s->d_a, line, n * sizeof(float)) ||
        atkgpu_detect_lines(s->stream, (const float *)s->d_a, n, n, 1, floor_db, 0, threshold_db, gap_bins, min_run,
                            (float *)s->d_b, d_rec, cap, d_found, NULL) ||
        atkgpu_d2h(s->stream, &found, d_found, sizeof found) || atkgpu_sync(s->stream))
        rc = ATKDSP_E_FFT;
    const size_t kept = rc ? 0 : (found < cap ? (size_t)found : cap);
    if (!rc && kept && (atkgpu_d2h(s->stream, out, d_rec, kept * sizeof(atkgpu_chan)) || atkgpu_sync(s->stream)))
        rc = ATKDSP_E_FFT;
    atk_release(pool->slot, si);
    if (rc) return rc;
    if (found > out_cap) return ATKDSP_E_CAP;          /* out holds the first out_cap found, in bin order */
    atk_sort_channels(out, kept);
    return (ptrdiff_t)kept;
}
