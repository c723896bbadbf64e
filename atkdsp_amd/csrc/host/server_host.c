/* The serving loop (include/atkgpu.h): many independent receiver streams,
 * each with its own format and filterbank shape, sending chunks of any size.
 *
 *   admission   a stream's device state (two history rows per element, taps,
 *               an output ring, and for a coherent array its W1/W2) is
 *               allocated when it is added and charged against a fixed HBM
 *               budget; a stream that does not fit is refused.
 *   submit      reserves a ticket and staging space under the lock, copies the
 *               chunk's bytes into the pinned staging ring (which the device
 *               reads directly, mapped) outside it, then publishes the chunk.
 *               Submitters of different streams copy in parallel. Each element
 *               of an array chunk starts on a 64-byte boundary. It allocates
 *               nothing and makes no HIP call. submit_ring takes the chunk
 *               from a block ring instead: staging first, a zero-copy peek
 *               copied into it, and the chunk published only once the ring's
 *               release shows the slot was not rewritten meanwhile.
 *   dispatcher  a thread forms batches continuously: at most one chunk per
 *               stream (each stream's chunks stay in order), taken round-robin
 *               over the streams that have work, while earlier batches still
 *               run. A batch is
 *                 on the ingest HIP stream: one ragged launch that copies each
 *                 chunk's kept prefix (the last keep samples of the stream's
 *                 previous row) and unpacks its bytes after it, every format
 *                 and element together;
 *                 on the compute HIP stream, after that: one fused filterbank
 *                 launch per (M, os, taps) class, and a W1 and a W2 launch
 *                 per coherent array stream;
 *                 on the display HIP stream, also after the ingest and beside
 *                 the filterbank: per (FFT length, detector) class, a frames
 *                 launch (one FFT team per frame) and a lines launch (one
 *                 block per line segment); the compute stream joins it before
 *                 the batch's completion event.
 *               Every launch reads one descriptor per block from mapped host
 *               memory (a single PCIe round trip, no search).
 *               Batch k's ingest writes only rows that batches up to k-2 read
 *               (a stream alternates two rows, one chunk per batch), so it
 *               waits for batch k-2's compute and overlaps batch k-1's: the
 *               PCIe-bound unpack runs beside the HBM-bound filterbank.
 *   completion  a thread polls each batch's event in order (sleeping a few
 *               microseconds between polls, so no core is held spinning),
 *               stamps the host monotonic clock and records each chunk's
 *               latency (completion minus submission), then frees staging.
 *
 * Every chunk of a stream is computed by the same kernels with the same
 * per-hop arithmetic whatever else is in its batch, so a stream's output is
 * identical alone and in the mix. The offset is fixed per stream (no
 * per-chunk re-estimation), which keeps the output independent of chunking.
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <sys/prctl.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <time.h>
#include "internal.h"
#include "../../../include/atkgpu.h"

int atkgpu_pfb_prototype(double *h, unsigned nchan, unsigned K, double atten_db);

#ifndef NINFLIGHT
#define NINFLIGHT 3            /* batches launched and not yet complete */
#endif
/* ordering events: batch k records oev[k % NOEV] after its compute, and batch
 * k + 2's ingest waits on it. Batch k + NOEV, which records it again, is
 * launched only once batch k + NOEV - NINFLIGHT = k + 3 is complete, and so
 * after batch k + 2's ingest (and its wait) has run: an event is never
 * recorded again while a wait on it may still be pending */
#ifndef NOEV
#define NOEV (2 * NINFLIGHT)
#endif
#define POLL_NS 10000          /* completion poll interval */
#define QCAP 256               /* pending chunks per stream */
/* chunk latencies: a histogram of log-spaced buckets, each 1/256 wider
 * than the last (0.4 % apart), from 0.1 us to ~8 s; every chunk is counted */
#define LAT_NB 4608
#define LAT_LO_MS 1e-4
#define LAT_LOG_STEP 0.00389864041565732   /* log(1 + 1/256) */
#define NEL_MAX 16             /* elements of an array stream */
#define NPLANS 7               /* FFT twiddles for 64..4096 points */
/* blocks of the ingest launch (each strides over the batch's ingest blocks):
 * 4 waves each, one 16-byte load a lane in flight, so INGEST_GRID * 4 KiB of
 * PCIe reads in flight. Enough for the link; an uncapped grid (hundreds of
 * blocks) keeps thousands of waves waiting on PCIe and measured 2-3x slower
 * filterbank and display kernels beside it. */
#ifndef INGEST_GRID
#define INGEST_GRID 64
#endif
/* 0: the ingest runs on the compute stream, in line with the rest */
#ifndef INGEST_STREAM
#define INGEST_STREAM 1
#endif
/* 1: the display lines run on their own HIP stream beside the filterbank;
 * 0: after it on the compute stream */
#ifndef DISPLAY_STREAM
#define DISPLAY_STREAM 1
#endif
#define DISP_BLK_CAP 4096      /* display line segments per batch */
#define FBLK_CAP 8192          /* display frame blocks per batch */
#define PFLOATS_CAP (32u << 20) /* display power spectra per batch, floats (128 MiB) */
#define UBLK_CAP 16384         /* ingest blocks per batch */
#define PBLK_CAP 16384         /* filterbank blocks per batch */
#define QREC_CAP 1088          /* channel records of one detect query (>= 4096 / 4 + 1) */
#define QSET_CAP 256           /* channel records per line of a detect_latest query */
#define DISP_COLS_MAX 4096

typedef struct {
    long long ticket;
    size_t off;                /* staging byte offset (monotonic) */
    size_t nbytes, n;          /* staged bytes, samples per element */
    size_t estride;            /* staging bytes between elements */
    double t_submit;
    int ready;                 /* bytes are in staging: the dispatcher may take it */
} chunk_t;

typedef struct {
    int used, removing;
    atkgpu_stream_cfg cfg;
    unsigned M, os, K, hop, ov, keep, nel;
    int kos, bps;
    size_t m_max, row_len, bytes;
    atkdsp_cf32 *rows[2];      /* each [nel][row_len] */
    int cur;
    float *taps;
    const void *tw;
    atkdsp_cf32 *out;          /* [out_slots][nel][M][m_max] */
    atkdsp_cf32 *R;            /* coherent: [M][nel][nel] */
    uint32_t *cnt;             /* coherent: [M] */
    float *pwr, *mean;         /* coherent: [nel][M], [M] */
    /* display (dN == 0: none) */
    unsigned dN, dhop, dF, dcols, dslots;
    int ddet;
    float *dwin, *dacc, *dring;    /* [dN], [2][nel][dN] (double-buffered running lines), [dslots][nel][dcols] */
    int dpar;                      /* the running-line buffer the next chunk reads */
    long long g_sub, g;            /* samples submitted / dispatched, per element */
    long long fnext;               /* next display frame to transform */
    long long lines_disp, lines_done;
    unsigned lc;
    long long b;
    long long submitted, dispatched, completed;
    int has_prev;              /* a chunk has been dispatched */
    size_t n_prev;             /* samples per element of the last dispatched chunk */
    size_t *slot_m;            /* hops of the result held in each output slot */
    chunk_t q[QCAP];
    unsigned qh, qn;           /* queue head, count */
} sstream;

/* one chunk in a batch, with everything the launch needs, taken from the
 * stream's state when the dispatcher admitted the chunk */
typedef struct {
    int sid, fmt;
    chunk_t c;
    unsigned M, os, kos, ov, keep, lc, nel;
    size_t m, row_len, m_max;
    long long b0;
    atkdsp_cf32 *row, *prev, *out;     /* element 0; element e at + e*row_len (rows), + e*M*m_max (out) */
    int has_prev;                      /* prev holds the stream's previous chunk (else the prefix is zero) */
    size_t n_prev;                     /* its samples per element */
    const float *taps;
    const void *tw;
    float off_re, off_im;
    atkdsp_cf32 *R;
    uint32_t *cnt;
    float *pwr, *mean;
    /* display: frames [fa, fa + nf) complete in this chunk */
    unsigned dN, dhop, dF, dcols, dslots;
    int ddet, nf, dpar;
    long long g0, fa, lines_after;
    const float *dwin;
    float *dacc, *dring;
} bitem;

typedef struct {
    void *event;               /* the batch's compute is done (polled by the completion thread) */
    void *uevent;              /* its ingest is done */
    void *devent;              /* its display lines are done */
    int n;
    bitem *items;              /* max_batch */
    void *desc_h, *desc_d;     /* mapped descriptor buffer */
    float *dP;                 /* display power spectra, PFLOATS_CAP floats */
} batch_t;

typedef struct {
    size_t off, nbytes;
    int released;
} stage_rec;

struct atkgpu_server {
    size_t budget, used_bytes;
    unsigned max_streams, max_batch, jobcap;   /* jobcap: element jobs per batch */
    sstream *st;
    unsigned char *stage_h, *stage_d;
    size_t stage_cap, stage_head, stage_tail;    /* monotonic byte counters */
    stage_rec *recs;
    size_t rec_cap, rec_head, rec_tail;
    void *stream, *ustream, *dstream;            /* compute, ingest, display */
    /* line queries (detect): their own stream, scratch and mapped results */
    pthread_mutex_t qmu;
    void *qstream;
    float *q_scratch;                            /* device, DISP_COLS_MAX floats */
    void *q_h, *q_d;                             /* mapped: records, count, floor */
    void *qs_h, *qs_d;                           /* mapped, per stream: line pointers, records, counts, floors */
    atkgpu_fftplan *plans[NPLANS];               /* twiddles for 64..4096 points */
    batch_t batch[NINFLIGHT];
    void *oev[NOEV];                             /* compute done, for the dispatcher's stream waits only */
    unsigned b_submit, b_done;                   /* batches launched / completed */
    unsigned rr;                                 /* round-robin start */
    pthread_mutex_t mu;
    pthread_cond_t cv_work, cv_launch, cv_done, cv_space;   /* dispatcher, completion, waiters, submitters */
    pthread_t th_disp, th_comp;
    int running, failed;
    uint64_t *lat_hist;                          /* LAT_NB buckets */
    size_t nlat;
    double lat_sum, lat_min, lat_max;
    long long chunks, samples, nbatches;
    double t_first, t_last;
};

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static size_t round16(size_t x) { return (x + 15) & ~(size_t)15; }
static size_t round64(size_t x) { return (x + 63) & ~(size_t)63; }

static int plan_index(unsigned M) {
    for (int i = 0; i < NPLANS; ++i)
        if (M == (64u << i)) return i;
    return -1;
}

/* descriptor buffer layout for one batch, one descriptor per launch block:
 * ingest (prefix copies and unpacks), filterbank, display */
static size_t desc_bytes(void) {
    return (size_t)UBLK_CAP * sizeof(atkgpu_unpack_job) + (size_t)PBLK_CAP * sizeof(atkgpu_pfb_job) +
           (size_t)FBLK_CAP * sizeof(atkgpu_frame_blk) + (size_t)DISP_BLK_CAP * sizeof(atkgpu_line_blk) + 256;
}

static unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

/* display frames complete once a stream has g samples */
static long long disp_frames_at(long long g, unsigned N, unsigned hop) {
    return g >= (long long)N ? (g - (long long)N) / hop + 1 : 0;
}

/* line segments of frames [fa, fa + nf): the runs between line boundaries
 * (multiples of F) */
static int disp_nseg(long long fa, int nf, unsigned F) {
    if (nf <= 0) return 0;
    const long long fe = fa + nf;
    return (int)((fe - 1) / F - fa / F + 1);
}

static void *disp_main(void *arg);
static void *comp_main(void *arg);

/* the detect_latest area: [n] line pointers, [n][QSET_CAP] records, [n] counts, [n] floors */
static size_t qset_bytes(unsigned n) {
    return (size_t)n * (sizeof(void *) + QSET_CAP * sizeof(atkgpu_chan) + sizeof(unsigned long long) + sizeof(float)) +
           64;
}

atkgpu_server *atkgpu_server_create(size_t hbm_budget, unsigned max_streams, unsigned max_batch, size_t staging) {
    if (max_streams < 1 || max_batch < 1 || staging < ((size_t)1 << 20)) return NULL;
    if (atkgpu_init() != 0) return NULL;
    atkgpu_server *v = (atkgpu_server *)calloc(1, sizeof *v);
    if (!v) return NULL;
    v->budget = hbm_budget;
    v->max_streams = max_streams;
    v->max_batch = max_batch;
    v->jobcap = max_batch + 4 * NEL_MAX;        /* room for array streams beside single ones */
    v->st = (sstream *)calloc(max_streams, sizeof(sstream));
    v->stage_cap = staging & ~(size_t)63;
    v->stage_h = (unsigned char *)atkgpu_hmalloc_mapped(staging, (void **)&v->stage_d);
    v->rec_cap = (size_t)max_streams * QCAP + 16;
    v->recs = (stage_rec *)calloc(v->rec_cap, sizeof(stage_rec));
    v->lat_hist = (uint64_t *)calloc(LAT_NB, sizeof(uint64_t));
    int ok = v->st && v->stage_h && v->recs && v->lat_hist && !atkgpu_stream_create(&v->stream) &&
             !atkgpu_stream_create(&v->ustream) && !atkgpu_stream_create(&v->dstream) &&
             !atkgpu_stream_create(&v->qstream) &&
             (v->q_scratch = (float *)atkgpu_dmalloc(DISP_COLS_MAX * sizeof(float))) &&
             (v->q_h = atkgpu_hmalloc_mapped(QREC_CAP * sizeof(atkgpu_chan) + 64, &v->q_d)) &&
             (v->qs_h = atkgpu_hmalloc_mapped(qset_bytes(max_streams), &v->qs_d));
    for (int i = 0; ok && i < NPLANS; ++i) ok = (v->plans[i] = atkgpu_fftplan_create(64u << i)) != NULL;
    for (int i = 0; ok && i < NOEV; ++i) ok = !atkgpu_event_create_order(&v->oev[i]);
    for (int i = 0; ok && i < NINFLIGHT; ++i) {
        batch_t *b = &v->batch[i];
        ok = !atkgpu_event_create_order(&b->event) &&
             !atkgpu_event_create_order(&b->uevent) &&
             !atkgpu_event_create_order(&b->devent) &&
             (b->items = (bitem *)calloc(max_batch, sizeof(bitem))) &&
             (b->desc_h = atkgpu_hmalloc_mapped(desc_bytes(), &b->desc_d)) &&
             (b->dP = (float *)atkgpu_dmalloc((size_t)PFLOATS_CAP * sizeof(float)));
    }
    if (!ok) { atkgpu_server_destroy(v); return NULL; }
    /* touch every staging page once now: the CPU mapping of pinned memory is
     * made page by page on first access, and left to the submit path that
     * made the first 1 GiB of chunks ~7x slower to stage (74 us instead of
     * 11 us per 128 KiB, measured) */
    memset(v->stage_h, 0, v->stage_cap);
    pthread_mutex_init(&v->mu, NULL);
    pthread_mutex_init(&v->qmu, NULL);
    pthread_cond_init(&v->cv_work, NULL);
    pthread_cond_init(&v->cv_launch, NULL);
    pthread_cond_init(&v->cv_done, NULL);
    pthread_cond_init(&v->cv_space, NULL);
    v->running = 1;
    pthread_create(&v->th_disp, NULL, disp_main, v);
    pthread_create(&v->th_comp, NULL, comp_main, v);
    return v;
}

static void stream_free(sstream *s) {
    atkgpu_dfree(s->rows[0]);
    atkgpu_dfree(s->rows[1]);
    atkgpu_dfree(s->taps);
    atkgpu_dfree(s->out);
    atkgpu_dfree(s->R);
    atkgpu_dfree(s->cnt);
    atkgpu_dfree(s->pwr);
    atkgpu_dfree(s->mean);
    atkgpu_dfree(s->dwin);
    atkgpu_dfree(s->dacc);
    atkgpu_dfree(s->dring);
    free(s->slot_m);
    memset(s, 0, sizeof *s);
}

void atkgpu_server_destroy(atkgpu_server *v) {
    if (!v) return;
    if (v->running) {
        atkgpu_server_drain(v);
        pthread_mutex_lock(&v->mu);
        v->running = 0;
        pthread_cond_broadcast(&v->cv_work);
        pthread_cond_broadcast(&v->cv_launch);
        pthread_cond_broadcast(&v->cv_done);
        pthread_cond_broadcast(&v->cv_space);
        pthread_mutex_unlock(&v->mu);
        pthread_join(v->th_disp, NULL);
        pthread_join(v->th_comp, NULL);
    }
    if (v->ustream) atkgpu_sync(v->ustream);
    if (v->dstream) atkgpu_sync(v->dstream);
    if (v->stream) atkgpu_sync(v->stream);
    for (unsigned i = 0; v->st && i < v->max_streams; ++i)
        if (v->st[i].used) stream_free(&v->st[i]);
    for (int i = 0; i < NOEV; ++i) atkgpu_event_destroy(v->oev[i]);
    for (int i = 0; i < NINFLIGHT; ++i) {
        atkgpu_event_destroy(v->batch[i].event);
        atkgpu_event_destroy(v->batch[i].uevent);
        atkgpu_event_destroy(v->batch[i].devent);
        atkgpu_dfree(v->batch[i].dP);
        free(v->batch[i].items);
        atkgpu_hfree(v->batch[i].desc_h);
    }
    for (int i = 0; i < NPLANS; ++i) atkgpu_fftplan_destroy(v->plans[i]);
    if (v->stream) atkgpu_stream_destroy(v->stream);
    if (v->ustream) atkgpu_stream_destroy(v->ustream);
    if (v->dstream) atkgpu_stream_destroy(v->dstream);
    if (v->qstream) atkgpu_stream_destroy(v->qstream);
    atkgpu_dfree(v->q_scratch);
    atkgpu_hfree(v->q_h);
    atkgpu_hfree(v->qs_h);
    atkgpu_hfree(v->stage_h);
    free(v->recs);
    free(v->lat_hist);
    free(v->st);
    free(v);
}

int atkgpu_server_add_stream(atkgpu_server *v, const atkgpu_stream_cfg *cfg) {
    if (!v || !cfg) return ATKDSP_E_ARG;
    const int bps = atkdsp_bytes_per_sample(cfg->fmt);
    const int kos = atkgpu_pfb_fused_kos(cfg->nchan, cfg->oversample, cfg->taps_per_poly);
    const unsigned nel = cfg->nel ? cfg->nel : 1;
    if (!bps || !kos || cfg->taps_per_poly < 2 || cfg->chunk_max < 1 || cfg->out_slots < 1 || !(cfg->fs > 0) ||
        nel > NEL_MAX || (cfg->coherent && nel < 2))
        return ATKDSP_E_ARG;
    const unsigned M = cfg->nchan, os = cfg->oversample, K = cfg->taps_per_poly;
    const unsigned hop = M / os, ov = M * K - hop;
    /* display: the frame length, hop, line and ring */
    const unsigned dN = cfg->disp_nfft;
    const unsigned dhop = cfg->disp_hop ? cfg->disp_hop : dN;
    const unsigned dcols = cfg->disp_cols ? cfg->disp_cols : dN;
    const unsigned dslots = cfg->disp_slots ? cfg->disp_slots : 4;
    if (dN) {
        if ((dN != 1024 && dN != 2048 && dN != 4096) || dhop > dN || cfg->disp_frames < 1 || dcols > dN ||
            cfg->disp_window < 0 || cfg->disp_window > 3 || cfg->disp_detector < 0 || cfg->disp_detector > 2)
            return ATKDSP_E_ARG;
        /* a chunk's frames, frame blocks and line segments must fit one batch
         * with room to spare */
        const size_t fmax = cfg->chunk_max / dhop + 1;
        if ((size_t)nel * (fmax / cfg->disp_frames + 2) > DISP_BLK_CAP / 2 ||
            (size_t)nel * fmax * dN > PFLOATS_CAP / 2 ||
            (size_t)nel * (fmax / (size_t)atkgpu_disp_teams(dN) + 1) > FBLK_CAP / 2)
            return ATKDSP_E_ARG;
    }
    /* the kept prefix holds the filterbank's history and, with a display,
     * the unfinished frame (< dN samples) */
    const unsigned keep = (ov + hop) > dN ? ov + hop : dN;
    /* a chunk's ingest and filterbank blocks must fit one batch with room to spare */
    if ((size_t)nel * (cdiv(keep, ATKGPU_RAGGED_SPB) + cdiv(cfg->chunk_max, ATKGPU_RAGGED_SPB)) > UBLK_CAP / 2 ||
        (size_t)nel * cdiv((cfg->chunk_max + hop) / hop, (size_t)atkgpu_pfb_group(M, os)) > PBLK_CAP / 2)
        return ATKDSP_E_ARG;
    /* output rows padded to 16 hops (128 bytes): whole sectors per segment */
    const size_t m_max = round16((cfg->chunk_max + hop) / hop + 1);
    /* rows padded to 8 samples (64 bytes): each element's row then starts
     * aligned, whatever chunk_max is (the ingest's word stores need 16) */
    const size_t row_len = ((size_t)keep + cfg->chunk_max + 7) & ~(size_t)7;
    size_t bytes = 2 * (size_t)nel * row_len * sizeof(atkdsp_cf32) + (size_t)kos * hop * sizeof(float) +
                   (size_t)cfg->out_slots * nel * M * m_max * sizeof(atkdsp_cf32);
    if (cfg->coherent)
        bytes += (size_t)M * nel * nel * sizeof(atkdsp_cf32) + (size_t)M * sizeof(uint32_t) +
                 (size_t)(nel + 1) * M * sizeof(float);
    if (dN) bytes += ((size_t)(2 * nel + 1) * dN + (size_t)dslots * nel * dcols) * sizeof(float);
    pthread_mutex_lock(&v->mu);
    int sid = -1;
    for (unsigned i = 0; i < v->max_streams; ++i)
        if (!v->st[i].used) { sid = (int)i; break; }
    if (sid < 0 || v->used_bytes + bytes > v->budget) {
        pthread_mutex_unlock(&v->mu);
        return ATKDSP_E_NOMEM;                      /* refused: no slot or over the HBM budget */
    }
    sstream *s = &v->st[sid];
    memset(s, 0, sizeof *s);
    s->used = 1;                                    /* reserve the slot while allocating */
    v->used_bytes += bytes;
    pthread_mutex_unlock(&v->mu);

    s->cfg = *cfg;
    s->cfg.nel = nel;
    s->M = M; s->os = os; s->K = K; s->hop = hop; s->ov = ov; s->keep = keep; s->nel = nel;
    s->kos = kos; s->bps = bps; s->m_max = m_max; s->row_len = row_len; s->bytes = bytes;
    s->tw = atkgpu_fftplan_twiddles(v->plans[plan_index(M)]);
    s->rows[0] = (atkdsp_cf32 *)atkgpu_dmalloc((size_t)nel * row_len * sizeof(atkdsp_cf32));
    s->rows[1] = (atkdsp_cf32 *)atkgpu_dmalloc((size_t)nel * row_len * sizeof(atkdsp_cf32));
    s->taps = (float *)atkgpu_dmalloc((size_t)kos * hop * sizeof(float));
    s->out = (atkdsp_cf32 *)atkgpu_dmalloc((size_t)cfg->out_slots * nel * M * m_max * sizeof(atkdsp_cf32));
    s->slot_m = (size_t *)calloc(cfg->out_slots, sizeof(size_t));
    int rc = (!s->rows[0] || !s->rows[1] || !s->taps || !s->out || !s->slot_m) ? ATKDSP_E_NOMEM : 0;
    if (!rc && cfg->coherent) {
        s->R = (atkdsp_cf32 *)atkgpu_dmalloc((size_t)M * nel * nel * sizeof(atkdsp_cf32));
        s->cnt = (uint32_t *)atkgpu_dmalloc((size_t)M * sizeof(uint32_t));
        s->pwr = (float *)atkgpu_dmalloc((size_t)nel * M * sizeof(float));
        s->mean = (float *)atkgpu_dmalloc((size_t)M * sizeof(float));
        if (!s->R || !s->cnt || !s->pwr || !s->mean) rc = ATKDSP_E_NOMEM;
    }
    float *w = NULL;
    if (!rc && dN) {
        s->dN = dN; s->dhop = dhop; s->dF = cfg->disp_frames; s->dcols = dcols; s->dslots = dslots;
        s->ddet = cfg->disp_detector;
        s->dwin = (float *)atkgpu_dmalloc((size_t)dN * sizeof(float));
        s->dacc = (float *)atkgpu_dmalloc((size_t)2 * nel * dN * sizeof(float));
        s->dring = (float *)atkgpu_dmalloc((size_t)dslots * nel * dcols * sizeof(float));
        w = (float *)malloc((size_t)dN * sizeof(float));
        if (!s->dwin || !s->dacc || !s->dring || !w) rc = ATKDSP_E_NOMEM;
        if (!rc) rc = atkdsp_window(cfg->disp_window, dN, w);
        if (!rc) rc = atkgpu_h2d(v->stream, s->dwin, w, (size_t)dN * sizeof(float));
    }
    double *h = (double *)malloc((size_t)M * K * sizeof(double));
    float *hf = (float *)calloc((size_t)kos * hop, sizeof(float));
    if (!rc && (!h || !hf)) rc = ATKDSP_E_NOMEM;
    if (!rc) rc = atkgpu_pfb_prototype(h, M, K, cfg->atten_db);
    if (!rc) {
        for (size_t q = 0; q < (size_t)M * K; ++q) hf[q] = (float)h[q];
        /* history rows start at zero (the reference callocs its history) */
        rc = atkgpu_h2d(v->stream, s->taps, hf, (size_t)kos * hop * sizeof(float));
        if (!rc) rc = atkgpu_fill_u32(v->stream, s->rows[0], 0, (size_t)nel * row_len * 2);
        if (!rc && cfg->coherent) rc = atkgpu_pfb_cov_reset(v->stream, s->R, s->cnt, nel, M, NULL);
        if (!rc) rc = atkgpu_sync(v->stream);
    }
    free(h);
    free(hf);
    free(w);
    if (rc) {
        pthread_mutex_lock(&v->mu);
        v->used_bytes -= bytes;
        stream_free(s);
        pthread_mutex_unlock(&v->mu);
        return rc < 0 ? rc : ATKDSP_E_NOMEM;
    }
    return sid;
}

int atkgpu_server_remove_stream(atkgpu_server *v, int sid) {
    if (!v || sid < 0 || (unsigned)sid >= v->max_streams) return ATKDSP_E_ARG;
    pthread_mutex_lock(&v->mu);
    sstream *s = &v->st[sid];
    if (!s->used) { pthread_mutex_unlock(&v->mu); return ATKDSP_E_ARG; }
    s->removing = 1;
    while (s->completed < s->submitted && !v->failed) pthread_cond_wait(&v->cv_done, &v->mu);
    v->used_bytes -= s->bytes;
    stream_free(s);
    pthread_mutex_unlock(&v->mu);
    return ATKDSP_OK;
}

/* staging space for nbytes at the ring's head (contiguous; a wrap skips the tail end) */
static int stage_alloc(atkgpu_server *v, size_t nbytes, size_t *off) {
    /* every chunk starts 64-byte aligned, so the device's vector loads of any
     * format are naturally aligned */
    nbytes = round64(nbytes);
    size_t head = round64(v->stage_head);
    const size_t pos = head % v->stage_cap;
    if (pos + nbytes > v->stage_cap) head += v->stage_cap - pos;       /* skip to the ring start */
    if (head + nbytes - v->stage_tail > v->stage_cap) return 0;
    if (v->rec_head - v->rec_tail >= v->rec_cap) return 0;
    *off = head;
    v->stage_head = head + nbytes;
    stage_rec *r = &v->recs[v->rec_head % v->rec_cap];
    r->off = head;
    r->nbytes = nbytes;
    r->released = 0;
    v->rec_head++;
    return 1;
}

static void stage_release(atkgpu_server *v, size_t off) {
    for (size_t i = v->rec_tail; i < v->rec_head; ++i) {
        stage_rec *r = &v->recs[i % v->rec_cap];
        if (r->off == off && !r->released) { r->released = 1; break; }
    }
    while (v->rec_tail < v->rec_head && v->recs[v->rec_tail % v->rec_cap].released) {
        const stage_rec *r = &v->recs[v->rec_tail % v->rec_cap];
        v->stage_tail = r->off + r->nbytes;
        v->rec_tail++;
    }
    if (v->rec_tail == v->rec_head) v->stage_tail = v->stage_head;
}

long long atkgpu_server_submit(atkgpu_server *v, int sid, const void *raw, size_t nbytes) {
    if (!v || !raw || sid < 0 || (unsigned)sid >= v->max_streams) return ATKDSP_E_ARG;
    pthread_mutex_lock(&v->mu);
    sstream *s = &v->st[sid];
    if (!s->used || s->removing || v->failed) { pthread_mutex_unlock(&v->mu); return v->failed ? ATKDSP_E_FFT : ATKDSP_E_ARG; }
    /* an array chunk: nel element blocks of n samples, element-major */
    const size_t n = nbytes / ((size_t)s->bps * s->nel);
    const size_t eb = n * (size_t)s->bps;             /* bytes of one element */
    const size_t estride = s->nel > 1 ? round64(eb) : eb;
    const size_t nb = estride * (s->nel - 1) + eb;
    if (n > s->cfg.chunk_max || nb + 64 > v->stage_cap) {
        pthread_mutex_unlock(&v->mu);
        return ATKDSP_E_ARG;
    }
    size_t off = 0;
    /* backpressure: wait for a queue entry and staging space */
    while (v->running && !v->failed && (s->qn >= QCAP || !stage_alloc(v, nb ? nb : 1, &off)))
        pthread_cond_wait(&v->cv_space, &v->mu);
    if (!v->running || v->failed) { pthread_mutex_unlock(&v->mu); return ATKDSP_E_FFT; }
    const double t = now_s();
    const long long ticket = s->submitted++;
    s->g_sub += (long long)n;
    chunk_t *c = &s->q[(s->qh + s->qn) % QCAP];
    c->ticket = ticket;
    c->off = off;
    c->nbytes = nb;
    c->n = n;
    c->estride = estride;
    c->t_submit = t;
    c->ready = 0;              /* queued in ticket order, not yet takeable */
    s->qn++;
    if (!v->chunks && !v->nlat) v->t_first = t;
    const unsigned nel = s->nel;
    pthread_mutex_unlock(&v->mu);
    /* the copy runs unlocked; the queue entry and staging span are this
     * call's alone until it is marked ready */
    unsigned char *dst = v->stage_h + (off % v->stage_cap);
    for (unsigned e = 0; e < nel && eb; ++e) memcpy(dst + e * estride, (const unsigned char *)raw + e * eb, eb);
    pthread_mutex_lock(&v->mu);
    c->ready = 1;
    pthread_cond_signal(&v->cv_work);
    pthread_mutex_unlock(&v->mu);
    return ticket;
}

/* One block of a ring consumer into the stream, copied once: from the ring
 * slot (a zero-copy peek) straight into pinned staging. The chunk is queued
 * only after the release confirms the producer did not rewrite the slot
 * during the copy; a block caught being rewritten gives its staging back,
 * counts as lost, and the next block is taken instead, as atkdsp_ring_read
 * does. A block of another format is not consumed. */
long long atkgpu_server_submit_ring(atkgpu_server *v, int sid, atkdsp_ring *r, int consumer, atkdsp_ring_meta *meta,
                                    uint64_t *lost) {
    uint64_t nlost = 0;
    if (lost) *lost = 0;
    if (!v || !r || sid < 0 || (unsigned)sid >= v->max_streams) return ATKDSP_E_ARG;
    const double t = now_s();
    for (;;) {
        atkdsp_ring_meta m;
        uint64_t ln = 0;
        const unsigned char *p = (const unsigned char *)atkdsp_ring_peek(r, consumer, &m, &ln);
        nlost += ln;
        if (lost) *lost = nlost;
        if (!p) {
            atkdsp_ring_counters rc;
            if (atkdsp_ring_stats(r, &rc) != ATKDSP_OK || consumer < 0 || consumer >= ATKDSP_RING_MAX_CONSUMERS ||
                !((rc.active_mask >> consumer) & 1u))
                return ATKDSP_E_ARG;
            return ATKGPU_RING_EMPTY;
        }
        pthread_mutex_lock(&v->mu);
        sstream *s = &v->st[sid];
        if (!s->used || s->removing || v->failed) {
            pthread_mutex_unlock(&v->mu);
            return v->failed ? ATKDSP_E_FFT : ATKDSP_E_ARG;
        }
        const size_t n = m.nbytes / ((size_t)s->bps * s->nel);
        const size_t eb = n * (size_t)s->bps;
        const size_t estride = s->nel > 1 ? round64(eb) : eb;
        const size_t nb = estride * (s->nel - 1) + eb;
        if (m.format != s->cfg.fmt || n > s->cfg.chunk_max || nb + 64 > v->stage_cap) {
            pthread_mutex_unlock(&v->mu);
            return ATKDSP_E_ARG;
        }
        size_t off = 0;
        while (v->running && !v->failed && !stage_alloc(v, nb ? nb : 1, &off)) pthread_cond_wait(&v->cv_space, &v->mu);
        if (!v->running || v->failed) { pthread_mutex_unlock(&v->mu); return ATKDSP_E_FFT; }
        const unsigned nel = s->nel;
        pthread_mutex_unlock(&v->mu);
        unsigned char *dst = v->stage_h + (off % v->stage_cap);
        for (unsigned e = 0; e < nel && eb; ++e) memcpy(dst + e * estride, p + e * eb, eb);
        const int intact = atkdsp_ring_release(r, consumer) == ATKDSP_OK;
        pthread_mutex_lock(&v->mu);
        if (!intact) {
            stage_release(v, off);
            pthread_cond_broadcast(&v->cv_space);
            pthread_mutex_unlock(&v->mu);
            if (lost) *lost = ++nlost;
            continue;
        }
        while (v->running && !v->failed && s->used && !s->removing && s->qn >= QCAP)
            pthread_cond_wait(&v->cv_space, &v->mu);
        if (!v->running || v->failed || !s->used || s->removing) {
            stage_release(v, off);
            pthread_cond_broadcast(&v->cv_space);
            const int failed = !v->running || v->failed;
            pthread_mutex_unlock(&v->mu);
            return failed ? ATKDSP_E_FFT : ATKDSP_E_ARG;
        }
        const long long ticket = s->submitted++;
        s->g_sub += (long long)n;
        chunk_t *c = &s->q[(s->qh + s->qn) % QCAP];
        c->ticket = ticket;
        c->off = off;
        c->nbytes = nb;
        c->n = n;
        c->estride = estride;
        c->t_submit = t;
        c->ready = 1;
        s->qn++;
        if (!v->chunks && !v->nlat) v->t_first = t;
        pthread_cond_signal(&v->cv_work);
        pthread_mutex_unlock(&v->mu);
        if (meta) *meta = m;
        return ticket;
    }
}

/* Launch batch k (the mutex is not held): on the ingest stream, after batch
 * k-2's compute, every element's kept prefix and unpack; after that ingest,
 * the fused filterbank per (M, os, kos) class and W1/W2 per coherent array on
 * the compute stream, and the display lines on the display stream, which the
 * compute stream then waits for. */
static int launch_batch(atkgpu_server *v, batch_t *bt, unsigned k) {
    const int n = bt->n;
    atkgpu_unpack_job *uj = (atkgpu_unpack_job *)bt->desc_h;
    atkgpu_pfb_job *pj = (atkgpu_pfb_job *)(uj + UBLK_CAP);
    char *dbase = (char *)bt->desc_d, *hbase = (char *)bt->desc_h;
#define DEV(p) ((void *)(dbase + ((char *)(p) - hbase)))
    int rc = 0;
    /* ingest, one launch: the kept prefix of each element (the last keep
     * samples of [prev, prev + keep + n_prev)) into [row, row + keep), and the
     * chunk's samples unpacked into [row + keep, row + keep + n), every format
     * together. Staging offsets are 64-byte aligned and keep is a multiple of
     * 64, so both ends of an unpack are 16-byte aligned for its word loads and
     * stores. */
    /* one descriptor per launch block (blk0 = the block's tile in its job) */
    int nu = 0;
    for (int i = 0; i < n; ++i) {
        const bitem *it = &bt->items[i];
        for (unsigned e = 0; e < it->nel; ++e) {
            atkgpu_unpack_job j;
            if (it->has_prev) {
                j.raw = it->prev + e * it->row_len + it->n_prev;
                j.dst = it->row + e * it->row_len;
                j.off_re = j.off_im = 0.0f;
                j.n = (int)it->keep;
                j.nblk = (j.n + ATKGPU_RAGGED_SPB - 1) / ATKGPU_RAGGED_SPB;
                j.fmt = ATKGPU_JOB_COPY;
                for (int t = 0; t < j.nblk; ++t) { j.blk0 = t; uj[nu++] = j; }
            }
            if (it->c.n == 0) continue;
            j.raw = v->stage_d + (it->c.off % v->stage_cap) + e * it->c.estride;
            j.dst = it->row + e * it->row_len + it->keep;
            j.off_re = it->off_re;
            j.off_im = it->off_im;
            j.n = (int)it->c.n;
            j.nblk = (j.n + ATKGPU_RAGGED_SPB - 1) / ATKGPU_RAGGED_SPB;
            j.fmt = it->fmt;
            if ((((size_t)j.raw) | ((size_t)j.dst)) & 15) rc = ATKDSP_E_ARG;
            for (int t = 0; t < j.nblk; ++t) { j.blk0 = t; uj[nu++] = j; }
        }
    }
    /* batch k's ingest overwrites the rows batch k-2 read. The wait uses
     * that batch's ordering event, which only this thread records and waits
     * on; the completion thread polls the other one. */
    void *ust = INGEST_STREAM ? v->ustream : v->stream;
    if (!rc && INGEST_STREAM && k >= 2)
        rc = atkgpu_stream_wait_event(v->ustream, v->oev[(k - 2) % NOEV]);
    if (!rc && nu) rc = atkgpu_unpack_jobs(ust, DEV(uj), nu, nu, 1, INGEST_GRID);
    if (!rc) rc = atkgpu_event_record(bt->uevent, ust);
    if (!rc && INGEST_STREAM) rc = atkgpu_stream_wait_event(v->stream, bt->uevent);
    void *dst = DISPLAY_STREAM ? v->dstream : v->stream;
    if (!rc && DISPLAY_STREAM) rc = atkgpu_stream_wait_event(v->dstream, bt->uevent);
    /* the filterbank, one launch per shape class present, one descriptor per
     * block (blk0 = its group in the element, nblk = the element's groups) */
    int np = 0;
    unsigned char done[4096];
    memset(done, 0, sizeof done);
    for (int i = 0; i < n && !rc; ++i) {
        if (done[i]) continue;
        const bitem *a = &bt->items[i];
        const int p0 = np;
        const int G = atkgpu_pfb_group(a->M, a->os);
        for (int k = i; k < n; ++k) {
            const bitem *it = &bt->items[k];
            if (done[k] || it->M != a->M || it->os != a->os || it->kos != a->kos) continue;
            done[k] = 1;
            if (!it->m) continue;
            for (unsigned e = 0; e < it->nel; ++e) {
                atkgpu_pfb_job j;
                j.s = it->row + e * it->row_len + (it->keep - it->ov - it->lc);
                j.out = it->out + (size_t)e * it->M * it->m_max;
                j.taps = it->taps;
                j.b0 = it->b0;
                j.nvalid = (int)(it->ov + it->lc + it->c.n);
                j.m = (int)it->m;
                j.opitch = (int)it->m_max;
                j.nblk = (int)((it->m + G - 1) / G);
                j.pad_ = 0;
                for (int t = 0; t < j.nblk; ++t) { j.blk0 = t; pj[np++] = j; }
            }
        }
        if (np > p0) rc = atkgpu_pfb_fused_jobs(v->stream, DEV(&pj[p0]), np - p0, np - p0, a->M, a->os, a->kos, a->tw);
    }
    /* W1 accumulated and W2 of this chunk, per coherent array: the element
     * outputs sit [nel][M][m_max] in the slot */
    for (int i = 0; i < n && !rc; ++i) {
        const bitem *it = &bt->items[i];
        if (!it->R || !it->m) continue;
        rc = atkgpu_pfb_cov(v->stream, it->out, it->nel, it->M, it->m, it->m_max, NULL, it->R, it->cnt);
        if (!rc) rc = atkgpu_pfb_power(v->stream, it->out, it->nel, it->M, it->m, it->m_max, it->pwr, it->mean);
    }
    /* display: per (FFT length, detector) class present, the frames launch
     * (each element's frames in order, spectra contiguous) and the lines
     * launch (one block per line segment; frame f of element e at row_e +
     * keep + f*hop - g0; running lines read from buffer dpar, written to the
     * other) */
    atkgpu_frame_blk *fb = (atkgpu_frame_blk *)(pj + PBLK_CAP);
    atkgpu_line_blk *lb = (atkgpu_line_blk *)(fb + FBLK_CAP);
    int nfb = 0, nb = 0;
    size_t pofs = 0;
    for (unsigned N = 1024; N <= 4096 && !rc; N *= 2) {
        const int TEAMS = atkgpu_disp_teams(N);
        for (int det = 0; det < 3 && !rc; ++det) {
            const int f0b = nfb, l0b = nb;
            for (int i = 0; i < n; ++i) {
                const bitem *it = &bt->items[i];
                if (!it->nf || it->dN != N || it->ddet != det) continue;
                const long long F = it->dF;
                for (unsigned e = 0; e < it->nel; ++e) {
                    const atkdsp_cf32 *row = it->row + e * it->row_len;
                    float *Pe = bt->dP + pofs;
                    pofs += (size_t)it->nf * N;
                    for (int t = 0; t < it->nf; t += TEAMS) {
                        atkgpu_frame_blk *d = &fb[nfb++];
                        d->first = row + (it->keep + (it->fa + t) * it->dhop - it->g0);
                        d->win = it->dwin;
                        d->P = Pe + (size_t)t * N;
                        d->nfr = it->nf - t < TEAMS ? it->nf - t : TEAMS;
                        d->fhop = (int)it->dhop;
                    }
                    const float *acc_in = it->dacc + ((size_t)it->dpar * it->nel + e) * N;
                    float *acc_out = it->dacc + ((size_t)(it->dpar ^ 1) * it->nel + e) * N;
                    for (long long f = it->fa, fe = it->fa + it->nf; f < fe;) {
                        long long se = (f / F + 1) * F;
                        if (se > fe) se = fe;
                        const int ends = se % F == 0;
                        atkgpu_line_blk *d = &lb[nb++];
                        d->P = Pe + (size_t)(f - it->fa) * N;
                        d->acc_in = (f % F) ? acc_in : NULL;
                        d->acc_out = ends ? NULL : acc_out;
                        d->line = ends ? it->dring + ((size_t)((se / F - 1) % it->dslots) * it->nel + e) * it->dcols
                                       : NULL;
                        d->nfr = (int)(se - f);
                        d->cols = (int)it->dcols;
                        d->scale = det == 2 ? 1.0f / (float)F : 1.0f;
                        d->pad_ = 0;
                        f = se;
                    }
                }
            }
            if (nfb > f0b)
                rc = atkgpu_disp_frames(dst, DEV(&fb[f0b]), nfb - f0b, N,
                                        atkgpu_fftplan_twiddles(v->plans[plan_index(N)]));
            if (!rc && nb > l0b) rc = atkgpu_disp_lines(dst, DEV(&lb[l0b]), nb - l0b, N, det);
        }
    }
#undef DEV
    if (!rc && nb && DISPLAY_STREAM) rc = atkgpu_event_record(bt->devent, v->dstream);
    if (!rc && nb && DISPLAY_STREAM) rc = atkgpu_stream_wait_event(v->stream, bt->devent);
    if (!rc) rc = atkgpu_event_record(v->oev[k % NOEV], v->stream);
    if (!rc) rc = atkgpu_event_record(bt->event, v->stream);
    return rc;
}

/* take up to max_batch chunks (and jobcap element jobs), one per stream,
 * round-robin; mutex held */
static int form_batch(atkgpu_server *v, batch_t *bt) {
    int n = 0;
    unsigned jobs = 0, dblk = 0, ublk = 0, pblk = 0, fblk = 0;
    size_t pfl = 0;
    const unsigned S = v->max_streams;
    for (unsigned t = 0; t < S && n < (int)v->max_batch && n < 4096; ++t) {
        const unsigned sid = (v->rr + t) % S;
        sstream *s = &v->st[sid];
        if (!s->used || !s->qn || !s->q[s->qh].ready) continue;
        if (jobs + s->nel > v->jobcap) continue;
        /* display frames this chunk completes, and their blocks */
        int nf = 0;
        unsigned nb = 0, nfb = 0;
        size_t npf = 0;
        if (s->dN) {
            const long long fe = disp_frames_at(s->g + (long long)s->q[s->qh].n, s->dN, s->dhop);
            nf = (int)(fe - s->fnext);
            nb = s->nel * (unsigned)disp_nseg(s->fnext, nf, s->dF);
            nfb = s->nel * cdiv((size_t)nf, (size_t)atkgpu_disp_teams(s->dN));
            npf = (size_t)s->nel * nf * s->dN;
            if (dblk + nb > DISP_BLK_CAP || fblk + nfb > FBLK_CAP || pfl + npf > PFLOATS_CAP) continue;
        }
        /* its ingest and filterbank blocks */
        const size_t cn = s->q[s->qh].n;
        const unsigned nu = s->nel * ((s->has_prev ? cdiv(s->keep, ATKGPU_RAGGED_SPB) : 0u) + cdiv(cn, ATKGPU_RAGGED_SPB));
        const unsigned np = s->nel * cdiv(((size_t)s->lc + cn) / s->hop, (size_t)atkgpu_pfb_group(s->M, s->os));
        if (ublk + nu > UBLK_CAP || pblk + np > PBLK_CAP) continue;
        ublk += nu;
        pblk += np;
        dblk += nb;
        fblk += nfb;
        pfl += npf;
        const chunk_t c = s->q[s->qh];
        s->qh = (s->qh + 1) % QCAP;
        s->qn--;
        jobs += s->nel;
        bitem *it = &bt->items[n++];
        it->sid = (int)sid;
        it->fmt = s->cfg.fmt;
        it->c = c;
        it->M = s->M;
        it->os = s->os;
        it->kos = (unsigned)s->kos;
        it->ov = s->ov;
        it->keep = s->keep;
        it->lc = s->lc;
        it->nel = s->nel;
        it->row_len = s->row_len;
        it->m_max = s->m_max;
        it->b0 = s->b;
        it->m = ((size_t)s->lc + c.n) / s->hop;
        it->row = s->rows[s->cur];
        it->prev = s->rows[s->cur ^ 1];
        it->has_prev = s->has_prev;
        it->n_prev = s->n_prev;
        const unsigned slot = (unsigned)(c.ticket % s->cfg.out_slots);
        it->out = s->out + (size_t)slot * s->nel * s->M * s->m_max;
        it->taps = s->taps;
        it->tw = s->tw;
        it->off_re = s->cfg.off_re;
        it->off_im = s->cfg.off_im;
        it->R = s->R;
        it->cnt = s->cnt;
        it->pwr = s->pwr;
        it->mean = s->mean;
        it->dN = s->dN;
        it->nf = nf;
        if (s->dN) {
            it->dhop = s->dhop; it->dF = s->dF; it->dcols = s->dcols; it->dslots = s->dslots;
            it->ddet = s->ddet;
            it->g0 = s->g;
            it->fa = s->fnext;
            it->dwin = s->dwin;
            it->dacc = s->dacc;
            it->dring = s->dring;
            it->dpar = s->dpar;
            if (nf) s->dpar ^= 1;          /* the chunk leaves its running line in the other buffer */
            s->fnext += nf;
            s->lines_disp = s->fnext / s->dF;
            it->lines_after = s->lines_disp;
        }
        s->g += (long long)c.n;
        /* advance the stream */
        s->slot_m[slot] = it->m;
        s->lc = (unsigned)(((size_t)s->lc + c.n) - it->m * s->hop);
        s->b += (long long)it->m;
        s->cur ^= 1;
        s->has_prev = 1;
        s->n_prev = c.n;
        s->dispatched++;
    }
    v->rr = (v->rr + 1) % S;
    bt->n = n;
    return n;
}

static void *disp_main(void *arg) {
    atkgpu_server *v = (atkgpu_server *)arg;
    pthread_mutex_lock(&v->mu);
    for (;;) {
        /* a free batch slot (at most NINFLIGHT batches in flight) */
        while (v->running && v->b_submit - v->b_done >= NINFLIGHT) pthread_cond_wait(&v->cv_done, &v->mu);
        if (!v->running) break;
        batch_t *bt = &v->batch[v->b_submit % NINFLIGHT];
        while (v->running && !form_batch(v, bt)) pthread_cond_wait(&v->cv_work, &v->mu);
        if (!v->running) break;
        pthread_mutex_unlock(&v->mu);
        const int rc = launch_batch(v, bt, v->b_submit);
        pthread_mutex_lock(&v->mu);
        if (rc) {
            v->failed = 1;
            pthread_cond_broadcast(&v->cv_done);
            pthread_cond_broadcast(&v->cv_launch);
            pthread_cond_broadcast(&v->cv_space);
            break;
        }
        v->b_submit++;
        v->nbatches++;
        pthread_cond_signal(&v->cv_launch);        /* wake the completion thread */
    }
    pthread_mutex_unlock(&v->mu);
    return NULL;
}

/* wait for an event by polling: a blocking wait either spins a core (which
 * the submitters and the dispatcher need) or sleeps on an interrupt */
static int poll_event(void *ev) {
    const struct timespec ts = {0, POLL_NS};
    for (;;) {
        const int q = atkgpu_event_query(ev);
        if (q) return q < 0 ? q : 0;
        nanosleep(&ts, NULL);
    }
}

static int lat_bucket(double ms) {
    if (!(ms > LAT_LO_MS)) return 0;
    const double b = log(ms / LAT_LO_MS) / LAT_LOG_STEP + 1.0;
    return b < (double)(LAT_NB - 1) ? (int)b : LAT_NB - 1;
}

/* the middle of bucket b (geometric), ms */
static double lat_value(int b) { return b == 0 ? LAT_LO_MS : LAT_LO_MS * exp(((double)b - 0.5) * LAT_LOG_STEP); }

static void lat_add(atkgpu_server *v, double ms) {
    v->lat_hist[lat_bucket(ms)]++;
    if (!v->nlat || ms < v->lat_min) v->lat_min = ms;
    if (!v->nlat || ms > v->lat_max) v->lat_max = ms;
    v->lat_sum += ms;
    v->nlat++;
}

/* the latency at sorted position q (0 .. n-1) of the recorded ones, within
 * the histogram's resolution and never outside [min, max] */
static double lat_at(const atkgpu_server *v, size_t q) {
    uint64_t seen = 0;
    for (int b = 0; b < LAT_NB; ++b) {
        seen += v->lat_hist[b];
        if (seen > q) {
            const double x = lat_value(b);
            return x < v->lat_min ? v->lat_min : x > v->lat_max ? v->lat_max : x;
        }
    }
    return v->lat_max;
}

static void *comp_main(void *arg) {
    atkgpu_server *v = (atkgpu_server *)arg;
    prctl(PR_SET_TIMERSLACK, 1000UL, 0, 0, 0);      /* wake within ~1 us of the poll interval */
    pthread_mutex_lock(&v->mu);
    for (;;) {
        while (v->running && v->b_done == v->b_submit && !v->failed) pthread_cond_wait(&v->cv_launch, &v->mu);
        if (v->b_done == v->b_submit && (!v->running || v->failed)) break;
        batch_t *bt = &v->batch[v->b_done % NINFLIGHT];
        pthread_mutex_unlock(&v->mu);
        const int rc = poll_event(bt->event);
        const double t = now_s();
        pthread_mutex_lock(&v->mu);
        if (rc) { v->failed = 1; pthread_cond_broadcast(&v->cv_done); pthread_cond_broadcast(&v->cv_space); break; }
        for (int i = 0; i < bt->n; ++i) {
            const bitem *it = &bt->items[i];
            sstream *s = &v->st[it->sid];
            s->completed++;
            if (it->dN && it->lines_after > s->lines_done) s->lines_done = it->lines_after;
            lat_add(v, (t - it->c.t_submit) * 1e3);
            v->chunks++;
            v->samples += (long long)it->c.n * it->nel;
            stage_release(v, it->c.off);
        }
        v->t_last = t;
        v->b_done++;
        pthread_cond_broadcast(&v->cv_done);
        pthread_cond_broadcast(&v->cv_space);
    }
    pthread_mutex_unlock(&v->mu);
    return NULL;
}

int atkgpu_server_wait(atkgpu_server *v, int sid, long long ticket) {
    if (!v || sid < 0 || (unsigned)sid >= v->max_streams) return ATKDSP_E_ARG;
    pthread_mutex_lock(&v->mu);
    sstream *s = &v->st[sid];
    if (!s->used || ticket < 0 || ticket >= s->submitted) { pthread_mutex_unlock(&v->mu); return ATKDSP_E_ARG; }
    while (s->completed <= ticket && !v->failed) pthread_cond_wait(&v->cv_done, &v->mu);
    const int rc = v->failed ? ATKDSP_E_FFT : ATKDSP_OK;
    pthread_mutex_unlock(&v->mu);
    return rc;
}

ptrdiff_t atkgpu_server_read(atkgpu_server *v, int sid, long long ticket, atkdsp_cf32 *out, size_t cap) {
    if (!v || !out || sid < 0 || (unsigned)sid >= v->max_streams) return ATKDSP_E_ARG;
    int rc = atkgpu_server_wait(v, sid, ticket);
    if (rc) return rc;
    pthread_mutex_lock(&v->mu);
    sstream *s = &v->st[sid];
    /* the slot still holds this ticket's result until out_slots later chunks run */
    if (ticket + (long long)s->cfg.out_slots < s->dispatched) { pthread_mutex_unlock(&v->mu); return ATKDSP_E_ARG; }
    const unsigned slot = (unsigned)(ticket % s->cfg.out_slots);
    const size_t m = s->slot_m[slot], m_max = s->m_max;
    const unsigned M = s->M, nel = s->nel;
    const atkdsp_cf32 *src = s->out + (size_t)slot * nel * M * m_max;
    pthread_mutex_unlock(&v->mu);
    if (cap < (size_t)nel * M * m) return ATKDSP_E_CAP;
    /* [nel][M] rows of m hops, from device rows of m_max */
    rc = atkgpu_d2h_2d_sync(out, m * sizeof(atkdsp_cf32), src, m_max * sizeof(atkdsp_cf32), m * sizeof(atkdsp_cf32),
                            (size_t)nel * M);
    if (rc) return ATKDSP_E_FFT;
    /* the copy ran beside the compute stream: if a chunk that reuses the slot
     * was dispatched meanwhile, what was copied may be torn */
    pthread_mutex_lock(&v->mu);
    const int overrun = ticket + (long long)s->cfg.out_slots < s->dispatched;
    pthread_mutex_unlock(&v->mu);
    return overrun ? ATKDSP_E_OVERRUN : (ptrdiff_t)m;
}

int atkgpu_server_read_cov(atkgpu_server *v, int sid, atkdsp_cf32 *R, uint32_t *counts) {
    if (!v || sid < 0 || (unsigned)sid >= v->max_streams) return ATKDSP_E_ARG;
    pthread_mutex_lock(&v->mu);
    sstream *s = &v->st[sid];
    if (!s->used || !s->R) { pthread_mutex_unlock(&v->mu); return ATKDSP_E_ARG; }
    while (s->completed < s->submitted && !v->failed) pthread_cond_wait(&v->cv_done, &v->mu);
    const size_t nR = (size_t)s->M * s->nel * s->nel, M = s->M;
    atkdsp_cf32 *dR = s->R;
    uint32_t *dc = s->cnt;
    const int failed = v->failed;
    pthread_mutex_unlock(&v->mu);
    if (failed) return ATKDSP_E_FFT;
    int rc = R ? atkgpu_d2h_sync(R, dR, nR * sizeof(atkdsp_cf32)) : 0;
    if (!rc && counts) rc = atkgpu_d2h_sync(counts, dc, M * sizeof(uint32_t));
    return rc ? ATKDSP_E_FFT : ATKDSP_OK;
}

int atkgpu_server_read_power(atkgpu_server *v, int sid, float *pwr, float *mean_pwr) {
    if (!v || sid < 0 || (unsigned)sid >= v->max_streams) return ATKDSP_E_ARG;
    pthread_mutex_lock(&v->mu);
    sstream *s = &v->st[sid];
    if (!s->used || !s->pwr || !s->completed) { pthread_mutex_unlock(&v->mu); return ATKDSP_E_ARG; }
    const size_t M = s->M, nel = s->nel;
    float *dp = s->pwr, *dm = s->mean;
    pthread_mutex_unlock(&v->mu);
    int rc = pwr ? atkgpu_d2h_sync(pwr, dp, nel * M * sizeof(float)) : 0;
    if (!rc && mean_pwr) rc = atkgpu_d2h_sync(mean_pwr, dm, M * sizeof(float));
    return rc ? ATKDSP_E_FFT : ATKDSP_OK;
}

long long atkgpu_server_lines(atkgpu_server *v, int sid) {
    if (!v || sid < 0 || (unsigned)sid >= v->max_streams) return ATKDSP_E_ARG;
    pthread_mutex_lock(&v->mu);
    const sstream *s = &v->st[sid];
    const long long r = (!s->used || !s->dN) ? ATKDSP_E_ARG : s->lines_done;
    pthread_mutex_unlock(&v->mu);
    return r;
}

int atkgpu_server_read_line(atkgpu_server *v, int sid, long long line, float *out, size_t cap) {
    if (!v || !out || sid < 0 || (unsigned)sid >= v->max_streams || line < 0) return ATKDSP_E_ARG;
    pthread_mutex_lock(&v->mu);
    sstream *s = &v->st[sid];
    /* the line must be completed by the chunks submitted so far */
    if (!s->used || !s->dN || line >= disp_frames_at(s->g_sub, s->dN, s->dhop) / s->dF) {
        pthread_mutex_unlock(&v->mu);
        return ATKDSP_E_ARG;
    }
    while (s->lines_done <= line && !v->failed) pthread_cond_wait(&v->cv_done, &v->mu);
    const int failed = v->failed;
    /* the ring still holds it until dslots later lines are dispatched */
    const int gone = line + (long long)s->dslots < s->lines_disp;
    const size_t len = (size_t)s->nel * s->dcols;
    const float *src = s->dring + (size_t)(line % s->dslots) * len;
    pthread_mutex_unlock(&v->mu);
    if (failed) return ATKDSP_E_FFT;
    if (gone) return ATKDSP_E_ARG;
    if (cap < len) return ATKDSP_E_CAP;
    if (atkgpu_d2h_sync(out, src, len * sizeof(float))) return ATKDSP_E_FFT;
    /* a line dispatched meanwhile may have reused the ring slot during the copy */
    pthread_mutex_lock(&v->mu);
    const int overrun = line + (long long)s->dslots < s->lines_disp;
    pthread_mutex_unlock(&v->mu);
    return overrun ? ATKDSP_E_OVERRUN : ATKDSP_OK;
}

/* channels of a served line, found on the device where the line lives
 * (atkdsp_detect_channels' policy; the floor, when NaN is passed, is the
 * median of the 5-bin-smoothed line as signal_id computes it): only the
 * records cross PCIe, strongest first */
ptrdiff_t atkgpu_server_detect(atkgpu_server *v, int sid, long long line, unsigned element, float floor_db,
                               float threshold_db, int gap_bins, int min_run, atkdsp_channel *out, size_t out_cap,
                               float *floor_used) {
    if (!v || (!out && out_cap) || sid < 0 || (unsigned)sid >= v->max_streams || line < 0) return ATKDSP_E_ARG;
    pthread_mutex_lock(&v->mu);
    sstream *s = &v->st[sid];
    if (!s->used || !s->dN || element >= s->nel || line >= disp_frames_at(s->g_sub, s->dN, s->dhop) / s->dF) {
        pthread_mutex_unlock(&v->mu);
        return ATKDSP_E_ARG;
    }
    while (s->lines_done <= line && !v->failed) pthread_cond_wait(&v->cv_done, &v->mu);
    const int failed = v->failed;
    const int gone = line + (long long)s->dslots < s->lines_disp;
    const size_t cols = s->dcols;
    const float *src = s->dring + (size_t)(line % s->dslots) * s->nel * cols + (size_t)element * cols;
    pthread_mutex_unlock(&v->mu);
    if (failed) return ATKDSP_E_FFT;
    if (gone) return ATKDSP_E_ARG;
    if (gap_bins < 2) gap_bins = 2;
    if (min_run < 2) min_run = 2;
    const size_t cap = out_cap < QREC_CAP ? out_cap : QREC_CAP;
    atkgpu_chan *rec_h = (atkgpu_chan *)v->q_h, *rec_d = (atkgpu_chan *)v->q_d;
    unsigned long long *found_h = (unsigned long long *)(rec_h + QREC_CAP);
    unsigned long long *found_d = (unsigned long long *)(rec_d + QREC_CAP);
    float *floor_h = (float *)(found_h + 1), *floor_d = (float *)(found_d + 1);
    pthread_mutex_lock(&v->qmu);
    int rc = atkgpu_detect_lines(v->qstream, src, cols, cols, 1, floor_db, floor_db != floor_db, threshold_db,
                                 gap_bins, min_run, v->q_scratch, rec_d, cap, found_d, floor_d) ||
             atkgpu_sync(v->qstream);
    const unsigned long long found = rc ? 0 : *found_h;
    const size_t kept = found < cap ? (size_t)found : cap;
    if (!rc) {
        memcpy(out, rec_h, kept * sizeof(atkgpu_chan));
        if (floor_used) *floor_used = *floor_h;
    }
    pthread_mutex_unlock(&v->qmu);
    if (rc) return ATKDSP_E_FFT;
    /* the query ran beside the display stream: the line's slot must not have
     * been reused by a line dispatched meanwhile */
    pthread_mutex_lock(&v->mu);
    const int overrun = line + (long long)s->dslots < s->lines_disp;
    pthread_mutex_unlock(&v->mu);
    if (overrun) return ATKDSP_E_OVERRUN;
    if (found > out_cap) return ATKDSP_E_CAP;
    atk_sort_channels(out, kept);
    return (ptrdiff_t)kept;
}

/* CFAR over served lines where they live: lines [first, first + n) of one
 * element, each a frame of dcols dB values, go through the handle's device
 * test and tracker as atkdsp_cfar_process would take them from the host.
 * The lines are copied into the handle's staging on its own stream; if the
 * ring overwrote the oldest of them meanwhile, ATKDSP_E_OVERRUN and the
 * handle is unchanged */
ptrdiff_t atkgpu_server_cfar(atkgpu_server *v, int sid, unsigned element, atkdsp_cfar *c, long long first, size_t n,
                             double frame_t0, double frame_dt, atkdsp_cfar_event *ev, size_t cap) {
    if (!v || !c || sid < 0 || (unsigned)sid >= v->max_streams || first < 0 || (!ev && n)) return ATKDSP_E_ARG;
    if (n == 0) return 0;
    if (n > atk_cfar_slab(c) || n > 256) return ATKDSP_E_ARG;
    if (cap < atkdsp_cfar_events_max(c, n)) return ATKDSP_E_CAP;
    const float *rows[256];
    pthread_mutex_lock(&v->mu);
    sstream *s = &v->st[sid];
    const long long last = first + (long long)n - 1;
    if (!s->used || !s->dN || element >= s->nel || s->dcols != atk_cfar_nbins(c) ||
        last >= disp_frames_at(s->g_sub, s->dN, s->dhop) / s->dF) {
        pthread_mutex_unlock(&v->mu);
        return ATKDSP_E_ARG;
    }
    while (s->lines_done <= last && !v->failed) pthread_cond_wait(&v->cv_done, &v->mu);
    const int failed = v->failed;
    const int gone = first + (long long)s->dslots < s->lines_disp || (long long)n > (long long)s->dslots;
    const size_t len = (size_t)s->nel * s->dcols;
    for (size_t i = 0; i < n; ++i)
        rows[i] = s->dring + (size_t)((first + (long long)i) % s->dslots) * len + (size_t)element * s->dcols;
    pthread_mutex_unlock(&v->mu);
    if (failed) return ATKDSP_E_FFT;
    if (gone) return ATKDSP_E_ARG;
    if (atk_cfar_stage_rows(c, rows, n)) return ATKDSP_E_FFT;
    pthread_mutex_lock(&v->mu);
    const int overrun = first + (long long)s->dslots < s->lines_disp;
    pthread_mutex_unlock(&v->mu);
    if (overrun) return ATKDSP_E_OVERRUN;
    return atk_cfar_process_staged(c, n, frame_t0, frame_dt, ev, cap);
}

/* the newest completed line (element 0) of each listed stream, all in one
 * launch: per stream i, out[i*per_cap ...] strongest first and counts[i]
 * (ATKDSP_E_CAP past min(per_cap, QSET_CAP) channels, the first ones then in
 * bin order; 0 with lines[i] = -1 when the stream has no line yet) */
int atkgpu_server_detect_latest(atkgpu_server *v, const int *sids, unsigned n, float threshold_db, int gap_bins,
                                int min_run, atkdsp_channel *out, size_t per_cap, ptrdiff_t *counts, long long *lines,
                                float *floors) {
    if (!v || !sids || !counts || n == 0 || n > v->max_streams || (!out && per_cap)) return ATKDSP_E_ARG;
    if (gap_bins < 2) gap_bins = 2;
    if (min_run < 2) min_run = 2;
    const size_t kcap = per_cap < QSET_CAP ? per_cap : QSET_CAP;
    const float **ptr_h = (const float **)v->qs_h;
    atkgpu_chan *rec_h = (atkgpu_chan *)(ptr_h + v->max_streams);
    unsigned long long *found_h = (unsigned long long *)(rec_h + (size_t)v->max_streams * QSET_CAP);
    float *floor_h = (float *)(found_h + v->max_streams);
    char *dbase = (char *)v->qs_d, *hbase = (char *)v->qs_h;
#define QDEV(p) ((void *)(dbase + ((char *)(p) - hbase)))
    unsigned slot[4096];
    long long line_of[4096];
    if (n > 4096) return ATKDSP_E_ARG;
    pthread_mutex_lock(&v->qmu);
    pthread_mutex_lock(&v->mu);
    unsigned nl = 0;
    size_t cols = 0;
    int rc = ATKDSP_OK;
    for (unsigned i = 0; i < n && !rc; ++i) {
        const int sid = sids[i];
        if (sid < 0 || (unsigned)sid >= v->max_streams || !v->st[sid].used || !v->st[sid].dN) { rc = ATKDSP_E_ARG; break; }
        const sstream *s = &v->st[sid];
        if (cols && s->dcols != cols) { rc = ATKDSP_E_ARG; break; }   /* one line length per query */
        cols = s->dcols;
        const long long L = s->lines_done - 1;
        if (lines) lines[i] = L;
        counts[i] = 0;
        if (L < 0) continue;
        if (L + (long long)s->dslots < s->lines_disp) {       /* its slot is already being reused */
            counts[i] = ATKDSP_E_OVERRUN;
            continue;
        }
        ptr_h[nl] = s->dring + (size_t)(L % s->dslots) * s->nel * cols;
        line_of[nl] = L;
        slot[nl++] = i;
    }
    const int failed = v->failed;
    pthread_mutex_unlock(&v->mu);
    if (!rc && failed) rc = ATKDSP_E_FFT;
    if (!rc && nl) {
        if (atkgpu_detect_line_set(v->qstream, (const float *const *)QDEV(ptr_h), cols, nl, threshold_db, gap_bins,
                                   min_run, (atkgpu_chan *)QDEV(rec_h), QSET_CAP, (unsigned long long *)QDEV(found_h),
                                   (float *)QDEV(floor_h)) ||
            atkgpu_sync(v->qstream))
            rc = ATKDSP_E_FFT;
    }
    /* lines dispatched during the query may have reused a slot: those
     * streams report ATKDSP_E_OVERRUN */
    unsigned char stale[4096];
    if (!rc) {
        pthread_mutex_lock(&v->mu);
        for (unsigned j = 0; j < nl; ++j) {
            const sstream *s = &v->st[sids[slot[j]]];
            stale[j] = line_of[j] + (long long)s->dslots < s->lines_disp;
        }
        pthread_mutex_unlock(&v->mu);
    }
    for (unsigned j = 0; !rc && j < nl; ++j) {
        const unsigned i = slot[j];
        if (stale[j]) {
            counts[i] = ATKDSP_E_OVERRUN;
            continue;
        }
        const unsigned long long found = found_h[j];
        const size_t kept = found < kcap ? (size_t)found : kcap;
        atkdsp_channel *o = out + (size_t)i * per_cap;
        memcpy(o, rec_h + (size_t)j * QSET_CAP, kept * sizeof(atkgpu_chan));
        if (found > kcap) {
            counts[i] = ATKDSP_E_CAP;
        } else {
            atk_sort_channels(o, kept);
            counts[i] = (ptrdiff_t)kept;
        }
        if (floors) floors[i] = floor_h[j];
    }
#undef QDEV
    pthread_mutex_unlock(&v->qmu);
    return rc;
}

int atkgpu_server_drain(atkgpu_server *v) {
    if (!v) return ATKDSP_E_ARG;
    pthread_mutex_lock(&v->mu);
    for (;;) {
        int pending = 0;
        for (unsigned i = 0; i < v->max_streams; ++i)
            if (v->st[i].used && v->st[i].completed < v->st[i].submitted) { pending = 1; break; }
        if (!pending || v->failed || !v->running) break;
        pthread_cond_wait(&v->cv_done, &v->mu);
    }
    const int rc = v->failed ? ATKDSP_E_FFT : ATKDSP_OK;
    pthread_mutex_unlock(&v->mu);
    return rc;
}

int atkgpu_server_stats(atkgpu_server *v, atkgpu_serve_stats *out, int reset) {
    if (!v || !out) return ATKDSP_E_ARG;
    pthread_mutex_lock(&v->mu);
    memset(out, 0, sizeof *out);
    out->chunks = v->chunks;
    out->samples = v->samples;
    out->batches = v->nbatches;
    out->hbm_used = v->used_bytes;
    out->hbm_budget = v->budget;
    for (unsigned i = 0; i < v->max_streams; ++i) out->streams += v->st[i].used ? 1u : 0u;
    out->busy_s = v->nlat ? v->t_last - v->t_first : 0.0;
    const size_t n = v->nlat;
    out->lat_count = (long long)n;
    out->lat_rel_res = exp(LAT_LOG_STEP) - 1.0;
    if (n) {
        out->mean_ms = v->lat_sum / (double)n;
        out->p50_ms = lat_at(v, (size_t)(0.50 * (double)(n - 1)));
        out->p99_ms = lat_at(v, (size_t)(0.99 * (double)(n - 1)));
        out->p999_ms = lat_at(v, (size_t)(0.999 * (double)(n - 1)));
        out->max_ms = v->lat_max;
    }
    if (reset) {
        memset(v->lat_hist, 0, LAT_NB * sizeof(uint64_t));
        v->lat_sum = 0.0;
        v->nlat = 0;
        v->chunks = 0;
        v->samples = 0;
        v->nbatches = 0;
    }
    pthread_mutex_unlock(&v->mu);
    return ATKDSP_OK;
}
