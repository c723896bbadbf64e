/* Copies between the caller's host buffers and the device, for the drop-in
 * entry points, whose callers pass ordinary (pageable) memory.
 *
 * Pageable memory cannot be the target of a DMA: the runtime stages it
 * itself, one piece at a time through one host thread (about 28 GB/s in and
 * 17 GB/s out on an MI355X host), and a pitched copy row by row (a 1024-row
 * device-to-host copy of 32 MB took 10.6 ms). Here a large copy moves in
 * 4 MiB pieces through a ring of pinned buffers: the DMA of one piece (about
 * 56 GB/s each way) overlaps the host copy of the next, and the host copies
 * are split over a small team of threads.
 *
 * Engines: a few, each a ring of XF_NBUF pinned pieces with an event per
 * piece, claimed by CAS like the device slots (one caller per engine at a
 * time) and made on first use. A piece is rewritten only after the event of
 * its last DMA has completed, whichever engine user recorded it. The copy team
 * is shared: a caller that finds it busy copies on its own thread. */
#define _GNU_SOURCE
#include <pthread.h>
#include <sched.h>
#include <stdint.h>
#include <string.h>
#include "internal.h"

#define XF_PIECE ((size_t)4 << 20)
#define XF_NBUF 4
#define XF_NENG 4
/* helpers beside the calling thread: up to XF_HELPERS, and fewer than the
 * library's thread count (atkdsp_get_threads: the host's cores unless the
 * caller set it), counted at the first large copy. Host copies of pageable
 * memory saturate near eight threads on an MI355X host (16 cores): a 50 ms
 * 1024-channel filterbank call took 1.05-1.08 ms with eight threads,
 * 1.19-1.22 ms with four and 1.08-1.11 ms with twelve */
#ifndef XF_HELPERS
#define XF_HELPERS 7
#endif
#define XF_SPLIT ((size_t)256 << 10)      /* smaller copies stay on one thread */
#define XF_STAGE_MIN ((size_t)1 << 20)    /* smaller copies go to the runtime directly */

/* ---- the copy team ------------------------------------------------------------ */
/* One job: nrows rows of width bytes, from src (row pitch spitch) to dst (row
 * pitch dpitch). Rows are shared out among the threads; a single row is cut
 * into byte ranges instead. */
typedef struct {
    char *dst;
    const char *src;
    size_t dpitch, spitch, width, nrows;
} xf_job;

static struct {
    pthread_mutex_t mu;
    pthread_cond_t cv;
    atomic_uint gen;                      /* one more per job                */
    atomic_int left;                      /* helpers still on the job        */
    atomic_int held;                      /* a caller owns the team          */
    xf_job job;
    int nhelp;                            /* helpers running                 */
} g_team = {PTHREAD_MUTEX_INITIALIZER, PTHREAD_COND_INITIALIZER};
static pthread_once_t g_team_once = PTHREAD_ONCE_INIT;

static inline void cpu_relax(void) {
#if defined(__x86_64__) || defined(__i386__)
    __builtin_ia32_pause();
#else
    sched_yield();
#endif
}

/* share k of p of a job */
static void job_part(const xf_job *j, int k, int p) {
    if (j->nrows == 1) {
        const size_t per = ((j->width + (size_t)p - 1) / (size_t)p + 63) & ~(size_t)63;
        const size_t a = (size_t)k * per;
        if (a >= j->width) return;
        const size_t n = j->width - a < per ? j->width - a : per;
        memcpy(j->dst + a, j->src + a, n);
        return;
    }
    const size_t per = (j->nrows + (size_t)p - 1) / (size_t)p;
    const size_t r0 = (size_t)k * per;
    const size_t r1 = r0 + per < j->nrows ? r0 + per : j->nrows;
    for (size_t r = r0; r < r1; ++r) memcpy(j->dst + r * j->dpitch, j->src + r * j->spitch, j->width);
}

static void *helper(void *arg) {
    const int id = (int)(intptr_t)arg;    /* 1 .. XF_HELPERS */
    unsigned seen = 0;
    for (;;) {
        unsigned g;
        /* spin a while (the next piece of a copy usually follows within tens
         * of microseconds), then sleep until the next job */
        int spins = 0;
        while ((g = atomic_load_explicit(&g_team.gen, memory_order_acquire)) == seen && spins < 20000) {
            cpu_relax();
            ++spins;
        }
        if (g == seen) {
            pthread_mutex_lock(&g_team.mu);
            while ((g = atomic_load_explicit(&g_team.gen, memory_order_acquire)) == seen)
                pthread_cond_wait(&g_team.cv, &g_team.mu);
            pthread_mutex_unlock(&g_team.mu);
        }
        seen = g;
        job_part(&g_team.job, id, g_team.nhelp + 1);
        atomic_fetch_sub_explicit(&g_team.left, 1, memory_order_acq_rel);
    }
    return NULL;
}

static void team_start(void) {
    pthread_attr_t at;
    pthread_attr_init(&at);
    pthread_attr_setdetachstate(&at, PTHREAD_CREATE_DETACHED);
    int n = 0;
    const int want = atkdsp_get_threads() - 1 < XF_HELPERS ? atkdsp_get_threads() - 1 : XF_HELPERS;
    for (int i = 1; i <= want; ++i) {
        pthread_t t;
        if (pthread_create(&t, &at, helper, (void *)(intptr_t)i) != 0) break;
        ++n;
    }
    pthread_attr_destroy(&at);
    g_team.nhelp = n;
}

static void team_copy(const xf_job *j) {
    const size_t bytes = j->width * j->nrows;
    int expect = 0;
    if (bytes >= XF_SPLIT) pthread_once(&g_team_once, team_start);
    if (bytes < XF_SPLIT || g_team.nhelp == 0 ||
        !atomic_compare_exchange_strong(&g_team.held, &expect, 1)) {
        job_part(j, 0, 1);
        return;
    }
    g_team.job = *j;
    atomic_store_explicit(&g_team.left, g_team.nhelp, memory_order_relaxed);
    pthread_mutex_lock(&g_team.mu);
    atomic_fetch_add_explicit(&g_team.gen, 1, memory_order_release);
    pthread_cond_broadcast(&g_team.cv);
    pthread_mutex_unlock(&g_team.mu);
    job_part(j, 0, g_team.nhelp + 1);
    while (atomic_load_explicit(&g_team.left, memory_order_acquire) > 0) cpu_relax();
    atomic_store(&g_team.held, 0);
}

/* ---- staging engines ------------------------------------------------------------ */
typedef struct {
    atomic_int busy;
    int state;                            /* 0 not made, 1 ready, -1 failed  */
    char *buf[XF_NBUF];
    void *ev[XF_NBUF];
    int armed[XF_NBUF];                   /* ev[b] recorded since last wait  */
    void *kev[2];                         /* atk_pipe: a piece's kernels done */
} xf_engine;

static xf_engine g_eng[XF_NENG];

static int engine_make(xf_engine *e) {
    for (int b = 0; b < XF_NBUF; ++b) {
        e->buf[b] = (char *)atkgpu_hmalloc(XF_PIECE);
        if (!e->buf[b] || atkgpu_event_create_order(&e->ev[b])) return -1;
        e->armed[b] = 0;
    }
    for (int i = 0; i < 2; ++i)
        if (atkgpu_event_create_order(&e->kev[i])) return -1;
    return 1;
}

/* an engine, or NULL when staging is unavailable (the caller then copies
 * directly) */
static xf_engine *engine_claim(void) {
    for (;;) {
        for (int i = 0; i < XF_NENG; ++i) {
            int expect = 0;
            xf_engine *e = &g_eng[i];
            if (!atomic_compare_exchange_strong(&e->busy, &expect, 1)) continue;
            if (e->state == 0) e->state = engine_make(e);
            if (e->state > 0) return e;
            atomic_store(&e->busy, 0);
            return NULL;
        }
        sched_yield();
    }
}

static void engine_release(xf_engine *e) { atomic_store(&e->busy, 0); }

/* piece b free for the host to write (its last DMA done) */
static int piece_wait(xf_engine *e, int b) {
    if (!e->armed[b]) return 0;
    e->armed[b] = 0;
    return atkgpu_event_sync(e->ev[b]);
}

static int piece_arm(xf_engine *e, int b, void *stream) {
    if (atkgpu_event_record(e->ev[b], stream)) return -4;
    e->armed[b] = 1;
    return 0;
}

int atk_put(void *stream, void *dst, const void *src, size_t bytes) {
    if (bytes < XF_STAGE_MIN) return atkgpu_h2d(stream, dst, src, bytes);
    xf_engine *e = engine_claim();
    if (!e) return atkgpu_h2d(stream, dst, src, bytes);
    int rc = 0;
    size_t k = 0;
    for (size_t a = 0; a < bytes && !rc; a += XF_PIECE, ++k) {
        const size_t n = bytes - a < XF_PIECE ? bytes - a : XF_PIECE;
        const int b = (int)(k % XF_NBUF);
        rc = piece_wait(e, b);
        if (rc) break;
        const xf_job j = {e->buf[b], (const char *)src + a, n, n, n, 1};
        team_copy(&j);
        rc = atkgpu_h2d(stream, (char *)dst + a, e->buf[b], n);
        if (!rc) rc = piece_arm(e, b, stream);
    }
    engine_release(e);
    return rc;
}

/* The pieces of a rows copy: whole rows per piece when a row fits in one,
 * otherwise each row cut into pieces. */
typedef struct {
    char *dst;
    const char *src;
    size_t dpitch, width, rows;
    size_t rpp;                           /* rows per piece (0: cut rows)    */
    size_t ppr;                           /* pieces per row when cut         */
} xf_rows;

static size_t rows_pieces(const xf_rows *r) {
    return r->rpp ? (r->rows + r->rpp - 1) / r->rpp : r->rows * r->ppr;
}

/* piece k: device bytes [*off, *off + *n) of the packed source, landing in
 * *nr host rows at *dst (row pitch dpitch, *w bytes each) */
static void rows_piece(const xf_rows *r, size_t k, size_t *off, size_t *n, char **dst, size_t *nr, size_t *w) {
    if (r->rpp) {
        const size_t r0 = k * r->rpp;
        *nr = r->rows - r0 < r->rpp ? r->rows - r0 : r->rpp;
        *off = r0 * r->width;
        *n = *nr * r->width;
        *dst = r->dst + r0 * r->dpitch;
        *w = r->width;
    } else {
        const size_t row = k / r->ppr, c0 = (k % r->ppr) * XF_PIECE;
        *nr = 1;
        *off = row * r->width + c0;
        *n = r->width - c0 < XF_PIECE ? r->width - c0 : XF_PIECE;
        *dst = r->dst + row * r->dpitch + c0;
        *w = *n;
    }
}

static int rows_issue(xf_engine *e, const xf_rows *r, size_t k, void *stream) {
    size_t off, n, nr, w;
    char *d;
    rows_piece(r, k, &off, &n, &d, &nr, &w);
    const int b = (int)(k % XF_NBUF);
    int rc = piece_wait(e, b);
    if (!rc) rc = atkgpu_d2h(stream, e->buf[b], r->src + off, n);
    if (!rc) rc = piece_arm(e, b, stream);
    return rc;
}

int atk_get_rows(void *stream, void *dst, size_t dpitch, const void *src, size_t width, size_t rows) {
    const size_t bytes = width * rows;
    if (!bytes) return 0;
    if (bytes < XF_STAGE_MIN) {
        if (dpitch == width) return atkgpu_d2h(stream, dst, src, bytes);
        return atkgpu_d2h_2d(stream, dst, dpitch, src, width, width, rows);
    }
    xf_engine *e = engine_claim();
    if (!e) {
        if (dpitch == width) return atkgpu_d2h(stream, dst, src, bytes);
        return atkgpu_d2h_2d(stream, dst, dpitch, src, width, width, rows);
    }
    xf_rows r = {(char *)dst, (const char *)src, dpitch, width, rows, 0, 0};
    if (dpitch == width) {                /* one packed run: pieces of bytes */
        r.width = bytes;
        r.rows = 1;
        r.dpitch = bytes;
    }
    if (r.width <= XF_PIECE)
        r.rpp = XF_PIECE / r.width;
    else
        r.ppr = (r.width + XF_PIECE - 1) / XF_PIECE;
    const size_t np = rows_pieces(&r);
    int rc = 0;
    for (size_t k = 0; k < np && k < XF_NBUF && !rc; ++k) rc = rows_issue(e, &r, k, stream);
    for (size_t k = 0; k < np && !rc; ++k) {
        const int b = (int)(k % XF_NBUF);
        rc = piece_wait(e, b);
        if (rc) break;
        size_t off, n, nr, w;
        char *d;
        rows_piece(&r, k, &off, &n, &d, &nr, &w);
        const xf_job j = {d, e->buf[b], r.dpitch, w, w, nr};
        team_copy(&j);
        if (k + XF_NBUF < np) rc = rows_issue(e, &r, k + XF_NBUF, stream);
    }
    engine_release(e);
    return rc;
}

int atk_get(void *stream, void *dst, const void *src, size_t bytes) {
    return atk_get_rows(stream, dst, bytes, src, bytes, 1);
}

/* ---- pipelined elementwise work ---------------------------------------------------- */
/* Pieces k = 0, 1, ...: the input goes through pinned piece k%2 to d_in on s;
 * the piece's kernels run on s; s2 waits for them (kev) and copies the output
 * into pinned piece 2 + k%2; the host copies piece k-1's output out meanwhile.
 * Without staging (small copies, or no engine), the same pieces run between
 * one copy in and one copy out, so the results do not depend on the route. */
int atk_pipe(void *s, void *s2, const void *in, size_t in_bpi, void *d_in, void *out, size_t out_bpi,
             const void *d_out, size_t n, size_t piece, atk_run_fn run, void *ctx) {
    if (!n) return 0;
    if (!piece) piece = n;
    const size_t big = in_bpi > out_bpi ? in_bpi : out_bpi;
    const size_t np = (n + piece - 1) / piece;
    xf_engine *e = NULL;
    if (s2 && piece * big <= XF_PIECE && n * (in_bpi + out_bpi) >= XF_STAGE_MIN && np > 1) e = engine_claim();
    if (!e) {
        int rc = atk_put(s, d_in, in, n * in_bpi);
        for (size_t k = 0; k < np && !rc; ++k) {
            const size_t i0 = k * piece, cnt = n - i0 < piece ? n - i0 : piece;
            rc = run(ctx, s, i0, cnt);
        }
        if (!rc) rc = atk_get(s, out, d_out, n * out_bpi);
        return rc;
    }
    int rc = 0;
    size_t drained = 0;                   /* pieces whose output is in `out` */
    for (size_t k = 0; k < np && !rc; ++k) {
        const size_t i0 = k * piece, cnt = n - i0 < piece ? n - i0 : piece;
        const int bi = (int)(k & 1), bo = 2 + (int)(k & 1);
        /* in */
        rc = piece_wait(e, bi);
        if (rc) break;
        const xf_job ji = {e->buf[bi], (const char *)in + i0 * in_bpi, cnt * in_bpi, cnt * in_bpi, cnt * in_bpi, 1};
        team_copy(&ji);
        rc = atkgpu_h2d(s, (char *)d_in + i0 * in_bpi, e->buf[bi], cnt * in_bpi);
        if (!rc) rc = piece_arm(e, bi, s);
        /* the piece's kernels, then its output on s2 */
        if (!rc) rc = run(ctx, s, i0, cnt);
        if (!rc) rc = atkgpu_event_record(e->kev[k & 1], s);
        if (!rc) rc = atkgpu_stream_wait_event(s2, e->kev[k & 1]);
        if (!rc) rc = piece_wait(e, bo);          /* piece k-2's output, drained already */
        if (!rc) rc = atkgpu_d2h(s2, e->buf[bo], (const char *)d_out + i0 * out_bpi, cnt * out_bpi);
        if (!rc) rc = piece_arm(e, bo, s2);
        /* the previous piece's output to the caller */
        if (!rc && k > 0) {
            const size_t p0 = (k - 1) * piece, pc = n - p0 < piece ? n - p0 : piece;
            const int pb = 2 + (int)((k - 1) & 1);
            rc = piece_wait(e, pb);
            if (!rc) {
                const xf_job jo = {(char *)out + p0 * out_bpi, e->buf[pb], pc * out_bpi, pc * out_bpi, pc * out_bpi, 1};
                team_copy(&jo);
                drained = k;
            }
        }
    }
    if (!rc && drained < np) {
        const size_t k = np - 1, p0 = k * piece, pc = n - p0;
        const int pb = 2 + (int)(k & 1);
        rc = piece_wait(e, pb);
        if (!rc) {
            const xf_job jo = {(char *)out + p0 * out_bpi, e->buf[pb], pc * out_bpi, pc * out_bpi, pc * out_bpi, 1};
            team_copy(&jo);
        }
    }
    engine_release(e);
    return rc;
}

/* ---- input in one run, output in ranges ---------------------------------------------- */
/* Pinned pieces 0 and 1 carry the input (alternately), 2 and 3 the ranges'
 * outputs (range r in piece 2 + r%2). A piece is rewritten only once the host
 * has waited for its last copy: the input piece two uploads back, the output
 * piece of range r-2, which the host wrote out before queuing range r. */
static int pipe_upload(xf_engine *e, void *s, char *d_in, const char *in, size_t *done, size_t upto, size_t *k) {
    int rc = 0;
    while (*done < upto && !rc) {
        const size_t n = upto - *done < XF_PIECE ? upto - *done : XF_PIECE;
        const int b = (int)(*k & 1);
        rc = piece_wait(e, b);
        if (rc) break;
        const xf_job j = {e->buf[b], in + *done, n, n, n, 1};
        team_copy(&j);
        rc = atkgpu_h2d(s, d_in + *done, e->buf[b], n);
        if (!rc) rc = piece_arm(e, b, s);
        *done += n;
        ++*k;
    }
    return rc;
}

static int pipe_drain(xf_engine *e, const atk_range *g, int b) {
    int rc = piece_wait(e, b);
    if (!rc && g->rows) {
        const xf_job j = {g->dst, e->buf[b], g->dpitch, g->width, g->width, g->rows};
        team_copy(&j);
    }
    return rc;
}

int atk_pipe_ranges(void *s, void *s2, void *d_in, const void *in, size_t in_bytes, size_t nr, atk_desc_fn desc,
                    atk_range_fn run, void *ctx) {
    xf_engine *e = (s2 && nr > 1) ? engine_claim() : NULL;
    int rc = 0;
    if (!e) {
        rc = atk_put(s, d_in, in, in_bytes);
        for (size_t r = 0; r < nr && !rc; ++r) {
            atk_range g;
            rc = desc(ctx, r, &g);
            if (!rc) rc = run(ctx, s, r);
            if (!rc && g.rows) rc = atk_get_rows(s, g.dst, g.dpitch, g.src, g.width, g.rows);
        }
        return rc;
    }
    size_t done = 0, k = 0;
    atk_range g[2];
    for (size_t r = 0; r < nr && !rc; ++r) {
        atk_range *cur = &g[r & 1];
        rc = desc(ctx, r, cur);
        if (!rc && cur->width * cur->rows > XF_PIECE) rc = -1;
        if (!rc) rc = pipe_upload(e, s, (char *)d_in, (const char *)in, &done,
                                  cur->need < in_bytes ? cur->need : in_bytes, &k);
        if (!rc) rc = run(ctx, s, r);
        const int bo = 2 + (int)(r & 1);
        if (!rc) rc = atkgpu_event_record(e->kev[r & 1], s);
        if (!rc) rc = atkgpu_stream_wait_event(s2, e->kev[r & 1]);
        if (!rc) rc = piece_wait(e, bo);                      /* range r-2's rows are out */
        if (!rc && cur->rows) rc = atkgpu_d2h(s2, e->buf[bo], cur->src, cur->width * cur->rows);
        if (!rc) rc = piece_arm(e, bo, s2);
        if (!rc && r > 0) rc = pipe_drain(e, &g[(r - 1) & 1], 2 + (int)((r - 1) & 1));
    }
    if (!rc) rc = pipe_drain(e, &g[(nr - 1) & 1], 2 + (int)((nr - 1) & 1));
    if (!rc) rc = pipe_upload(e, s, (char *)d_in, (const char *)in, &done, in_bytes, &k);
    engine_release(e);
    return rc;
}
