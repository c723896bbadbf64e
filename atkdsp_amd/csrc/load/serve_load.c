/* serve_load: replay the many-stream receiver workload against the serving
 * loop (include/atkgpu.h) and print one JSON line of throughput and chunk
 * latency.
 *
 * Streams (default 128): rates 2.4 / 10 / 20 / 40 MSPS (a quarter each),
 * formats cu8 / ci8 / ci16 / cf32, filterbank M in {64, 256, 1024} with 8..16
 * taps per branch at 2x. Chunks of 1..65536 samples (uniform) arrive as a
 * Poisson process at each stream's real rate (mean gap = mean chunk / fs),
 * times `--speed` (1 = real time; larger compresses time to find the
 * throughput ceiling). Submitter threads sleep until each arrival and submit
 * it; latency is measured by the server from each submit call to the
 * completion of the chunk's batch.
 *
 * With --mix uniform every stream is instead 40 MSPS ci16 into a 1024-channel,
 * 16-tap, 2x oversampled filterbank (the full-rate single-receiver shape).
 * --arrays A adds A coherent arrays of 8 cf32 elements at 20 MSPS into
 * 2048-channel, 12-tap, 2x filterbanks with W1 / W2 on the device; a chunk
 * holds all 8 elements.
 * Every single stream also makes display lines (--display 1, the default):
 * 4096-point Hann frames at hop 2048, max-hold over 50 ms of frames, each line
 * decimated to 1024 columns; a chunk's latency includes the lines it
 * completes.
 *
 * With --ring 1 the chunks go through a block ring per stream (atkdsp_ring_*,
 * --ring-blocks B blocks of the largest chunk, default 64, about 50 ms of a
 * stream's arrivals, in ordinary host memory), as from a
 * receiver process: the arrival threads become producers that claim, copy and
 * commit each chunk as a block (t_utc_ns = the commit time on the monotonic
 * clock), and --consumers C threads (default 4) poll their streams' rings with
 * atkgpu_server_submit_ring, sleeping RING_POLL_US when all of them are
 * empty. Producers, consumers and the server's two threads together stay
 * within the 16 cores the serving host gives the process: with more runnable
 * threads than cores, the scheduler parks one for a whole time slice (about
 * 10 ms measured), and a parked consumer or dispatcher is a latency spike. Latency is then timed from the submit_ring call; the JSON adds the
 * blocks lost (a producer that lapped its consumer), the delay from commit to
 * the call that took the block, and the time a submit_ring call took.
 *
 * With --detect 1 a consumer thread runs atkgpu_server_detect on every display
 * line as it completes (each stream's lines in turn; the floor is the median
 * of the smoothed line, found on the device; threshold 6 dB, gap 4, min run
 * 2), as a signal-finding consumer of the lines would; the JSON adds the
 * lines it covered, the channels found and the call latency. With --detect 2
 * the consumer instead asks every 50 ms (one line period) for the newest line
 * of every stream at once (atkgpu_server_detect_latest: one launch), and the
 * latency is that of the whole query. With --detect 3 the consumer runs CFAR
 * (atkgpu_server_cfar: CA, 16 + 16 cells, 4 guard, Pfa 1e-6, two-line
 * average, the burst tracker) over every completed line of every display
 * stream, each stream through a handle of its own made before the clock
 * starts; the JSON's detect_channels then counts the events.
 *
 * Before the clock starts every stream gets one chunk of CHUNK_MAX samples and
 * the server is drained and its statistics reset: the figures are the steady
 * state, without the first launch of each kernel (code object loading) or the
 * first touch of ring pages.
 *
 * usage: serve_load [--seconds S] [--streams N] [--speed X] [--budget-gib G]
 *                   [--max-batch B] [--seed K] [--mix c4|uniform] [--arrays A]
 *                   [--display 0|1] [--ring 0|1] [--consumers C] [--ring-blocks B]
 *                   [--detect 0|1|2|3]
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/resource.h>
#include <time.h>
#include "atkgpu.h"

#define NTHREADS 8
#define CHUNK_MAX 65536
#define SRC_SAMPLES (1u << 18)
#define RING_POLL_US 20

typedef struct {
    int sid, fmt, bps;
    unsigned nel;
    double fs;
    unsigned M, K;
    unsigned char *src;        /* SRC_SAMPLES samples of this stream's format */
    size_t pos;                /* read position in src, samples */
    unsigned long long rng;
    double next_t;
    long long submitted;
    void *ring_mem;            /* --ring: the stream's block ring and its consumer */
    atkdsp_ring *ring;
    int cons;
} lstream;

static atkgpu_server *g_srv;
static lstream *g_st;
static int g_ns;
static double g_t0, g_seconds, g_speed;
static int g_ring, g_consumers = 4, g_ring_blocks = 64, g_detect;
static long long g_det_lines, g_det_chans, g_det_missed;
static double *g_det_lat;
static size_t g_det_n, g_det_cap;
static volatile int g_producers_done;
static pthread_mutex_t g_pick_mu = PTHREAD_MUTEX_INITIALIZER;
static long long g_picked, g_lost;
static double g_pick_sum, g_pick_max, g_call_sum, g_call_max, g_pass_max, g_late_max;

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static unsigned long long splitmix(unsigned long long *x) {
    unsigned long long z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double urand(unsigned long long *x) { return ((double)(splitmix(x) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

/* noise at -30 dBFS plus a few tones, quantised to the stream's format */
static void make_source(lstream *s) {
    s->src = (unsigned char *)malloc((size_t)SRC_SAMPLES * s->bps * s->nel);
    const double sig = 0.0224;
    double pr[3], pi[3], wr[3], wi[3], a[3];
    for (int i = 0; i < 3; ++i) {
        const double f = (urand(&s->rng) - 0.5) * 0.9;
        a[i] = pow(10.0, -(10 + 40 * urand(&s->rng)) / 20);
        pr[i] = 1.0; pi[i] = 0.0;
        wr[i] = cos(6.283185307179586 * f); wi[i] = sin(6.283185307179586 * f);
    }
    for (size_t n = 0; n < SRC_SAMPLES; ++n) {
        /* Box-Muller pair */
        const double rr = sqrt(-2.0 * log(urand(&s->rng))), th = 6.283185307179586 * urand(&s->rng);
        double re = sig * rr * cos(th), im = sig * rr * sin(th);
        for (int i = 0; i < 3; ++i) {
            re += a[i] * pr[i];
            im += a[i] * pi[i];
            const double nr = pr[i] * wr[i] - pi[i] * wi[i];
            pi[i] = pr[i] * wi[i] + pi[i] * wr[i];
            pr[i] = nr;
        }
        for (unsigned e = 1; e < s->nel; ++e) {       /* array elements: the same wave, phase-shifted */
            const double c = cos(0.7 * e), sn = sin(0.7 * e);
            float *p = (float *)(s->src + (size_t)e * SRC_SAMPLES * 8);
            p[2 * n] = (float)(re * c - im * sn);
            p[2 * n + 1] = (float)(re * sn + im * c);
        }
        switch (s->fmt) {
        case ATKDSP_FMT_CU8:
            s->src[2 * n] = (unsigned char)fmin(255, fmax(0, round(127.5 + 127.5 * re)));
            s->src[2 * n + 1] = (unsigned char)fmin(255, fmax(0, round(127.5 + 127.5 * im)));
            break;
        case ATKDSP_FMT_CI8: {
            signed char *p = (signed char *)s->src;
            p[2 * n] = (signed char)fmin(127, fmax(-128, round(128 * re)));
            p[2 * n + 1] = (signed char)fmin(127, fmax(-128, round(128 * im)));
            break;
        }
        case ATKDSP_FMT_CI16: {
            short *p = (short *)s->src;
            p[2 * n] = (short)fmin(32767, fmax(-32768, round(32768 * re)));
            p[2 * n + 1] = (short)fmin(32767, fmax(-32768, round(32768 * im)));
            break;
        }
        default: {
            float *p = (float *)s->src;
            p[2 * n] = (float)re;
            p[2 * n + 1] = (float)im;
        }
        }
    }
}

static void *submitter(void *arg) {
    const int tid = (int)(long)arg;
    unsigned char *tmp = (unsigned char *)malloc((size_t)CHUNK_MAX * 8 * 8);
    double late = 0;               /* the most this thread started a chunk after its arrival time */
    for (;;) {
        /* the earliest next arrival among this thread's streams */
        int best = -1;
        for (int i = tid; i < g_ns; i += NTHREADS)
            if (best < 0 || g_st[i].next_t < g_st[best].next_t) best = i;
        if (best < 0) break;
        lstream *s = &g_st[best];
        if (s->next_t > g_seconds) break;
        /* sleep until the arrival (no spinning: the server's threads need the
         * cores, and latency is timed from the submit call, not the due time) */
        const double due = g_t0 + s->next_t;
        const double d = due - now_s();
        if (d > 0) {
            struct timespec ts = {(time_t)d, (long)((d - (double)(time_t)d) * 1e9)};
            nanosleep(&ts, NULL);
        } else if (-d > late) {
            late = -d;
        }
        const size_t n = 1 + (size_t)(splitmix(&s->rng) % CHUNK_MAX);
        const size_t bytes = n * (size_t)s->bps * s->nel;
        const unsigned char *p;
        if (s->nel == 1 && s->pos + n <= SRC_SAMPLES) {
            p = s->src + s->pos * s->bps;
        } else {
            /* element-major chunk: each element's n samples from its own source */
            for (unsigned e = 0; e < s->nel; ++e) {
                const unsigned char *src = s->src + (size_t)e * SRC_SAMPLES * s->bps;
                unsigned char *d = tmp + (size_t)e * n * s->bps;
                const size_t a = SRC_SAMPLES - s->pos < n ? SRC_SAMPLES - s->pos : n;
                memcpy(d, src + s->pos * s->bps, a * s->bps);
                if (a < n) memcpy(d + a * s->bps, src, (n - a) * s->bps);
            }
            p = tmp;
        }
        const size_t first = s->pos;
        s->pos = (s->pos + n) % SRC_SAMPLES;
        if (g_ring) {
            void *blk = atkdsp_ring_claim(s->ring);
            if (!blk) break;
            memcpy(blk, p, bytes);
            atkdsp_ring_meta m;
            memset(&m, 0, sizeof m);
            m.first_sample = first;
            m.t_utc_ns = (int64_t)(now_s() * 1e9);
            m.fs_hz = s->fs;
            m.format = s->fmt;
            m.nbytes = (uint32_t)bytes;
            if (atkdsp_ring_commit(s->ring, &m) != ATKDSP_OK) break;
        } else if (atkgpu_server_submit(g_srv, s->sid, p, bytes) < 0) {
            break;
        }
        s->submitted++;
        /* Poisson arrivals at the stream's rate: exponential gaps, mean chunk / fs */
        const double gap = -log(urand(&s->rng)) * ((CHUNK_MAX + 1) / 2.0) / (s->fs * g_speed);
        s->next_t += gap;
    }
    free(tmp);
    pthread_mutex_lock(&g_pick_mu);
    if (late > g_late_max) g_late_max = late;
    pthread_mutex_unlock(&g_pick_mu);
    return NULL;
}

/* --ring: take every committed block of this thread's streams into the
 * server until the producers are done and the rings are empty */
static void *ring_consumer(void *arg) {
    const int tid = (int)(long)arg;
    long long picked = 0, lost_n = 0;
    double sum = 0, mx = 0, csum = 0, cmx = 0, pmx = 0;
    for (;;) {
        const int done = g_producers_done;
        int took = 0;
        const double p0 = now_s();
        for (int i = tid; i < g_ns; i += g_consumers) {
            lstream *s = &g_st[i];
            for (;;) {
                atkdsp_ring_meta m;
                uint64_t lost = 0;
                const double c0 = now_s();
                const long long t = atkgpu_server_submit_ring(g_srv, s->sid, s->ring, s->cons, &m, &lost);
                const double c1 = now_s();
                lost_n += (long long)lost;
                if (t < 0) break;
                csum += c1 - c0;
                if (c1 - c0 > cmx) cmx = c1 - c0;
                const double d = c0 - 1e-9 * (double)m.t_utc_ns;
                sum += d;
                if (d > mx) mx = d;
                ++picked;
                took = 1;
            }
        }
        if (now_s() - p0 > pmx) pmx = now_s() - p0;
        if (!took) {
            if (done) break;
            struct timespec ts = {0, RING_POLL_US * 1000L};
            nanosleep(&ts, NULL);
        }
    }
    pthread_mutex_lock(&g_pick_mu);
    g_picked += picked;
    g_lost += lost_n;
    g_pick_sum += sum;
    if (mx > g_pick_max) g_pick_max = mx;
    g_call_sum += csum;
    if (cmx > g_call_max) g_call_max = cmx;
    if (pmx > g_pass_max) g_pass_max = pmx;
    pthread_mutex_unlock(&g_pick_mu);
    return NULL;
}

static int cmp_dbl(const void *a, const void *b) {
    const double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

/* --detect: every completed line of every display stream, in turn */
static void *line_detector(void *arg) {
    (void)arg;
    long long *next = (long long *)calloc((size_t)g_ns, sizeof(long long));
    atkdsp_channel ch[512];
    for (;;) {
        const int done = g_producers_done;
        int took = 0;
        for (int i = 0; i < g_ns; ++i) {
            const long long nl = atkgpu_server_lines(g_srv, g_st[i].sid);
            if (nl <= 0) continue;
            if (nl - next[i] > 3) {                   /* fell behind the line ring: skip to the newest */
                g_det_missed += nl - 1 - next[i];
                next[i] = nl - 1;
            }
            for (; next[i] < nl; ++next[i]) {
                float fl;
                const double t0 = now_s();
                const ptrdiff_t n = atkgpu_server_detect(g_srv, g_st[i].sid, next[i], 0, NAN, 6.0f, 4, 2, ch, 512, &fl);
                const double t1 = now_s();
                if (n < 0 && n != ATKDSP_E_CAP) continue;
                if (g_det_n < g_det_cap) g_det_lat[g_det_n++] = (t1 - t0) * 1e3;
                g_det_lines++;
                g_det_chans += n < 0 ? 512 : n;
                took = 1;
            }
        }
        if (!took) {
            if (done) break;
            struct timespec ts = {0, 200000L};
            nanosleep(&ts, NULL);
        }
    }
    free(next);
    return NULL;
}

/* --detect 3: every completed line of every display stream through that
 * stream's CFAR handle, on the device; a stream that fell behind the line ring
 * skips to its newest line and its handle starts afresh */
static atkdsp_cfar **g_cfar;
static void *line_cfar(void *arg) {
    (void)arg;
    long long *next = (long long *)calloc((size_t)g_ns, sizeof(long long));
    atkdsp_cfar_event *ev = (atkdsp_cfar_event *)malloc(4 * 256 * sizeof(atkdsp_cfar_event));
    for (;;) {
        const int done = g_producers_done;
        int took = 0;
        for (int i = 0; i < g_ns; ++i) {
            if (!g_cfar[i]) continue;
            const long long nl = atkgpu_server_lines(g_srv, g_st[i].sid);
            if (nl <= next[i]) continue;
            if (nl - next[i] > 3) {
                g_det_missed += nl - 1 - next[i];
                next[i] = nl - 1;
                atkdsp_cfar_reset(g_cfar[i]);
            }
            const size_t n = (size_t)(nl - next[i]);
            const double t0 = now_s();
            const ptrdiff_t got = atkgpu_server_cfar(g_srv, g_st[i].sid, 0, g_cfar[i], next[i], n,
                                                     0.05 * (double)next[i], 0.05, ev, 4 * 256);
            const double t1 = now_s();
            if (got < 0) {                            /* overwritten meanwhile: count them missed */
                g_det_missed += (long long)n;
                atkdsp_cfar_reset(g_cfar[i]);
            } else {
                if (g_det_n < g_det_cap) g_det_lat[g_det_n++] = (t1 - t0) * 1e3;
                g_det_lines += (long long)n;
                g_det_chans += got;
            }
            next[i] += (long long)n;
            took = 1;
        }
        if (!took) {
            if (done) break;
            struct timespec ts = {0, 200000L};
            nanosleep(&ts, NULL);
        }
    }
    free(ev);
    free(next);
    return NULL;
}

/* --detect 2: the newest line of every display stream, every 50 ms, one query */
static void *line_detector_latest(void *arg) {
    (void)arg;
    int *sids = (int *)malloc((size_t)g_ns * sizeof(int));
    ptrdiff_t *counts = (ptrdiff_t *)malloc((size_t)g_ns * sizeof(ptrdiff_t));
    long long *lines = (long long *)malloc((size_t)g_ns * sizeof(long long));
    atkdsp_channel *ch = (atkdsp_channel *)malloc((size_t)g_ns * 64 * sizeof(atkdsp_channel));
    int n = 0;
    for (int i = 0; i < g_ns; ++i)
        if (g_st[i].nel == 1) sids[n++] = g_st[i].sid;
    while (!g_producers_done && n) {
        const double t0 = now_s();
        if (atkgpu_server_detect_latest(g_srv, sids, (unsigned)n, 6.0f, 4, 2, ch, 64, counts, lines, NULL) == 0) {
            const double t1 = now_s();
            if (g_det_n < g_det_cap) g_det_lat[g_det_n++] = (t1 - t0) * 1e3;
            for (int i = 0; i < n; ++i)
                if (lines[i] >= 0) {
                    g_det_lines++;
                    g_det_chans += counts[i] < 0 ? 64 : counts[i];
                }
        }
        struct timespec ts = {0, 50000000L};
        nanosleep(&ts, NULL);
    }
    free(sids);
    free(counts);
    free(lines);
    free(ch);
    return NULL;
}

int main(int argc, char **argv) {
    double seconds = 5.0, speed = 1.0, budget_gib = 32.0;
    int streams = 128, max_batch = 256;
    unsigned long long seed = 4000;
    int uniform = 0, arrays = 0, display = 1;
    for (int i = 1; i + 1 < argc; i += 2) {
        if (!strcmp(argv[i], "--seconds")) seconds = atof(argv[i + 1]);
        else if (!strcmp(argv[i], "--streams")) streams = atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "--speed")) speed = atof(argv[i + 1]);
        else if (!strcmp(argv[i], "--budget-gib")) budget_gib = atof(argv[i + 1]);
        else if (!strcmp(argv[i], "--max-batch")) max_batch = atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "--seed")) seed = strtoull(argv[i + 1], NULL, 10);
        else if (!strcmp(argv[i], "--mix")) uniform = !strcmp(argv[i + 1], "uniform");
        else if (!strcmp(argv[i], "--arrays")) arrays = atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "--display")) display = atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "--ring")) g_ring = atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "--consumers")) g_consumers = atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "--ring-blocks")) g_ring_blocks = atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "--detect")) g_detect = atoi(argv[i + 1]);
    }
    if (g_consumers < 1 || g_consumers > NTHREADS) { fprintf(stderr, "serve_load: --consumers 1..%d\n", NTHREADS); return 2; }
    g_srv = atkgpu_server_create((size_t)(budget_gib * (1u << 30)), (unsigned)(streams + arrays),
                                 (unsigned)max_batch, (size_t)1024 << 20);
    if (!g_srv) { fprintf(stderr, "serve_load: no server (no GPU?)\n"); return 2; }
    static const double rates[4] = {2.4e6, 10e6, 20e6, 40e6};
    static const int fmts[4] = {ATKDSP_FMT_CU8, ATKDSP_FMT_CI8, ATKDSP_FMT_CI16, ATKDSP_FMT_CF32};
    static const unsigned Ms[3] = {64, 256, 1024};
    g_st = (lstream *)calloc(streams + arrays, sizeof(lstream));
    int admitted = 0, refused = 0;
    double total_rate = 0;
    for (int s = 0; s < streams + arrays; ++s) {
        lstream *ls = &g_st[admitted];
        ls->nel = 1;
        ls->rng = seed + (unsigned long long)s;
        ls->fs = rates[s % 4];
        ls->fmt = fmts[(s / 4) % 4];
        ls->bps = atkdsp_bytes_per_sample(ls->fmt);
        ls->M = Ms[(s / 16) % 3];
        ls->K = 8 + (unsigned)(splitmix(&ls->rng) % 9);
        if (uniform) {
            ls->fs = 40e6;
            ls->fmt = ATKDSP_FMT_CI16;
            ls->bps = 4;
            ls->M = 1024;
            ls->K = 16;
        }
        if (s >= streams) {                          /* a coherent 8-element array */
            ls->fs = 20e6;
            ls->fmt = ATKDSP_FMT_CF32;
            ls->bps = 8;
            ls->M = 2048;
            ls->K = 12;
            ls->nel = 8;
        }
        atkgpu_stream_cfg cfg = {ls->fmt, ls->fs, ls->M, 2, ls->K, 80.0, 0.0f, 0.0f, CHUNK_MAX, 2, ls->nel,
                                 ls->nel > 1};
        if (display && ls->nel == 1) {
            cfg.disp_nfft = 4096;
            cfg.disp_hop = 2048;
            cfg.disp_window = ATKDSP_WIN_HANN;
            cfg.disp_detector = ATKDSP_DET_MAX;
            const double fr = floor(0.05 * ls->fs / 2048 + 0.5);      /* 50 ms of frames */
            cfg.disp_frames = fr < 1 ? 1u : (unsigned)fr;
            cfg.disp_cols = 1024;
            cfg.disp_slots = 4;
        }
        const int sid = atkgpu_server_add_stream(g_srv, &cfg);
        if (sid < 0) { ++refused; continue; }
        ls->sid = sid;
        make_source(ls);
        if (g_ring) {
            const size_t bb = (size_t)CHUNK_MAX * ls->bps * ls->nel;
            const size_t rb = atkdsp_ring_bytes(bb, (size_t)g_ring_blocks, 1);
            if (!rb || posix_memalign(&ls->ring_mem, 64, rb)) { fprintf(stderr, "serve_load: ring\n"); return 2; }
            memset(ls->ring_mem, 0, rb);
            ls->ring = atkdsp_ring_init(ls->ring_mem, rb, bb, (size_t)g_ring_blocks, 1);
            ls->cons = ls->ring ? atkdsp_ring_open(ls->ring, 0) : -1;
            if (ls->cons < 0) { fprintf(stderr, "serve_load: ring\n"); return 2; }
        }
        ls->next_t = -log(urand(&ls->rng)) * ((CHUNK_MAX + 1) / 2.0) / (ls->fs * speed);
        total_rate += ls->fs * ls->nel;
        ++admitted;
    }
    g_ns = admitted;
    g_seconds = seconds;
    g_speed = speed;
    fprintf(stderr, "serve_load: %d streams admitted (%d refused), %.1f MS/s offered x%.1f, %.1f s\n", admitted,
            refused, total_rate / 1e6, speed, seconds);
    /* warm-up: one full chunk per stream, then a clean slate */
    for (int s = 0; s < admitted; ++s) {
        lstream *ls = &g_st[s];
        const size_t n = ls->nel == 1 ? CHUNK_MAX : CHUNK_MAX / 8;
        if (atkgpu_server_submit(g_srv, ls->sid, ls->src, n * (size_t)ls->bps * ls->nel) < 0) {
            fprintf(stderr, "serve_load: warm-up submit failed\n");
            return 2;
        }
    }
    atkgpu_server_drain(g_srv);
    {
        atkgpu_serve_stats w;
        atkgpu_server_stats(g_srv, &w, 1);
    }
    pthread_t th[NTHREADS], rc[NTHREADS], dth;
    if (g_detect && display) {
        g_det_cap = 1u << 22;
        g_det_lat = (double *)malloc(g_det_cap * sizeof(double));
    }
    if (g_detect == 3 && display) {
        atkdsp_cfar_config ccfg;
        atkdsp_cfar_defaults(&ccfg);
        ccfg.k_frames = 2;
        g_cfar = (atkdsp_cfar **)calloc((size_t)g_ns, sizeof(atkdsp_cfar *));
        for (int i = 0; i < admitted; ++i)
            if (g_st[i].nel == 1) g_cfar[i] = atkdsp_cfar_create(1024, &ccfg);
    }
    struct rusage ru0, ru1;
    getrusage(RUSAGE_SELF, &ru0);
    const double w0 = now_s();
    g_t0 = now_s() + 0.05;
    for (long t = 0; t < NTHREADS; ++t) pthread_create(&th[t], NULL, submitter, (void *)t);
    if (g_ring)
        for (long t = 0; t < g_consumers; ++t) pthread_create(&rc[t], NULL, ring_consumer, (void *)t);
    if (g_det_lat)
        pthread_create(&dth, NULL, g_detect == 3 ? line_cfar : g_detect == 2 ? line_detector_latest : line_detector,
                       NULL);
    for (int t = 0; t < NTHREADS; ++t) pthread_join(th[t], NULL);
    g_producers_done = 1;
    if (g_ring)
        for (int t = 0; t < g_consumers; ++t) pthread_join(rc[t], NULL);
    atkgpu_server_drain(g_srv);
    if (g_det_lat) pthread_join(dth, NULL);
    getrusage(RUSAGE_SELF, &ru1);
    const double cpu_s = (double)(ru1.ru_utime.tv_sec - ru0.ru_utime.tv_sec) +
                         1e-6 * (double)(ru1.ru_utime.tv_usec - ru0.ru_utime.tv_usec) +
                         (double)(ru1.ru_stime.tv_sec - ru0.ru_stime.tv_sec) +
                         1e-6 * (double)(ru1.ru_stime.tv_usec - ru0.ru_stime.tv_usec);
    const double cpus = cpu_s / (now_s() - w0);
    double dp50 = 0, dp99 = 0, dmax = 0;
    if (g_det_n) {
        qsort(g_det_lat, g_det_n, sizeof(double), cmp_dbl);
        dp50 = g_det_lat[(size_t)(0.50 * (double)(g_det_n - 1))];
        dp99 = g_det_lat[(size_t)(0.99 * (double)(g_det_n - 1))];
        dmax = g_det_lat[g_det_n - 1];
    }
    atkgpu_serve_stats st;
    atkgpu_server_stats(g_srv, &st, 0);
    long long lines = 0;
    for (int s = 0; s < admitted; ++s) {
        const long long l = atkgpu_server_lines(g_srv, g_st[s].sid);
        if (l > 0) lines += l;
    }
    printf("{\"mix\": \"%s\", \"arrays\": %d, \"streams\": %d, \"refused\": %d, \"offered_msps\": %.2f, \"speed\": %.2f, \"seconds\": %.2f, "
           "\"chunks\": %lld, \"samples\": %lld, \"batches\": %lld, \"msamples_per_s\": %.2f, "
           "\"p50_ms\": %.4f, \"p99_ms\": %.4f, \"p999_ms\": %.4f, \"max_ms\": %.4f, \"mean_ms\": %.4f, "
           "\"latency_chunks\": %lld, \"latency_resolution\": %.4f, "
           "\"hbm_used_gib\": %.3f, \"hbm_budget_gib\": %.1f, \"display\": %d, \"lines\": %lld, "
           "\"ring\": %d, \"ring_lost\": %lld, \"ring_pickup_mean_ms\": %.4f, \"ring_pickup_max_ms\": %.4f, "
           "\"ring_call_mean_us\": %.2f, \"ring_call_max_ms\": %.4f, \"consumers\": %d, \"ring_blocks\": %d, \"ring_pass_max_ms\": %.3f, \"arrival_late_max_ms\": %.3f, "
           "\"detect\": %d, \"detect_lines\": %lld, \"detect_missed\": %lld, \"detect_channels\": %lld, "
           "\"detect_p50_ms\": %.4f, \"detect_p99_ms\": %.4f, \"detect_max_ms\": %.4f, \"detect_latency_calls\": %zu, "
           "\"host_cpus_busy\": %.2f}\n",
           uniform ? "uniform" : "c4", arrays, admitted, refused, total_rate / 1e6, speed, seconds, st.chunks, st.samples, st.batches,
           st.busy_s > 0 ? st.samples / st.busy_s / 1e6 : 0.0, st.p50_ms, st.p99_ms, st.p999_ms, st.max_ms,
           st.mean_ms, st.lat_count, st.lat_rel_res, st.hbm_used / 1073741824.0, st.hbm_budget / 1073741824.0, display, lines, g_ring, g_lost,
           g_picked ? 1e3 * g_pick_sum / (double)g_picked : 0.0, 1e3 * g_pick_max,
           g_picked ? 1e6 * g_call_sum / (double)g_picked : 0.0, 1e3 * g_call_max, g_ring ? g_consumers : 0, g_ring ? g_ring_blocks : 0, 1e3 * g_pass_max, 1e3 * g_late_max,
           g_det_lat ? g_detect : 0, g_det_lines, g_det_missed, g_det_chans, dp50, dp99, dmax, (size_t)g_det_n, cpus);
    if (g_cfar)
        for (int i = 0; i < g_ns; ++i) atkdsp_cfar_destroy(g_cfar[i]);
    free(g_cfar);
    atkgpu_server_destroy(g_srv);
    for (int s = 0; s < admitted; ++s) {
        free(g_st[s].src);
        free(g_st[s].ring_mem);
    }
    free(g_det_lat);
    free(g_st);
    return 0;
}
