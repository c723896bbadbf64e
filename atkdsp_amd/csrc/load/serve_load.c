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
 * usage: serve_load [--seconds S] [--streams N] [--speed X] [--budget-gib G]
 *                   [--max-batch B] [--seed K] [--mix c4|uniform] [--arrays A]
 *                   [--display 0|1]
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "atkgpu.h"

#define NTHREADS 8
#define CHUNK_MAX 65536
#define SRC_SAMPLES (1u << 18)

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
} lstream;

static atkgpu_server *g_srv;
static lstream *g_st;
static int g_ns;
static double g_t0, g_seconds, g_speed;

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
        s->pos = (s->pos + n) % SRC_SAMPLES;
        if (atkgpu_server_submit(g_srv, s->sid, p, bytes) < 0) break;
        s->submitted++;
        /* Poisson arrivals at the stream's rate: exponential gaps, mean chunk / fs */
        const double gap = -log(urand(&s->rng)) * ((CHUNK_MAX + 1) / 2.0) / (s->fs * g_speed);
        s->next_t += gap;
    }
    free(tmp);
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
    }
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
        ls->next_t = -log(urand(&ls->rng)) * ((CHUNK_MAX + 1) / 2.0) / (ls->fs * speed);
        total_rate += ls->fs * ls->nel;
        ++admitted;
    }
    g_ns = admitted;
    g_seconds = seconds;
    g_speed = speed;
    fprintf(stderr, "serve_load: %d streams admitted (%d refused), %.1f MS/s offered x%.1f, %.1f s\n", admitted,
            refused, total_rate / 1e6, speed, seconds);
    pthread_t th[NTHREADS];
    g_t0 = now_s() + 0.05;
    for (long t = 0; t < NTHREADS; ++t) pthread_create(&th[t], NULL, submitter, (void *)t);
    for (int t = 0; t < NTHREADS; ++t) pthread_join(th[t], NULL);
    atkgpu_server_drain(g_srv);
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
           "\"hbm_used_gib\": %.3f, \"hbm_budget_gib\": %.1f, \"display\": %d, \"lines\": %lld}\n",
           uniform ? "uniform" : "c4", arrays, admitted, refused, total_rate / 1e6, speed, seconds, st.chunks, st.samples, st.batches,
           st.busy_s > 0 ? st.samples / st.busy_s / 1e6 : 0.0, st.p50_ms, st.p99_ms, st.p999_ms, st.max_ms,
           st.mean_ms, st.hbm_used / 1073741824.0, st.hbm_budget / 1073741824.0, display, lines);
    atkgpu_server_destroy(g_srv);
    for (int s = 0; s < admitted; ++s) free(g_st[s].src);
    free(g_st);
    return 0;
}
