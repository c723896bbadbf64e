/* The NCO (reference src/nco.c:27-98, include/atkdsp.h:309-319).
 *
 *   atkdsp_nco_init / _set_freq   host arithmetic on the caller's struct,
 *                                 the reference's formulas
 *   atkdsp_nco_mix                the rotation on the GPU through the
 *                                 library's pool; the carried phase is
 *                                 wrapped on the host as the reference's
 *                                 atk_wrap does
 *   atkgpu_nco_*                  a device-resident mixer whose FP64 phase
 *                                 lives in device memory across calls (and is
 *                                 logged after each), for long runs fed from
 *                                 device buffers
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "internal.h"
#include "../../../include/atkgpu.h"

static double host_wrap(double ph) {
    ph = fmod(ph, 2.0 * ATK_PI);
    if (ph >= ATK_PI) ph -= 2.0 * ATK_PI;
    if (ph < -ATK_PI) ph += 2.0 * ATK_PI;
    return ph;
}

void atkdsp_nco_set_freq(atkdsp_nco *nco, double freq_hz, double sample_rate) {
    if (!nco) return;
    /* -2*pi*f/fs: a channel at +f mixes down to DC */
    nco->step = sample_rate > 0.0 ? -2.0 * ATK_PI * freq_hz / sample_rate : 0.0;
}

void atkdsp_nco_init(atkdsp_nco *nco, double freq_hz, double sample_rate) {
    if (!nco) return;
    nco->phase = 0.0;
    atkdsp_nco_set_freq(nco, freq_hz, sample_rate);
}

/* one piece of a pipelined mix: samples [i0, i0 + cnt) of the slot's buffers,
 * at absolute index base + i0 */
#define NCO_PIECE ((size_t)1 << 19)
typedef struct {
    const atk_slot *s;
    size_t base;
    double ph0, step;
} nco_piece_ctx;

static int nco_piece(void *vc, void *st, size_t i0, size_t cnt) {
    const nco_piece_ctx *c = (const nco_piece_ctx *)vc;
    return atkgpu_nco_mix(st, (const atkdsp_cf32 *)c->s->d_a + i0, (atkdsp_cf32 *)c->s->d_b + i0, cnt, c->base + i0,
                          c->ph0, c->step);
}

void atkdsp_nco_mix(atkdsp_nco *nco, const atkdsp_cf32 *in, atkdsp_cf32 *out, size_t n) {
    if (!nco || !in || !out) return;
    if (nco->step == 0.0) {
        if (out != in) memmove(out, in, n * sizeof(atkdsp_cf32));
        return;
    }
    if (n == 0) return;
    atk_pool *pool = atk_pool_get();
    if (!pool) return;                         /* no device: a void call does nothing */
    const double ph0 = nco->phase, step = nco->step;
    const int si = atk_claim(pool->slot, pool->nslot);
    atk_slot *s = &pool->slot[si];
    int rc = 0;
    /* each sample's phase comes from its absolute index, so the pieces of
     * the pipelined copy cannot change a value; in place (out == in) is safe:
     * a piece's output is written only after its input was read */
    for (size_t a = 0; a < n && !rc; a += pool->chunk) {
        const size_t c = n - a < pool->chunk ? n - a : pool->chunk;
        nco_piece_ctx ctx = {s, a, ph0, step};
        rc = atk_pipe(s->stream, s->stream2, in + a, sizeof(atkdsp_cf32), s->d_a, out + a, sizeof(atkdsp_cf32), s->d_b,
                      c, NCO_PIECE, nco_piece, &ctx);
    }
    atk_release(pool->slot, si);
    if (!rc) nco->phase = host_wrap(ph0 + step * (double)n);
}

struct atkgpu_nco {
    void *stream;
    atkdsp_cf32 *d_in, *d_out;
    double *d_state;           /* two {phase, step} slots, used in turn */
    double *d_log;             /* the phase after each run */
    size_t chunk_max, log_max, nlog;
    int par;
    long long samples;
};

void atkgpu_nco_destroy(atkgpu_nco *v) {
    if (!v) return;
    if (v->stream) atkgpu_sync(v->stream);
    atkgpu_dfree(v->d_in);
    atkgpu_dfree(v->d_out);
    atkgpu_dfree(v->d_state);
    atkgpu_dfree(v->d_log);
    if (v->stream) atkgpu_stream_destroy(v->stream);
    free(v);
}

atkgpu_nco *atkgpu_nco_create(double freq_hz, double sample_rate, size_t chunk_max, size_t log_max) {
    if (chunk_max < 1 || !(sample_rate > 0)) return NULL;
    if (atkgpu_init() != 0) return NULL;
    atkgpu_nco *v = (atkgpu_nco *)calloc(1, sizeof *v);
    if (!v) return NULL;
    v->chunk_max = chunk_max;
    v->log_max = log_max;
    atkdsp_nco h;
    atkdsp_nco_init(&h, freq_hz, sample_rate);
    const double st[4] = {h.phase, h.step, h.phase, h.step};
    int ok = !atkgpu_stream_create(&v->stream);
    if (ok) {
        v->d_in = (atkdsp_cf32 *)atkgpu_dmalloc(chunk_max * sizeof(atkdsp_cf32));
        v->d_out = (atkdsp_cf32 *)atkgpu_dmalloc(chunk_max * sizeof(atkdsp_cf32));
        v->d_state = (double *)atkgpu_dmalloc(4 * sizeof(double));
        v->d_log = (double *)atkgpu_dmalloc((log_max ? log_max : 1) * sizeof(double));
        ok = v->d_in && v->d_out && v->d_state && v->d_log;
    }
    if (ok) ok = !atkgpu_h2d(v->stream, v->d_state, st, sizeof st) && !atkgpu_sync(v->stream);
    if (!ok) {
        atkgpu_nco_destroy(v);
        return NULL;
    }
    return v;
}

void *atkgpu_nco_input(atkgpu_nco *v) { return v ? v->d_in : NULL; }
void *atkgpu_nco_output(atkgpu_nco *v) { return v ? v->d_out : NULL; }

int atkgpu_nco_upload(atkgpu_nco *v, const atkdsp_cf32 *in, size_t n) {
    if (!v || !in || n > v->chunk_max) return ATKDSP_E_ARG;
    return atk_put(v->stream, v->d_in, in, n * sizeof(atkdsp_cf32)) ? ATKDSP_E_FFT : ATKDSP_OK;
}

int atkgpu_nco_run(atkgpu_nco *v, size_t n) {
    if (!v || n > v->chunk_max) return ATKDSP_E_ARG;
    double *logp = v->nlog < v->log_max ? v->d_log + v->nlog : NULL;
    const int rc = atkgpu_nco_run_dev(v->stream, v->d_in, v->d_out, n, v->d_state + 2 * v->par,
                                      v->d_state + 2 * (v->par ^ 1), logp);
    if (rc) return ATKDSP_E_FFT;
    if (logp) v->nlog++;
    v->par ^= 1;
    v->samples += (long long)n;
    return ATKDSP_OK;
}

int atkgpu_nco_sync(atkgpu_nco *v) {
    if (!v) return ATKDSP_E_ARG;
    return atkgpu_sync(v->stream) ? ATKDSP_E_FFT : ATKDSP_OK;
}

ptrdiff_t atkgpu_nco_read(atkgpu_nco *v, atkdsp_cf32 *out, size_t n) {
    if (!v || !out || n > v->chunk_max) return ATKDSP_E_ARG;
    int rc = atk_get(v->stream, out, v->d_out, n * sizeof(atkdsp_cf32));
    if (!rc) rc = atkgpu_sync(v->stream);
    return rc ? ATKDSP_E_FFT : (ptrdiff_t)n;
}

int atkgpu_nco_phase(atkgpu_nco *v, double *phase, long long *samples) {
    if (!v) return ATKDSP_E_ARG;
    double st[2];
    int rc = atkgpu_d2h(v->stream, st, v->d_state + 2 * v->par, sizeof st);
    if (!rc) rc = atkgpu_sync(v->stream);
    if (rc) return ATKDSP_E_FFT;
    if (phase) *phase = st[0];
    if (samples) *samples = v->samples;
    return ATKDSP_OK;
}

ptrdiff_t atkgpu_nco_log(atkgpu_nco *v, double *phases, size_t cap) {
    if (!v || !phases) return ATKDSP_E_ARG;
    const size_t k = v->nlog < cap ? v->nlog : cap;
    int rc = k ? atkgpu_d2h(v->stream, phases, v->d_log, k * sizeof(double)) : 0;
    if (!rc) rc = atkgpu_sync(v->stream);
    return rc ? ATKDSP_E_FFT : (ptrdiff_t)k;
}
