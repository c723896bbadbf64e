/* The overlap-save extractor (reference include/atkdsp.h:3252-3289,
 * src/channelizer.c:473-788), the arbitrary-ratio resampler it lands its
 * channels with (include/atkdsp.h:368-377, src/resample.c:100-252) and the
 * Kaiser lowpass design both use (include/atkdsp.h:505-506,
 * src/design.c:5-62).
 *
 * Extractor. Per element the device keeps the stream s = hist(ov) ++
 * leftover(lc) ++ in contiguous in one of two slab buffers; block i is
 * s[i*hop, i*hop + nfft) with hop = ov = nfft/2, so one batched FFT with input
 * distance hop transforms every block of a sub-call. Each channel then runs
 * the fold, its M-point inverse FFT, the FP64 mix to absolute time and, when
 * its rate differs from out_rate, the resampler, all on the device; the next
 * history s[m*hop, m*hop + ov) is copied to the other slab buffer and the
 * < hop leftover kept in a host mirror (a call that completes no block never
 * touches the device). Every block depends only on its own samples, and the
 * resampler's output positions only on counts, so outputs do not depend on
 * how the input is cut into calls.
 *
 * Resampler. The prototype, its forward differences and the history live on
 * the device; the position bookkeeping (base, K, in_off) stays on the host,
 * which counts a call's outputs with the reference's own loop before the
 * device computes them, one thread per output. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "internal.h"

/* ---- Kaiser lowpass (src/design.c:5-62) ---------------------------------------- */
static double design_i0(double x) {
    double sum = 1.0, term = 1.0;
    const double y = x * 0.5;
    for (int k = 1; k < 500; ++k) {
        term *= (y / (double)k);
        const double t2 = term * term;
        sum += t2;
        if (t2 < sum * 1e-17) break;
    }
    return sum;
}

static double design_beta(double a) {
    if (a > 50.0) return 0.1102 * (a - 8.7);
    if (a >= 21.0) return 0.5842 * pow(a - 21.0, 0.4) + 0.07886 * (a - 21.0);
    return 0.0;
}

static size_t design_length(double atten_db, double transition_hz, double rate) {
    if (transition_hz <= 0.0 || rate <= 0.0) return 0;
    const double dw = 2.0 * ATK_PI * transition_hz / rate;
    const double n = (atten_db > 8.0 ? (atten_db - 8.0) : 1.0) / (2.285 * dw);
    size_t N = (size_t)ceil(n) + 1;
    if (N < 3) N = 3;
    if ((N & 1) == 0) ++N;
    return N;
}

ptrdiff_t atkdsp_design_lowpass(double fp_hz, double fs_hz, double atten_db, double rate, float *out,
                                size_t out_cap) {
    if (rate <= 0.0 || fp_hz < 0.0 || fs_hz <= fp_hz || fs_hz > rate * 0.5 + 1e-9) return ATKDSP_E_ARG;
    const size_t N = design_length(atten_db, fs_hz - fp_hz, rate);
    if (!out) return (ptrdiff_t)N;
    if (out_cap < N) return ATKDSP_E_CAP;
    const double beta = design_beta(atten_db);
    const double i0b = design_i0(beta);
    const double fc = 0.5 * (fp_hz + fs_hz) / rate;
    const double M = (double)(N - 1) * 0.5;
    double sum = 0.0;
    for (int pass = 0; pass < 2; ++pass) {
        for (size_t n = 0; n < N; ++n) {
            const double t = (double)n - M;
            const double x = 2.0 * fc * t;
            const double sinc = (t == 0.0) ? 1.0 : sin(ATK_PI * x) / (ATK_PI * x);
            const double r = t / M;
            const double rr = r * r;
            const double w = design_i0(beta * sqrt(rr < 1.0 ? 1.0 - rr : 0.0)) / i0b;
            const double h = 2.0 * fc * sinc * w;
            if (pass == 0) sum += h;
            else out[n] = (float)(h / sum);
        }
        if (sum == 0.0) sum = 1.0;
    }
    return (ptrdiff_t)N;
}

/* ---- the resampler core: one set of taps, nel element histories --------------- */
typedef struct {
    size_t P, Q;
    float *d_p, *d_d;          /* device, P*Q each */
    atkdsp_cf32 *d_hist[2];    /* device, [nel][Q] each, used in turn */
    int hcur;
    unsigned nel;
    double in_rate, out_rate, step, base;
    uint64_t K, in_off;
} gpu_rs;

static void rs_free(gpu_rs *r) {
    atkgpu_dfree(r->d_p);
    atkgpu_dfree(r->d_d);
    atkgpu_dfree(r->d_hist[0]);
    atkgpu_dfree(r->d_hist[1]);
    memset(r, 0, sizeof *r);
}

static int rs_reset(gpu_rs *r, void *stream) {
    r->base = 0.0;
    r->K = 0;
    r->in_off = 0;
    r->hcur = 0;
    return atkgpu_fill_u32(stream, r->d_hist[0], 0, (size_t)r->nel * r->Q * 2);
}

/* the prototype at the interpolated rate, cut at the lower Nyquist, gain P
 * folded in, forward differences with p[PQ] = 0 (src/resample.c:130-167) */
static int rs_init(gpu_rs *r, double in_rate, double out_rate, double atten_db, unsigned nphase, unsigned nel,
                   void *stream) {
    memset(r, 0, sizeof *r);
    if (in_rate <= 0.0 || out_rate <= 0.0 || nel < 1) return ATKDSP_E_ARG;
    if (atten_db <= 0.0) atten_db = 60.0;
    const size_t P = nphase ? nphase : 64u;
    const double fs_proto = in_rate * (double)P;
    const double nyq = 0.5 * (in_rate < out_rate ? in_rate : out_rate);
    const double fp = 0.8 * nyq, fstop = nyq;
    const ptrdiff_t nt = atkdsp_design_lowpass(fp, fstop, atten_db, fs_proto, NULL, 0);
    if (nt <= 0) return ATKDSP_E_ARG;
    r->P = P;
    r->Q = ((size_t)nt + P - 1) / P;
    r->nel = nel;
    const size_t PQ = r->P * r->Q;
    float *p = (float *)calloc(PQ, sizeof(float));
    float *d = (float *)malloc(PQ * sizeof(float));
    int rc = (!p || !d) ? ATKDSP_E_NOMEM : 0;
    if (!rc && atkdsp_design_lowpass(fp, fstop, atten_db, fs_proto, p, (size_t)nt) != nt) rc = ATKDSP_E_ARG;
    if (!rc) {
        for (size_t j = 0; j < PQ; ++j) p[j] *= (float)P;
        for (size_t j = 0; j + 1 < PQ; ++j) d[j] = p[j + 1] - p[j];
        d[PQ - 1] = -p[PQ - 1];
        r->d_p = (float *)atkgpu_dmalloc(PQ * sizeof(float));
        r->d_d = (float *)atkgpu_dmalloc(PQ * sizeof(float));
        r->d_hist[0] = (atkdsp_cf32 *)atkgpu_dmalloc((size_t)nel * r->Q * sizeof(atkdsp_cf32));
        r->d_hist[1] = (atkdsp_cf32 *)atkgpu_dmalloc((size_t)nel * r->Q * sizeof(atkdsp_cf32));
        if (!r->d_p || !r->d_d || !r->d_hist[0] || !r->d_hist[1]) rc = ATKDSP_E_NOMEM;
    }
    if (!rc) rc = atkgpu_h2d(stream, r->d_p, p, PQ * sizeof(float));
    if (!rc) rc = atkgpu_h2d(stream, r->d_d, d, PQ * sizeof(float));
    r->in_rate = in_rate;
    r->out_rate = out_rate;
    r->step = in_rate / out_rate;
    if (!rc) rc = rs_reset(r, stream);
    if (!rc) rc = atkgpu_sync(stream);
    free(p);
    free(d);
    if (rc) rs_free(r);
    return rc < 0 ? rc : (rc ? ATKDSP_E_FFT : 0);
}

/* output K sits at base + K*step - in_off within the call's inputs; count the
 * outputs before n (the reference's loop, src/resample.c:214-243) */
static size_t rs_count(const gpu_rs *r, size_t n) {
    size_t k = 0;
    uint64_t K = r->K;
    const double inoff = (double)r->in_off;
    double local = r->base + (double)K * r->step - inoff;
    while (local < (double)n) {
        ++k;
        ++K;
        local = r->base + (double)K * r->step - inoff;
    }
    return k;
}

/* outputs of a call over n inputs per element (element e at d_in + e*ipitch)
 * into d_out + e*opitch; advances the position and the history */
static int rs_run(gpu_rs *r, void *stream, const atkdsp_cf32 *d_in, size_t ipitch, size_t n, unsigned nel,
                  atkdsp_cf32 *d_out, size_t opitch, size_t count) {
    if (!n) return 0;
    const int rc = atkgpu_rs_run(stream, d_in, ipitch, n, nel, r->d_hist[r->hcur], r->d_hist[r->hcur ^ 1], r->d_p,
                                 r->d_d, (unsigned)r->P, (unsigned)r->Q, r->base, r->step, (unsigned long long)r->K,
                                 (unsigned long long)r->in_off, d_out, opitch, count);
    if (rc) return ATKDSP_E_FFT;
    r->hcur ^= 1;
    r->K += count;
    r->in_off += n;
    return 0;
}

/* ---- the drop-in resampler ---------------------------------------------------------- */
#define RS_SLAB ((size_t)1 << 20)      /* input samples per sub-call */

struct atkdsp_arb_resampler {
    gpu_rs rs;
    void *stream;
    atkdsp_cf32 *d_in, *d_out;
    size_t ocap;                       /* outputs a sub-call can make */
};

void atkdsp_arb_resampler_destroy(atkdsp_arb_resampler *r) {
    if (!r) return;
    if (r->stream) atkgpu_sync(r->stream);
    rs_free(&r->rs);
    atkgpu_dfree(r->d_in);
    atkgpu_dfree(r->d_out);
    if (r->stream) atkgpu_stream_destroy(r->stream);
    free(r);
}

atkdsp_arb_resampler *atkdsp_arb_resampler_create(double in_rate, double out_rate, double atten_db,
                                                  unsigned nphase) {
    if (in_rate <= 0.0 || out_rate <= 0.0) return NULL;
    if (atkgpu_init() != 0) return NULL;
    atkdsp_arb_resampler *r = (atkdsp_arb_resampler *)calloc(1, sizeof *r);
    if (!r) return NULL;
    if (atkgpu_stream_create(&r->stream) || rs_init(&r->rs, in_rate, out_rate, atten_db, nphase, 1, r->stream)) {
        atkdsp_arb_resampler_destroy(r);
        return NULL;
    }
    r->ocap = (size_t)((double)RS_SLAB * out_rate / in_rate) + 64;
    r->d_in = (atkdsp_cf32 *)atkgpu_dmalloc(RS_SLAB * sizeof(atkdsp_cf32));
    r->d_out = (atkdsp_cf32 *)atkgpu_dmalloc(r->ocap * sizeof(atkdsp_cf32));
    if (!r->d_in || !r->d_out) { atkdsp_arb_resampler_destroy(r); return NULL; }
    return r;
}

void atkdsp_arb_resampler_reset(atkdsp_arb_resampler *r) {
    if (!r) return;
    rs_reset(&r->rs, r->stream);
    atkgpu_sync(r->stream);
}

/* re-anchor so the next output stays where it is, then change the step
 * (src/resample.c:185-192) */
void atkdsp_arb_resampler_set_ratio(atkdsp_arb_resampler *r, double ratio) {
    if (!r || ratio <= 0.0) return;
    r->rs.base = r->rs.base + (double)r->rs.K * r->rs.step;
    r->rs.K = 0;
    r->rs.out_rate = r->rs.in_rate * ratio;
    r->rs.step = 1.0 / ratio;
}

double atkdsp_arb_resampler_ratio(const atkdsp_arb_resampler *r) {
    return (r && r->rs.step > 0.0) ? 1.0 / r->rs.step : 0.0;
}

size_t atkdsp_arb_resampler_out_max(const atkdsp_arb_resampler *r, size_t n_in) {
    if (!r || r->rs.step <= 0.0) return 0;
    const double local = r->rs.base + (double)r->rs.K * r->rs.step - (double)r->rs.in_off;
    if (local >= (double)n_in) return 0;
    return (size_t)(((double)n_in - local) / r->rs.step) + 2u;
}

ptrdiff_t atkdsp_arb_resampler_process(atkdsp_arb_resampler *r, const atkdsp_cf32 *in, size_t n,
                                       atkdsp_cf32 *out, size_t out_cap) {
    if (!r || (!in && n) || !out) return ATKDSP_E_ARG;
    if (rs_count(&r->rs, n) > out_cap) return ATKDSP_E_CAP;
    size_t done = 0;
    for (size_t pos = 0; pos < n;) {
        size_t c = n - pos < RS_SLAB ? n - pos : RS_SLAB;
        size_t k = rs_count(&r->rs, c);
        while (k > r->ocap && c > 1) {                 /* a raised ratio: fewer inputs per sub-call */
            c /= 2;
            k = rs_count(&r->rs, c);
        }
        if (k > r->ocap) return ATKDSP_E_FFT;
        int rc = atk_put(r->stream, r->d_in, in + pos, c * sizeof(atkdsp_cf32));
        if (!rc) rc = rs_run(&r->rs, r->stream, r->d_in, 0, c, 1, r->d_out, 0, k);
        if (!rc && k) rc = atk_get(r->stream, out + done, r->d_out, k * sizeof(atkdsp_cf32));
        if (!rc) rc = atkgpu_sync(r->stream);
        if (rc) return ATKDSP_E_FFT;
        done += k;
        pos += c;
    }
    return (ptrdiff_t)done;
}

/* ---- the extractor --------------------------------------------------------------------- */
typedef struct {
    int active;
    double offset, bw, out_rate, chan_rate, delta;
    unsigned decim, shift;
    size_t M, ntaps, ovd, nvalid;
    atkdsp_cf32 *d_H;          /* nfft: the channel filter's frequency response */
    atkgpu_fftplan *iplan;     /* M-point */
    atkdsp_cf32 *d_iscr;       /* its scratch */
    atkdsp_cf32 *d_Y;          /* folds, inverse-transformed in place: [max_nel*mcap][M] */
    atkdsp_cf32 *d_V;          /* pre-resampler samples: [max_nel][mcap*nvalid] */
    int rs_on;
    gpu_rs rs;
    atkdsp_cf32 *d_out;        /* resampled outputs, [max_nel][ocap] */
    size_t ocap;
} osx_chan;

/* input samples per element per sub-call: about 2M samples over all elements */
#define OSX_SLAB_TOTAL ((size_t)1 << 21)
#define OSX_SLAB_MIN ((size_t)1 << 16)

struct atkdsp_osx {
    double fs, atten;
    size_t nfft, ov, hop;
    unsigned max_nel;
    int64_t blk;
    unsigned lc;
    void *stream;
    atkgpu_fftplan *plan;      /* nfft forward */
    size_t slab, spitch, mcap; /* input per sub-call, samples per element row, blocks per sub-call */
    atkdsp_cf32 *d_s[2];       /* rows, max_nel * spitch */
    int cur;
    atkdsp_cf32 *d_X;          /* max_nel * mcap * nfft forward transforms */
    atkdsp_cf32 *d_scr;        /* forward plan scratch */
    atkdsp_cf32 *lo;           /* host mirror of the leftover, max_nel * hop */
    osx_chan *chans;
    int nslots, cap_slots;
    /* the channel descriptors of a sub-call, device-visible, two sets used in
     * turn; a set is rewritten only after the launches that read it are done */
    atkgpu_osx_chan *desc_h[2];
    void *desc_d[2];
    void *desc_ev[2];
    int desc_used[2], desc_cap, dcur;
};

static void osx_free_chan(osx_chan *c) {
    if (!c->active) return;
    atkgpu_dfree(c->d_H);
    atkgpu_dfree(c->d_Y);
    atkgpu_dfree(c->d_V);
    atkgpu_fftplan_destroy(c->iplan);
    atkgpu_dfree(c->d_iscr);
    if (c->rs_on) rs_free(&c->rs);
    atkgpu_dfree(c->d_out);
    memset(c, 0, sizeof *c);
}

void atkdsp_osx_destroy(atkdsp_osx *x) {
    if (!x) return;
    if (x->stream) atkgpu_sync(x->stream);
    for (int i = 0; i < x->nslots; ++i) osx_free_chan(&x->chans[i]);
    free(x->chans);
    atkgpu_fftplan_destroy(x->plan);
    atkgpu_dfree(x->d_s[0]);
    atkgpu_dfree(x->d_s[1]);
    atkgpu_dfree(x->d_X);
    atkgpu_dfree(x->d_scr);
    free(x->lo);
    for (int i = 0; i < 2; ++i) {
        atkgpu_hfree(x->desc_h[i]);
        atkgpu_event_destroy(x->desc_ev[i]);
    }
    if (x->stream) atkgpu_stream_destroy(x->stream);
    free(x);
}

atkdsp_osx *atkdsp_osx_create(double fs, size_t nfft, double atten_db, unsigned max_nel) {
    if (!(fs > 0.0) || nfft < 64 || (nfft & (nfft - 1)) || max_nel < 1) return NULL;
    if (atten_db < 20.0) atten_db = 20.0;
    if (atkgpu_init() != 0) return NULL;
    atkdsp_osx *x = (atkdsp_osx *)calloc(1, sizeof *x);
    if (!x) return NULL;
    x->fs = fs;
    x->atten = atten_db;
    x->nfft = nfft;
    x->max_nel = max_nel;
    x->ov = nfft / 2;
    x->hop = nfft - x->ov;
    /* a sub-call takes a whole number of hops */
    size_t slab = OSX_SLAB_TOTAL / max_nel;
    if (slab < OSX_SLAB_MIN) slab = OSX_SLAB_MIN;
    x->slab = ((slab + x->hop - 1) / x->hop) * x->hop;
    x->mcap = x->slab / x->hop + 1;
    x->spitch = x->ov + x->hop + x->slab;
    const size_t nx = (size_t)max_nel * x->mcap * nfft;
    int ok = !atkgpu_stream_create(&x->stream) && (x->plan = atkgpu_fftplan_create(nfft)) != NULL;
    if (ok) {
        x->d_s[0] = (atkdsp_cf32 *)atkgpu_dmalloc((size_t)max_nel * x->spitch * sizeof(atkdsp_cf32));
        x->d_s[1] = (atkdsp_cf32 *)atkgpu_dmalloc((size_t)max_nel * x->spitch * sizeof(atkdsp_cf32));
        x->d_X = (atkdsp_cf32 *)atkgpu_dmalloc(nx * sizeof(atkdsp_cf32));
        const size_t scr = atkgpu_fftplan_scratch(x->plan, x->mcap);
        x->d_scr = scr ? (atkdsp_cf32 *)atkgpu_dmalloc(scr * sizeof(atkdsp_cf32)) : NULL;
        x->lo = (atkdsp_cf32 *)calloc((size_t)max_nel * x->hop, sizeof(atkdsp_cf32));
        ok = x->d_s[0] && x->d_s[1] && x->d_X && (!scr || x->d_scr) && x->lo;
    }
    x->desc_cap = 64;
    for (int i = 0; ok && i < 2; ++i)
        ok = !atkgpu_event_create_order(&x->desc_ev[i]) &&
             (x->desc_h[i] = (atkgpu_osx_chan *)atkgpu_hmalloc_mapped(x->desc_cap * sizeof(atkgpu_osx_chan),
                                                                      &x->desc_d[i])) != NULL;
    /* the history starts at zero (the reference callocs it) */
    if (ok) ok = !atkgpu_fill_u32(x->stream, x->d_s[0], 0, (size_t)max_nel * x->spitch * 2) && !atkgpu_sync(x->stream);
    if (!ok) { atkdsp_osx_destroy(x); return NULL; }
    return x;
}

void atkdsp_osx_reset(atkdsp_osx *x) {
    if (!x) return;
    atkgpu_fill_u32(x->stream, x->d_s[x->cur], 0, (size_t)x->max_nel * x->spitch * 2);
    memset(x->lo, 0, (size_t)x->max_nel * x->hop * sizeof(atkdsp_cf32));
    x->lc = 0;
    x->blk = 0;
    for (int i = 0; i < x->nslots; ++i)
        if (x->chans[i].active && x->chans[i].rs_on) rs_reset(&x->chans[i].rs, x->stream);
    atkgpu_sync(x->stream);
}

/* the largest power-of-two decimation keeping fs/decim >= max(out_rate,
 * 1.25 bw) and dividing ov and nfft (src/channelizer.c:560-570) */
static unsigned osx_pick_decim(const atkdsp_osx *x, double bw, double out_rate) {
    const double target = (out_rate > 1.25 * bw) ? out_rate : 1.25 * bw;
    unsigned decim = 1;
    for (;;) {
        const unsigned nd = decim * 2;
        if (x->fs / (double)nd < target) break;
        if (x->ov % nd != 0 || x->nfft % nd != 0) break;
        decim = nd;
    }
    return decim;
}

/* the integer bin the roll moves by, and the residual in Hz (src/channelizer.c:573-579) */
static void osx_tune(const atkdsp_osx *x, osx_chan *c, double offset_hz) {
    const long s = (long)floor(offset_hz / x->fs * (double)x->nfft + 0.5);
    long sh = s % (long)x->nfft;
    if (sh < 0) sh += (long)x->nfft;
    c->offset = offset_hz;
    c->shift = (unsigned)sh;
    c->delta = offset_hz - (double)s * x->fs / (double)x->nfft;
}

/* forward DFT of n (a power of two) points in FP64, radix-2 in place; the
 * twiddles come from cos/sin of exact multiples of 2 pi/n */
static void dft_dbl(double *re, double *im, size_t n) {
    for (size_t i = 1, j = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        const size_t half = len / 2, stride = n / len;
        for (size_t i = 0; i < n; i += len)
            for (size_t k = 0; k < half; ++k) {
                const double a = -2.0 * ATK_PI * (double)(k * stride) / (double)n;
                const double wr = cos(a), wi = sin(a);
                const size_t u = i + k, v = i + k + half;
                const double xr = re[v] * wr - im[v] * wi, xi = re[v] * wi + im[v] * wr;
                re[v] = re[u] - xr; im[v] = im[u] - xi;
                re[u] += xr; im[u] += xi;
            }
    }
}

int atkdsp_osx_add(atkdsp_osx *x, double offset_hz, double bw_hz, double out_rate) {
    if (!x || !(bw_hz > 0.0) || !(out_rate > 0.0)) return ATKDSP_E_ARG;
    const unsigned decim = osx_pick_decim(x, bw_hz, out_rate);
    const double chan_rate = x->fs / (double)decim;
    /* the channel's lowpass: passband bw/2, stopband chan_rate - bw/2, at most
     * ov + 1 taps (renormalised to unit DC gain when capped) */
    const double fp = 0.5 * bw_hz;
    double fstop = chan_rate - 0.5 * bw_hz;
    if (fstop > 0.5 * x->fs) fstop = 0.5 * x->fs;
    if (fstop <= fp) return ATKDSP_E_ARG;
    const ptrdiff_t nt = atkdsp_design_lowpass(fp, fstop, x->atten, x->fs, NULL, 0);
    if (nt <= 0) return ATKDSP_E_ARG;
    float *taps = (float *)malloc((size_t)nt * sizeof(float));
    double *hr = (double *)calloc(x->nfft, sizeof(double));
    double *hi = (double *)calloc(x->nfft, sizeof(double));
    atkdsp_cf32 *H = (atkdsp_cf32 *)malloc(x->nfft * sizeof(atkdsp_cf32));
    if (!taps || !hr || !hi || !H) { free(taps); free(hr); free(hi); free(H); return ATKDSP_E_NOMEM; }
    atkdsp_design_lowpass(fp, fstop, x->atten, x->fs, taps, (size_t)nt);
    size_t ntaps = (size_t)nt;
    if (ntaps > x->ov + 1) {
        ntaps = x->ov + 1;
        double s = 0.0;
        for (size_t i = 0; i < ntaps; ++i) s += taps[i];
        if (s == 0.0) s = 1.0;
        for (size_t i = 0; i < ntaps; ++i) taps[i] = (float)(taps[i] / s);
    }
    /* H = FFT of the taps zero-padded to nfft, in FP64, rounded once to float */
    for (size_t i = 0; i < ntaps; ++i) hr[i] = (double)taps[i];
    dft_dbl(hr, hi, x->nfft);
    for (size_t k = 0; k < x->nfft; ++k) {
        H[k].re = (float)hr[k];
        H[k].im = (float)hi[k];
    }
    free(taps);
    free(hr);
    free(hi);
    /* a free slot, or a new one */
    int slot = -1;
    for (int i = 0; i < x->nslots; ++i)
        if (!x->chans[i].active) { slot = i; break; }
    if (slot < 0) {
        if (x->nslots >= x->cap_slots) {
            const int ncap = x->cap_slots ? x->cap_slots * 2 : 8;
            osx_chan *nc = (osx_chan *)realloc(x->chans, (size_t)ncap * sizeof(osx_chan));
            if (!nc) { free(H); return ATKDSP_E_NOMEM; }
            memset(nc + x->cap_slots, 0, (size_t)(ncap - x->cap_slots) * sizeof(osx_chan));
            x->chans = nc;
            x->cap_slots = ncap;
        }
        slot = x->nslots++;
    }
    osx_chan *c = &x->chans[slot];
    memset(c, 0, sizeof *c);
    c->active = 1;
    c->bw = bw_hz;
    c->out_rate = out_rate;
    c->decim = decim;
    c->chan_rate = chan_rate;
    c->ntaps = ntaps;
    c->M = x->nfft / decim;
    c->ovd = x->ov / decim;
    c->nvalid = c->M - c->ovd;
    osx_tune(x, c, offset_hz);
    int rc = 0;
    c->d_H = (atkdsp_cf32 *)atkgpu_dmalloc(x->nfft * sizeof(atkdsp_cf32));
    c->iplan = atkgpu_fftplan_create(c->M);
    if (!c->d_H || !c->iplan) rc = ATKDSP_E_NOMEM;
    if (!rc) {
        const size_t scr = atkgpu_fftplan_scratch(c->iplan, (size_t)x->max_nel * x->mcap);
        if (scr && !(c->d_iscr = (atkdsp_cf32 *)atkgpu_dmalloc(scr * sizeof(atkdsp_cf32)))) rc = ATKDSP_E_NOMEM;
        c->d_Y = (atkdsp_cf32 *)atkgpu_dmalloc((size_t)x->max_nel * x->mcap * c->M * sizeof(atkdsp_cf32));
        c->d_V = (atkdsp_cf32 *)atkgpu_dmalloc((size_t)x->max_nel * x->mcap * c->nvalid * sizeof(atkdsp_cf32));
        if (!c->d_Y || !c->d_V) rc = ATKDSP_E_NOMEM;
    }
    /* room for every slot's descriptor (grown here, a configuration call) */
    if (!rc && x->nslots > x->desc_cap) {
        int cap = x->desc_cap;
        while (cap < x->nslots) cap *= 2;
        if (atkgpu_sync(x->stream)) rc = ATKDSP_E_FFT;
        for (int i = 0; i < 2 && !rc; ++i) {
            void *dd = NULL;
            atkgpu_osx_chan *h = (atkgpu_osx_chan *)atkgpu_hmalloc_mapped(cap * sizeof(atkgpu_osx_chan), &dd);
            if (!h) { rc = ATKDSP_E_NOMEM; break; }
            atkgpu_hfree(x->desc_h[i]);
            x->desc_h[i] = h;
            x->desc_d[i] = dd;
            x->desc_used[i] = 0;
        }
        if (!rc) x->desc_cap = cap;
    }
    if (!rc) rc = atkgpu_h2d(x->stream, c->d_H, H, x->nfft * sizeof(atkdsp_cf32)) ? ATKDSP_E_FFT : 0;
    if (!rc && fabs(chan_rate - out_rate) > 1e-6 * out_rate) {
        c->rs_on = 1;
        rc = rs_init(&c->rs, chan_rate, out_rate, x->atten, 64, x->max_nel, x->stream);
        if (rc) c->rs_on = 0;
        if (!rc) {
            c->ocap = (size_t)((double)(x->mcap * c->nvalid) * out_rate / chan_rate) + 64;
            c->ocap = (c->ocap + 63) & ~(size_t)63;
            c->d_out = (atkdsp_cf32 *)atkgpu_dmalloc((size_t)x->max_nel * c->ocap * sizeof(atkdsp_cf32));
            if (!c->d_out) rc = ATKDSP_E_NOMEM;
        }
    }
    /* the upload of H (and the resampler's set-up) finish before H is freed */
    if (atkgpu_sync(x->stream) && !rc) rc = ATKDSP_E_FFT;
    free(H);
    if (rc) {
        osx_free_chan(c);
        return rc;
    }
    return slot;
}

int atkdsp_osx_retune(atkdsp_osx *x, int ch, double offset_hz) {
    if (!x || ch < 0 || ch >= x->nslots || !x->chans[ch].active) return ATKDSP_E_ARG;
    osx_tune(x, &x->chans[ch], offset_hz);
    return ATKDSP_OK;
}

int atkdsp_osx_remove(atkdsp_osx *x, int ch) {
    if (!x || ch < 0 || ch >= x->nslots || !x->chans[ch].active) return ATKDSP_E_ARG;
    atkgpu_sync(x->stream);
    osx_free_chan(&x->chans[ch]);
    return ATKDSP_OK;
}

int atkdsp_osx_channels(const atkdsp_osx *x) { return x ? x->nslots : 0; }

double atkdsp_osx_chan_rate(const atkdsp_osx *x, int ch) {
    if (!x || ch < 0 || ch >= x->nslots || !x->chans[ch].active) return 0.0;
    return x->chans[ch].out_rate;
}

size_t atkdsp_osx_out_max(const atkdsp_osx *x, int ch, size_t n) {
    if (!x || ch < 0 || ch >= x->nslots || !x->chans[ch].active) return 0;
    return (size_t)((double)n * x->chans[ch].out_rate / x->fs) + 64;
}

/* one sub-call over ins[e][pos, pos+c): the forward FFT of every block, then
 * every channel's fold, mix and resampler in one launch each */
static int osx_sub(atkdsp_osx *x, const atkdsp_cf32 *const *ins, unsigned nel, size_t pos, size_t c,
                   atkdsp_cf32 *const *outs, size_t *out_counts) {
    const size_t hop = x->hop, ov = x->ov, nfft = x->nfft;
    const size_t avail = (size_t)x->lc + c;
    const size_t ms = avail / hop;
    const int ns = x->nslots;
    if (ms == 0) {
        for (unsigned e = 0; e < nel; ++e)
            if (c) memcpy(x->lo + (size_t)e * hop + x->lc, ins[e] + pos, c * sizeof(atkdsp_cf32));
        x->lc = (unsigned)avail;
        return 0;
    }
    atkdsp_cf32 *S = x->d_s[x->cur];
    int rc = 0;
    for (unsigned e = 0; e < nel && !rc; ++e) {
        atkdsp_cf32 *se = S + (size_t)e * x->spitch;
        if (x->lc) rc = atkgpu_h2d(x->stream, se + ov, x->lo + (size_t)e * hop, x->lc * sizeof(atkdsp_cf32));
        if (!rc && c) rc = atk_put(x->stream, se + ov + x->lc, ins[e] + pos, c * sizeof(atkdsp_cf32));
        /* every block of the element: nfft samples, hop apart */
        if (!rc) rc = atkgpu_fft_batch(x->stream, x->plan, se, hop, NULL, ms, 0, x->d_X + (size_t)e * x->mcap * nfft,
                                       x->d_scr);
    }
    /* this sub-call's channel descriptors, in the set not in flight */
    const int b = x->dcur;
    x->dcur ^= 1;
    if (!rc && x->desc_used[b] && atkgpu_event_sync(x->desc_ev[b])) rc = ATKDSP_E_FFT;
    atkgpu_osx_chan *dh = x->desc_h[b];
    int nd = 0, max_nvalid = 0, max_count = 0, max_q = 0;
    size_t got[x->nslots > 0 ? x->nslots : 1];
    int slot_of[x->nslots > 0 ? x->nslots : 1];
    for (int ci = 0; ci < ns && !rc; ++ci) {
        osx_chan *ch = &x->chans[ci];
        if (!ch->active) continue;
        atkgpu_osx_chan *d = &dh[nd];
        memset(d, 0, sizeof *d);
        d->H = ch->d_H;
        d->Y = ch->d_Y;
        d->V = ch->d_V;
        d->delta = ch->delta;
        d->M = (int)ch->M;
        d->decim = (int)ch->decim;
        d->ovd = (int)ch->ovd;
        d->nvalid = (int)ch->nvalid;
        d->sh = (int)ch->shift;
        d->vpitch = (long long)(x->mcap * ch->nvalid);
        const size_t nv = ms * ch->nvalid;
        got[nd] = nv;
        if (ch->rs_on) {
            const size_t k = rs_count(&ch->rs, nv);
            if (k > ch->ocap) { rc = ATKDSP_E_FFT; break; }
            d->rs = 1;
            d->out = ch->d_out;
            d->ocap = (long long)ch->ocap;
            d->p = ch->rs.d_p;
            d->d = ch->rs.d_d;
            d->hist_in = ch->rs.d_hist[ch->rs.hcur];
            d->hist_out = ch->rs.d_hist[ch->rs.hcur ^ 1];
            d->base = ch->rs.base;
            d->step = ch->rs.step;
            d->inoff = (double)ch->rs.in_off;
            d->K0 = (unsigned long long)ch->rs.K;
            d->P = (int)ch->rs.P;
            d->Q = (int)ch->rs.Q;
            d->count = (int)k;
            got[nd] = k;
            if ((int)k > max_count) max_count = (int)k;
            if ((int)ch->rs.Q > max_q) max_q = (int)ch->rs.Q;
            /* the position and the history advance with this sub-call */
            ch->rs.hcur ^= 1;
            ch->rs.K += k;
            ch->rs.in_off += nv;
        }
        if (d->nvalid > max_nvalid) max_nvalid = d->nvalid;
        slot_of[nd] = ci;
        ++nd;
    }
    const void *dd = x->desc_d[b];
    if (!rc && nd) rc = atkgpu_osx_fold_all(x->stream, x->d_X, dd, nd, nel, ms, x->mcap * nfft, (unsigned)nfft);
    for (int i = 0; i < nd && !rc; ++i) {
        osx_chan *ch = &x->chans[slot_of[i]];
        rc = atkgpu_fft_batch(x->stream, ch->iplan, ch->d_Y, ch->M, NULL, (size_t)nel * ms, 1, ch->d_Y, ch->d_iscr);
    }
    if (!rc && nd)
        rc = atkgpu_osx_mix_all(x->stream, dd, nd, nel, ms, max_nvalid, (unsigned)hop, (unsigned)ov, (unsigned)nfft,
                                (long long)x->blk, x->fs);
    if (!rc && nd && max_q) rc = atkgpu_osx_rs_all(x->stream, dd, nd, nel, ms, max_count, max_q);
    if (!rc) rc = atkgpu_event_record(x->desc_ev[b], x->stream);
    x->desc_used[b] = 1;
    for (int i = 0; i < nd && !rc; ++i) {
        const int ci = slot_of[i];
        const osx_chan *ch = &x->chans[ci];
        const atkdsp_cf32 *src = ch->rs_on ? ch->d_out : ch->d_V;
        const size_t spitch = ch->rs_on ? ch->ocap : x->mcap * ch->nvalid;
        for (unsigned e = 0; e < nel && !rc && got[i]; ++e)
            rc = atkgpu_d2h(x->stream, outs[(size_t)e * ns + ci] + out_counts[ci], src + (size_t)e * spitch,
                            got[i] * sizeof(atkdsp_cf32));
        out_counts[ci] += got[i];
    }
    /* next history: s[ms hop, ms hop + ov) into the other row */
    atkdsp_cf32 *T = x->d_s[x->cur ^ 1];
    for (unsigned e = 0; e < nel && !rc; ++e)
        rc = atkgpu_d2d(x->stream, T + (size_t)e * x->spitch, S + (size_t)e * x->spitch + ms * hop,
                        ov * sizeof(atkdsp_cf32));
    if (rc) return rc < 0 && rc != -4 ? rc : ATKDSP_E_FFT;
    /* next leftover: the tail of this sub-call's input */
    const size_t rem = avail - ms * hop;
    const size_t from = ms * hop - x->lc;
    for (unsigned e = 0; e < nel; ++e)
        if (rem) memcpy(x->lo + (size_t)e * hop, ins[e] + pos + from, rem * sizeof(atkdsp_cf32));
    x->cur ^= 1;
    x->lc = (unsigned)rem;
    x->blk += (int64_t)ms;
    return 0;
}

ptrdiff_t atkdsp_osx_process_n(atkdsp_osx *x, const atkdsp_cf32 *const *ins, unsigned nel, size_t n,
                               atkdsp_cf32 *const *outs, size_t *out_counts, size_t cap_each) {
    if (!x || !ins || nel < 1 || nel > x->max_nel || !out_counts) return ATKDSP_E_ARG;
    const int ns = x->nslots;
    for (int ci = 0; ci < ns; ++ci) out_counts[ci] = 0;
    for (unsigned e = 0; e < nel; ++e)
        if (n && !ins[e]) return ATKDSP_E_ARG;
    const size_t m = ((size_t)x->lc + n) / x->hop;
    if (m) {
        /* every active channel needs its outputs, and they must fit */
        for (int ci = 0; ci < ns; ++ci) {
            const osx_chan *ch = &x->chans[ci];
            if (!ch->active) continue;
            if (!outs) return ATKDSP_E_ARG;
            for (unsigned e = 0; e < nel; ++e)
                if (!outs[(size_t)e * ns + ci]) return ATKDSP_E_ARG;
            const size_t nv = m * ch->nvalid;
            const size_t total = ch->rs_on ? rs_count(&ch->rs, nv) : nv;
            if (total > cap_each) return ATKDSP_E_CAP;
        }
    }
    for (size_t pos = 0; pos < n || (n == 0 && pos == 0);) {
        const size_t c = n - pos < x->slab ? n - pos : x->slab;
        const int rc = osx_sub(x, ins, nel, pos, c, outs, out_counts);
        if (rc) return rc;
        pos += c;
        if (n == 0) break;
    }
    if (atkgpu_sync(x->stream)) return ATKDSP_E_FFT;
    return (ptrdiff_t)ns;
}

ptrdiff_t atkdsp_osx_process(atkdsp_osx *x, const atkdsp_cf32 *in, size_t n, atkdsp_cf32 *const *outs,
                             size_t *out_counts, size_t cap_each) {
    const atkdsp_cf32 *ins[1] = {in};
    return atkdsp_osx_process_n(x, ins, 1, n, outs, out_counts, cap_each);
}

int atkdsp_osx_describe(const atkdsp_osx *x, char *buf, size_t cap) {
    if (!x || !buf || cap == 0) return ATKDSP_E_ARG;
    int active = 0;
    for (int i = 0; i < x->nslots; ++i)
        if (x->chans[i].active) ++active;
    return snprintf(buf, cap, "OSX %.6g MSPS: %zu-pt FFT, ov %zu, %d channel(s), %.0f dB, <=%u el; hip gfx950",
                    x->fs / 1e6, x->nfft, x->ov, active, x->atten, x->max_nel);
}
