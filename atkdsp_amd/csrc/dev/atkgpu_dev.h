/*
 * atkgpu_dev.h — the thin C ABI between the C host layer (atkdsp_* entry
 * points, src in ../host) and the HIP kernels for gfx950 (this directory).
 *
 * Plain C: device pointers are void*, streams and plans are opaque. Every
 * launcher returns 0 or a negative ATKDSP_E_* code (E_FFT for a failed HIP
 * call). Nothing here allocates except the *_create / alloc functions, which
 * the host layer only calls from its own *_create paths or a one-time init.
 */
#ifndef ATKGPU_DEV_H
#define ATKGPU_DEV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- runtime ------------------------------------------------------------ */
int         atkgpu_init(void);                 /* 0, or E_FFT when no device */
const char *atkgpu_device_name(void);
void       *atkgpu_dmalloc(size_t bytes);      /* device memory, NULL on failure */
void        atkgpu_dfree(void *p);
void       *atkgpu_hmalloc(size_t bytes);      /* pinned host memory */
void        atkgpu_hfree(void *p);
int         atkgpu_stream_create(void **s);    /* default priority */
void        atkgpu_stream_destroy(void *s);
int         atkgpu_sync(void *s);
int         atkgpu_h2d(void *s, void *dst, const void *src, size_t bytes);
int         atkgpu_d2h(void *s, void *dst, const void *src, size_t bytes);
int         atkgpu_d2d(void *s, void *dst, const void *src, size_t bytes);
int         atkgpu_d2h_2d(void *s, void *dst, size_t dpitch, const void *src,
                          size_t spitch, size_t width, size_t height);
int         atkgpu_d2d_2d(void *s, void *dst, size_t dpitch, const void *src, size_t spitch,
                          size_t width, size_t height);
int         atkgpu_fill_u32(void *s, void *dst, uint32_t v, size_t count);
int         atkgpu_event_create(void **e);
void        atkgpu_event_destroy(void *e);
int         atkgpu_event_record(void *e, void *s);
int         atkgpu_event_elapsed_ms(void *e0, void *e1, float *ms);
int         atkgpu_event_sync(void *e);
int         atkgpu_event_query(void *e);           /* 1 done, 0 pending, <0 error */
int         atkgpu_event_create_order(void **e);   /* no timing: for stream ordering */
int         atkgpu_stream_wait_event(void *s, void *e);
int         atkgpu_d2h_sync(void *dst, const void *src, size_t bytes);
int         atkgpu_d2h_2d_sync(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width,
                               size_t height);

/* ---- unpack --------------------------------------------------------------- */
/* number of double pairs in the mean partial buffer */
#define ATKGPU_MEAN_PARTS 1024
/* samples one pass of the pooled entry points moves */
#define ATKGPU_POOL_CHUNK ((size_t)1 << 22)
/* unpack forms its mean as FP64 sums over passes of this many samples (one
 * launch each), added in pass order */
#define ATKGPU_MEAN_PASS ((size_t)1 << 19)

/* The folded constants of a format, formed in float exactly as
 * src/convert.c:46-68 does (formats by the ABI's values: 0 cu8, 1 ci8,
 * 2 ci16, 3 ci16q11, 4 cf32). */
static inline void atkgpu_unpack_consts(int fmt, float off_re, float off_im, float *cre, float *cim, float *k) {
    switch (fmt) {
    case 0:
        *k = 1.0f / 127.5f;
        *cre = 127.5f + off_re / *k;
        *cim = 127.5f + off_im / *k;
        break;
    case 1:
        *k = 1.0f / 128.0f;
        *cre = off_re / *k;
        *cim = off_im / *k;
        break;
    case 2:
    case 3:
        *k = fmt == 2 ? (1.0f / 32768.0f) : (1.0f / 2048.0f);
        *cre = off_re / *k;
        *cim = off_im / *k;
        break;
    default:
        *k = 1.0f;
        *cre = 0.0f;
        *cim = 0.0f;
        break;
    }
}
/* out[i] = ((float)raw - c) * k per component (cf32 input: x - off).
 * When sum2 (2 doubles, device) is non-NULL it receives the FP64 sums of out,
 * formed through partials (2*ATKGPU_MEAN_PARTS doubles) in a fixed order. */
int atkgpu_unpack(void *s, const void *raw, int fmt, size_t n, float cre, float cim,
                  float k, float off_re, float off_im, void *out,
                  double *partials, double *sum2);
/* one-pole DC block in place over cf32 x[n]: st <- st + a(x - st); x -= st.
 * state (device, 2 floats) in/out. scratch: >= 4*(n/256+1)+16 doubles. */
int atkgpu_dc_pole(void *s, void *x, size_t n, float alpha, float *state, double *scratch);

/* ---- FFT plans -------------------------------------------------------------- */
typedef struct atkgpu_fftplan atkgpu_fftplan;
atkgpu_fftplan *atkgpu_fftplan_create(size_t n);
void            atkgpu_fftplan_destroy(atkgpu_fftplan *p);
/* device scratch (complex samples) one batch of `batch` transforms needs */
size_t          atkgpu_fftplan_scratch(const atkgpu_fftplan *p, size_t batch);
/* forward (inverse: conj trick and 1/n) transforms of `batch` inputs
 * in + t*in_dist (contiguous, optionally times window) into out + t*n.
 * in and out may alias only when in_dist == n and window == NULL. */
int atkgpu_fft_batch(void *s, const atkgpu_fftplan *p, const void *in, size_t in_dist,
                     const float *window, size_t batch, int inverse, void *out,
                     void *scratch);

/* ---- spectrum lines ----------------------------------------------------------- */
/* acc planes (device): reduce uses float line[n]; stats uses float max, min and
 * double sum, sum2 (see spectrum.hip). first != 0 starts a fresh accumulation.
 * part: device scratch of the frame groups' partials, at least
 * atkgpu_spec_part_entries(n) doubles. */
size_t atkgpu_spec_part_entries(size_t n);
int atkgpu_spec_accum(void *s, const void *X, size_t n, size_t frames, int detector,
                      int first, float *line, void *part);
int atkgpu_spec_reduce_db(void *s, const float *line, size_t n, int detector,
                          long long frames, float *out_db);
int atkgpu_spec_stats_accum(void *s, const void *X, size_t n, size_t frames, int first,
                            float *mx, float *mn, double *su, double *s2, void *part);
int atkgpu_spec_stats_db(void *s, size_t n, long long frames, const float *mx,
                         const float *mn, const double *su, const double *s2,
                         float *max_db, float *avg_db, float *min_db, float *sk);
int atkgpu_power_db(void *s, const void *X, size_t n, float *out_db);

/* ---- display ------------------------------------------------------------------ */
/* detection on spectrum lines (reference src/detect.c:4-114); a record is
 * atkdsp_channel's layout */
typedef struct {
    int start_bin, end_bin;
    float centroid_bin, peak_db, snr_db;
} atkgpu_chan;
/* nlines lines of n bins, line l at lines + l*lstride; scratch n floats per
 * line; out cap records per line (bin order), found[l] the runs found,
 * floors[l] the floor used (the median of the smoothed line when
 * compute_floor, else floor_db); gap >= 2, min_run >= 2 */
int atkgpu_detect_lines(void *s, const float *lines, size_t lstride, size_t n, unsigned nlines, float floor_db,
                        int compute_floor, float threshold_db, int gap, int min_run, float *scratch,
                        atkgpu_chan *out, size_t cap, unsigned long long *found, float *floors);
/* nlines lines of n <= 4096 bins at ptrs[l] (a device-visible table), each
 * with its floor computed (the median of its smoothed line): one launch */
int atkgpu_detect_line_set(void *s, const float *const *ptrs, size_t n, unsigned nlines, float threshold_db, int gap,
                           int min_run, atkgpu_chan *out, size_t cap, unsigned long long *found, float *floors);
/* the median of x[0..n) into *out (device) */
int atkgpu_median(void *s, const float *x, size_t n, float *out);
int atkgpu_decimate_max(void *s, const float *in, size_t n, float *out, size_t m);
int atkgpu_db_to_pixels(void *s, const float *db, size_t n, float lo, float hi,
                        const uint32_t *lut, uint32_t *out);
int atkgpu_stitch_max(void *s, const float *seg, size_t nseg, double seg_lo,
                      double seg_step, double keep_lo, double keep_hi, double out_lo,
                      double out_step, float *out, size_t nout);

/* ---- polyphase filterbank -------------------------------------------------------- */
/* fold: for hops i in [0, m) of element e, stream s_e = sbuf + e*spitch
 * (hist ++ leftover ++ input, already contiguous on device):
 *   fold[(e*m + i)*M + q] = sum_k h[kM + q] * s_e[i*hop + kM + q], k ascending (FP32) */
int atkgpu_pfb_fold(void *s, const void *sbuf, size_t spitch, unsigned nel, size_t m,
                    unsigned M, unsigned hop, unsigned K, const float *taps, void *fold);
/* rotation exp(-2 pi i k ((b0+i+1) hop mod M)/M), cal[e*M+k] for e < cal_nel
 * (FP32 complex), then out[(e*M + k)*m + i] (channel-major slab). */
int atkgpu_pfb_rotate_store(void *s, const void *X, unsigned nel, size_t m, unsigned M,
                            unsigned hop, long long b0, const void *twM, const void *cal,
                            unsigned cal_nel, void *out);
/* The fused kernel (fold, FFT, rotation, cal and the channel-major store in
 * one launch) for M in 64..2048 (power of two), os 1 or 2, K*os <= 32.
 * atkgpu_pfb_fused_kos returns the padded tap-row count (8/16/24/32) the
 * taps buffer must hold (KOS*hop floats, zero beyond K*os), or 0 when the
 * shape needs the general path. Element e streams from sbuf + e*spitch with
 * nvalid valid samples; hop i of element e, channel k goes to
 * out[(e*M + k)*opitch + ocol + i]. */
int atkgpu_pfb_fused_kos(unsigned M, unsigned os, unsigned K);
int atkgpu_pfb_fused(void *s, const void *sbuf, size_t spitch, size_t nvalid, unsigned nel,
                     size_t m, long long b0, unsigned M, unsigned os, unsigned kos,
                     const float *taps, const void *tw, const void *cal, unsigned cal_nel,
                     void *out, size_t opitch, size_t ocol);
/* The same kernel over raw rows (element e at raw + e*spitch samples of the
 * format's size): each sample is converted as it is loaded, bit for bit the
 * unpack with the fixed offset (off_re, off_im); window samples before zlo
 * are the stream's zero history. For M 64 or 128, os 1 or 2, kos 8/16/32 and
 * the cu8, ci8, ci16 and ci16q11 formats (atkgpu_pfb_fused_raw_supported). */
int atkgpu_pfb_fused_raw_supported(unsigned M, unsigned os, unsigned kos, int fmt);
int atkgpu_pfb_fused_raw(void *s, const void *raw, size_t spitch, size_t nvalid, unsigned nel, size_t m,
                         long long b0, unsigned M, unsigned os, unsigned kos, int fmt, float off_re,
                         float off_im, int zlo, const float *taps, const void *tw, void *out, size_t opitch);
/* ---- chain kernels (chain.hip): S stream rows spitch samples apart ---- */
/* raw (stream s at raw + s*raw_pitch bytes) -> cf32 at row0 + s*spitch, with the
 * per-stream offset off[2s..2s+1] (NULL: none); mean[2s..] = block mean before
 * the offset; update != 0 makes it the next block's offset. partials: S*parts*2. */
int atkgpu_chain_unpack(void *s, const void *raw, size_t raw_pitch, int fmt, size_t n, void *row0,
                        size_t spitch, float *off, double *partials, int parts, double *mean,
                        int update, unsigned S);
/* the chain's display reduction straight from the raw blocks (frames >= 1,
 * N in 1024/2048/4096; every frame f at row0 + s*spitch + f*fhop: window,
 * FFT, |X|^2/N^2, detector -> partial[S][nfb][N] -> line_db[S][N]),
 * unpacking each block into its row on the way: row0 is stream 0's frame 0,
 * the block starts dlen samples after it; then the dB lines and the mean of
 * every block (FP64 sums in mparts, [S][64][2]) into mean and, with update,
 * the next block's offset */
int atkgpu_chain_display_unpack(void *st, void *row0, size_t spitch, const void *raw, size_t raw_pitch, int fmt,
                                size_t n, size_t dlen, float *off, unsigned S, size_t frames, size_t fhop, unsigned N,
                                int nfb, const float *win, const void *tw, int det, float *partial, float *line_db,
                                double *mparts, double *mean, int update);
/* decimate line_db[S][N] to dec[S][cols]; stitch into trace[S][nout] (NULL: none)
 * with stream s's segment starting at seg_lo0 + s*seg_lo_step */
int atkgpu_chain_post(void *s, const float *line_db, unsigned S, unsigned N, unsigned cols, float *dec,
                      float *trace, unsigned nout, double seg_lo0, double seg_lo_step, double seg_step,
                      double out_lo, double out_step);
/* copy keep samples from src + s*spitch to dst + s*spitch for every stream */
int atkgpu_chain_keep(void *s, const void *src, size_t spitch, void *dst, size_t keep, unsigned S);

/* ---- ragged batches of independent streams (serve.hip) ----
 * Job arrays live in device-visible memory (mapped pinned host memory is
 * fine). Filterbank jobs are per block: entry b is the element block b
 * works on, with the block's tile in blk0 and the element's tile count in
 * nblk. Unpack jobs are per job (job j owns launch blocks [blk0, blk0 + nblk))
 * or, with per_block set, per block (entry b's blk0 is block b's tile). */
typedef struct {               /* layout of PfbJob in pfb_fused.hpp */
    const void *s;             /* stream rows (cf32), hop i's window at s + i*hop */
    void *out;                 /* channel k, hop i at out[k*opitch + i] */
    const float *taps;         /* kos*hop floats, zero-padded */
    long long b0;              /* global hop index of hop 0 */
    int nvalid, m, opitch, blk0, nblk, pad_;
} atkgpu_pfb_job;
typedef struct {
    const void *raw;           /* device-visible raw bytes, 16-byte aligned */
    void *dst;                 /* cf32 out, 16-byte aligned */
    float off_re, off_im;      /* offset folded into the constants */
    int n, blk0, nblk, fmt;    /* fmt: ATKDSP_FMT_* of this job's bytes, or ATKGPU_JOB_COPY */
} atkgpu_unpack_job;
#define ATKGPU_JOB_COPY 255        /* an unpack-launch job that copies n cf32 samples raw -> dst */
typedef struct {
    const void *src;
    void *dst;
    int n, blk0, nblk, pad_;
} atkgpu_copy_job;
#define ATKGPU_RAGGED_SPB 2048     /* samples a block of the ragged unpack/copy handles */
/* served display, two launches per (FFT length, detector) class:
 *   frames: each FFT team of a block transforms one frame (window, N-point
 *           FFT, |X|^2/n^2 in float) and writes its power spectrum in display
 *           order to P + t*N for the block's t-th frame;
 *   lines:  one block per line segment of one element's chunk (its frames
 *           belong to one line): starting from the running line of the
 *           line's earlier frames (acc_in) or the first frame, the detector
 *           over the segment's spectra in frame order; then the running line
 *           for the next chunk (acc_out), or, when the segment ends its line,
 *           the line's dB decimated to cols columns (`line`). The running
 *           line is double buffered across chunks (read one buffer, write the
 *           other), so a chunk's segments are independent blocks. */
typedef struct {
    const void *first;         /* cf32: the block's first frame */
    const float *win;          /* N window values, or NULL */
    float *P;                  /* power spectra out, N floats per frame */
    int nfr, fhop;             /* frames (at most the block's teams), frame hop */
} atkgpu_frame_blk;
typedef struct {
    const float *P;            /* the segment's spectra, N floats per frame */
    const float *acc_in;       /* running line before the segment, or NULL */
    float *acc_out;            /* running line after it, or NULL when it ends a line */
    float *line;               /* the finished line's columns, or NULL */
    int nfr, cols;
    float scale;               /* 1/frames-per-line for the average detector, else 1 */
    int pad_;
} atkgpu_line_blk;
/* ---- the overlap-save extractor and the arbitrary resampler (osx.hip) ----
 * fold: for nt blocks t (forward transforms at X + t*nfft),
 *   Yf[t*M + k] = sum_j X[t][(k + j*M + sh) & (nfft-1)] * H[k + j*M], j < decim;
 * mix: valid[e*vpitch + i*(M-ovd) + k] = ys[(e*m+i)*M + ovd + k] / decim times
 *   exp(-2 pi i (((sh*start) mod nfft)/nfft + delta*nabs/fs)) in FP64, start =
 *   (blk0+i)*hop - ov, nabs = (blk0+i)*hop + k*decim;
 * rs_run: outputs K0 .. K0+count-1 of a call over n inputs per element
 *   (element e at in + e*ipitch), history hist_in ([nel][Q]) -> hist_out. */
int atkgpu_rs_run(void *s, const void *in, size_t ipitch, size_t n, unsigned nel, const void *hist_in,
                  void *hist_out, const float *p, const float *d, unsigned P, unsigned Q, double base,
                  double step, unsigned long long K0, unsigned long long in_off, void *out, size_t opitch,
                  size_t count);
/* every channel of one extractor sub-call in one launch per stage; the
 * descriptors (one per active channel) sit in device-visible memory */
typedef struct {
    const void *H;             /* nfft: the channel filter's frequency response */
    void *Y;                   /* folds, then their inverse transforms: row t = e*m + i at Y + t*M */
    void *V;                   /* pre-resampler samples: element e, block i at V + e*vpitch + i*nvalid */
    void *out;                 /* resampled outputs: element e at out + e*ocap */
    const float *p, *d;        /* resampler prototype and differences (rs) */
    const void *hist_in;       /* [nel][Q] history before the sub-call (rs) */
    void *hist_out;            /* [nel][Q] history after it (rs) */
    double delta, base, step, inoff;
    unsigned long long K0;
    int M, decim, ovd, nvalid, sh, rs, P, Q, count, pad_;
    long long vpitch, ocap;
} atkgpu_osx_chan;
int atkgpu_osx_fold_all(void *s, const void *X, const void *chans, int nch, unsigned nel, size_t m, size_t xpitch,
                        unsigned nfft);
int atkgpu_osx_mix_all(void *s, const void *chans, int nch, unsigned nel, size_t m, int max_nvalid, unsigned hop,
                       unsigned ov, unsigned nfft, long long blk0, double fs);
int atkgpu_osx_rs_all(void *s, const void *chans, int nch, unsigned nel, size_t nin, int max_count, int max_q);
int atkgpu_disp_teams(unsigned N);          /* frames one block of the frames launch transforms */
int atkgpu_disp_frames(void *s, const void *blks, int nblk, unsigned N, const void *tw);
int atkgpu_disp_lines(void *s, const void *blks, int nblk, unsigned N, int det);
int atkgpu_pfb_group(unsigned M, unsigned os);      /* hops per block of the fused kernel */
/* the matrix-core FP32 filterbank (pfb_mm.hpp): same arguments and output as
 * atkgpu_pfb_fused; atkgpu_pfb_mm_k gives the tap rows per branch it runs for
 * (M, os, kos), 0 when the shape is not covered */
int atkgpu_pfb_mm_k(unsigned M, unsigned os, unsigned kos);
int atkgpu_pfb_mm(void *s, const void *sbuf, size_t spitch, size_t nvalid, unsigned nel, size_t m,
                  long long b0, unsigned M, unsigned os, unsigned kos, const float *taps, const void *tw,
                  const void *cal, unsigned cal_nel, void *out, size_t opitch, size_t ocol);
int atkgpu_pfb_fused_jobs(void *s, const void *jobs, int njobs, int nblocks, unsigned M, unsigned os,
                          unsigned kos, const void *tw);
/* any mix of formats and prefix copies; max_grid > 0 caps the launch's blocks */
int atkgpu_unpack_jobs(void *s, const void *jobs, int njobs, int nblocks, int per_block, int max_grid);
int atkgpu_copy_jobs(void *s, const void *jobs, int njobs, int nblocks);
int atkgpu_copy_bytes_jobs(void *s, const void *jobs, int njobs, int nblocks);   /* n counts bytes, 4096 per block */

/* ---- the filterbank's matrix-core form (pfb_mx.hip) ----
 * raw rows of 8/12-bit samples -> fold on MFMA (exact BF16/FP16 samples,
 * hi+lo or single taps, FP32 accumulation) -> the fused FFT/rotation/store.
 * form: 1 BF16 hi+lo, 2 BF16, 3 FP16 hi+lo, 4 FP16 */
int    atkgpu_pfb_mx_supported(unsigned M, unsigned os, unsigned kos, int fmt, int form);
size_t atkgpu_pfb_mx_tpad_count(unsigned M, unsigned os, unsigned kos, int form);   /* 16-bit entries */
int    atkgpu_pfb_mx_tpad(void *s, const float *taps, unsigned M, unsigned os, unsigned kos, int form,
                          float tscale, void *tpad);
int    atkgpu_pfb_mx(void *s, const void *raw, size_t spitch, size_t nvalid, unsigned nel, size_t m, long long b0,
                     unsigned M, unsigned os, unsigned kos, int fmt, int form, const void *tpad,
                     const float *tapsum, const float *taps, int zlo, float scale, const void *offs,
                     const void *tw, const void *cal, unsigned cal_nel, void *out, size_t opitch);

/* ---- the NCO (nco.hip): FP64 phase, FP32 rotation ----
 * mix: samples base .. base+n-1 of a call whose phase is ph0 (any chunking
 * gives the same values); run: the phase lives in device memory,
 * st_in = {phase, step} -> st_out (and *logp when set) = the wrapped next phase */
int atkgpu_nco_mix(void *s, const void *in, void *out, size_t n, size_t base, double ph0, double step);
int atkgpu_nco_run_dev(void *s, const void *in, void *out, size_t n, const double *st_in, double *st_out,
                   double *logp);

/* pinned host memory the device can read (and its device address) */
void *atkgpu_hmalloc_mapped(size_t bytes, void **dev_ptr);

/* W1 / W2 over a device [nel][nchan][m] block */
int atkgpu_pfb_cov(void *s, const void *y, unsigned nel, unsigned nchan, size_t m, size_t ypitch,
                   const uint8_t *gate, void *R, uint32_t *counts);
int atkgpu_pfb_cov_reset(void *s, void *R, uint32_t *counts, unsigned nel, unsigned nchan,
                         const uint8_t *gate);
int atkgpu_pfb_power(void *s, const void *y, unsigned nel, unsigned nchan, size_t m, size_t ypitch,
                     float *pwr, float *mean_pwr);
/* the plan's twiddle table (n complex floats, device) */
const void *atkgpu_fftplan_twiddles(const atkgpu_fftplan *p);

/* ---- CFAR ------------------------------------------------------------------ */
/* raw[i] = 10^(max(db[i], -300) / 10) in FP64 */
int atkgpu_cfar_raw(void *s, const float *db, size_t n, double *raw);
/* the CFAR test of nframes frames of nbins bins; raw holds k-1+nframes rows,
 * the k-1 frames before the first; cells / ok / sided are the per-bin
 * reference-cell table (N per bin), reg flags the bins whose table is the
 * plain window at n_guard (left i-g-1-q, right i+g+1+q). Writes X (k > 1),
 * thr, noise (nframes x nbins) and hits. mode 0 CA, 1 OS, 2 GO, 3 SO;
 * N <= 512, n_guard + N/2 <= 512. */
/* out[2q], out[2q+1] = X[idx[q]], noise[idx[q]] for q < n (all device) */
int atkgpu_cfar_gather(void *s, const double *X, const double *noise, const unsigned *idx, size_t n, double *out);
int atkgpu_cfar_test(void *s, const double *raw, int nbins, int nframes, int k, int mode, int N, int os_k,
                     int n_guard, double alpha, double alpha_ca, double et, const int *cells, const unsigned char *ok,
                     const unsigned char *sided, const unsigned char *reg, double *X, double *thr, double *noise,
                     unsigned char *hits);

#ifdef __cplusplus
}
#endif
#endif
