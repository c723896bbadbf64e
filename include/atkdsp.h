/*
 * atkdsp.h — the wideband front-end subset of the atkdsp C ABI (ABI 26,
 * version "0.25.1"), implemented on one MI355X.
 *
 * Each declaration below replaces the reference function of the same name;
 * the citation gives the reference declaration (include/atkdsp.h:LINE) and
 * the reference implementation it follows (src/FILE:LINES). Signatures,
 * argument meaning, return codes and streaming state are the reference's, so
 * the reference ctypes binding (python/atkdsp/__init__.py) loads this library
 * unchanged when ATKDSP_LIB points at it.
 *
 * Contract kept from the reference header (include/atkdsp.h:5-41):
 *   - plain C calling convention, opaque handles, caller-owned host buffers;
 *   - nothing in a *_process / *_exec / reduce call allocates: device
 *     buffers, pinned staging and HIP streams are made in *_create;
 *   - streaming state is explicit and block boundaries are invisible: the
 *     filterbank's history lives in device memory and its hop counter is int64;
 *   - NaN means "unmeasured" (no fast-math anywhere).
 *
 * The library also exports every other name of the reference ABI (334 in
 * all; 88 are implemented here) as a documented stub: *_create returns NULL, signed counts return
 * ATKDSP_E_ARG, unsigned counts 0, floating results NaN, void functions do
 * nothing. Those are outside the wideband front end.
 */
#ifndef ATKDSP_H
#define ATKDSP_H

#include <stddef.h>
#include <stdint.h>

#define ATKDSP_API __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif

#define ATKDSP_ABI_VERSION 26          /* reference include/atkdsp.h:63 */
#define ATKDSP_VERSION_STRING "0.25.1" /* reference include/atkdsp.h:64 */

/* error codes: reference include/atkdsp.h:202-206 */
#define ATKDSP_OK       0
#define ATKDSP_E_ARG   -1
#define ATKDSP_E_NOMEM -2   /* also: device memory exhausted at create */
#define ATKDSP_E_CAP   -3
#define ATKDSP_E_FFT   -4   /* also: a HIP launch or copy failed       */
#define ATKDSP_E_OVERRUN -6 /* a ring reader was lapped: the data is gone */
#define ATKDSP_E_EMPTY -7   /* nothing to return yet */

/* complex float32, numpy complex64 layout: reference include/atkdsp.h:212 */
typedef struct { float re, im; } atkdsp_cf32;

/* reference include/atkdsp.h:215-235 */
enum atkdsp_format { ATKDSP_FMT_CU8 = 0, ATKDSP_FMT_CI8 = 1, ATKDSP_FMT_CI16 = 2,
                     ATKDSP_FMT_CI16Q11 = 3, ATKDSP_FMT_CF32 = 4 };
enum atkdsp_detector { ATKDSP_DET_MAX = 0, ATKDSP_DET_MIN = 1, ATKDSP_DET_AVG = 2 };
enum atkdsp_window { ATKDSP_WIN_RECT = 0, ATKDSP_WIN_HANN = 1, ATKDSP_WIN_HAMMING = 2,
                     ATKDSP_WIN_BLACKMANHARRIS = 3 };

/* ---- identity: reference include/atkdsp.h:238-245, src/atkdsp.c:43-80 ---- */
ATKDSP_API int         atkdsp_abi_version(void);
ATKDSP_API const char *atkdsp_version_string(void);
/* contains "abi 26", then "hip gfx950" and the device name */
ATKDSP_API const char *atkdsp_build_info(void);
ATKDSP_API int         atkdsp_bytes_per_sample(int fmt);
/* kept for the binding; the GPU result does not depend on it */
ATKDSP_API int         atkdsp_set_threads(int n);
ATKDSP_API int         atkdsp_get_threads(void);

/* ---- unpack: reference include/atkdsp.h:252-254, src/convert.c:7-17,113-120.
 * Bit-exact conversion; the one-pole DC block runs as a chunked scan whose
 * chunk boundaries carry an FP64 state (within 2e-6 of the serial pole). */
ATKDSP_API ptrdiff_t atkdsp_unpack(const void *raw, size_t nbytes, int fmt,
                                   atkdsp_cf32 *out, size_t out_cap,
                                   float dc_alpha, atkdsp_cf32 *dc_state);
/* reference include/atkdsp.h:270-273, src/convert.c:33-111,122-128.
 * Bit-exact; the mean is an FP64 fixed-order tree sum. */
ATKDSP_API ptrdiff_t atkdsp_unpack_dc(const void *raw, size_t nbytes, int fmt,
                                      atkdsp_cf32 *out, size_t out_cap,
                                      float off_re, float off_im,
                                      double *mean_re, double *mean_im);

/* ---- FFT: reference include/atkdsp.h:384-389, src/fft.c:31-131.
 * Any n >= 2. A plan may be shared by several caller threads: it owns a fixed
 * pool of device slots (stream + buffers) claimed per call. */
typedef struct atkdsp_fft atkdsp_fft;
ATKDSP_API atkdsp_fft *atkdsp_fft_create(size_t n);
ATKDSP_API void        atkdsp_fft_destroy(atkdsp_fft *p);
ATKDSP_API size_t      atkdsp_fft_length(const atkdsp_fft *p);
ATKDSP_API int         atkdsp_fft_exec(const atkdsp_fft *p, const atkdsp_cf32 *in,
                                       atkdsp_cf32 *out, int inverse);
/* reference include/atkdsp.h:392, src/fft.c:133-152 (create-time, host) */
ATKDSP_API int atkdsp_window(int kind, size_t n, float *out);
/* reference include/atkdsp.h:397-398, src/fft.c:155-174 */
ATKDSP_API int atkdsp_power_db(const atkdsp_fft *p, const atkdsp_cf32 *in,
                               const float *window, float *out_db);
/* reference include/atkdsp.h:406-409, src/fft.c:177-295 */
ATKDSP_API ptrdiff_t atkdsp_spectrum_reduce(const atkdsp_fft *p, const atkdsp_cf32 *in,
                                            size_t n_samples, size_t hop,
                                            const float *window, int detector,
                                            float *out_line);
/* reference include/atkdsp.h:428-432, src/fft.c:297-431 */
ATKDSP_API ptrdiff_t atkdsp_spectrum_stats(const atkdsp_fft *p, const atkdsp_cf32 *in,
                                           size_t n_samples, size_t hop,
                                           const float *window,
                                           float *max_db, float *avg_db,
                                           float *min_db, float *sk);
/* reference include/atkdsp.h:441-442, src/fft.c:433-444 (host arithmetic) */
ATKDSP_API int atkdsp_window_stats(const float *window, size_t n,
                                   double *coherent_gain, double *enbw_bins);

/* ---- display: reference include/atkdsp.h:459-464, src/display.c:4-37 ---- */
ATKDSP_API void atkdsp_db_to_pixels(const float *db, size_t n, float lo, float hi,
                                    const uint32_t *lut256, uint32_t *out);
ATKDSP_API void atkdsp_decimate_max(const float *in, size_t n, float *out, size_t m);
/* reference include/atkdsp.h:493-497, src/detect.c:116-134 */
ATKDSP_API void atkdsp_stitch_max(const float *seg, size_t nseg,
                                  double seg_lo_hz, double seg_hz_per_bin,
                                  double keep_lo_hz, double keep_hi_hz,
                                  double out_lo_hz, double out_hz_per_bin,
                                  float *out, size_t nout);

/* ---- detection on a spectrum line: reference include/atkdsp.h:465-486,
 * src/detect.c:4-114. Both run on the GPU (one block per line): the median is
 * an exact radix select, so it returns the reference's value (at most the
 * sign of a zero median may differ); detect_channels smooths, thresholds,
 * bridges gaps and measures runs on the device and orders the records on the
 * host with the reference's qsort comparator. smooth_scratch must be given,
 * as in the reference, but is not written. Lines longer than 2^21 bins (and
 * median inputs longer than 2^23) are refused: ATKDSP_E_ARG / NaN. ---- */
ATKDSP_API float atkdsp_median(const float *x, size_t n, float *scratch);
typedef struct {
    int   start_bin, end_bin;   /* [start, end) */
    float centroid_bin;         /* power-weighted centre */
    float peak_db;
    float snr_db;               /* peak - floor */
} atkdsp_channel;
ATKDSP_API ptrdiff_t atkdsp_detect_channels(const float *line, size_t n, float floor_db,
                                            float threshold_db, int gap_bins, int min_run,
                                            float *smooth_scratch,
                                            atkdsp_channel *out, size_t out_cap);

/* ---- CFAR threshold design: reference include/atkdsp.h:1733-1860,
 * src/cfar.c:23-268. The defaults, the equivalent independent cells of a
 * window-correlated periodogram, that correlation for a window, and the
 * threshold factor alpha for a false-alarm rate (host code, design time; the
 * solver is this library's own and agrees with the reference's to its
 * stated accuracy). ---- */
#define ATKDSP_CFAR_CA 0
#define ATKDSP_CFAR_OS 1
#define ATKDSP_CFAR_GO 2
#define ATKDSP_CFAR_SO 3
#define ATKDSP_CFAR_RHO_MAX 4
typedef struct {
    int32_t mode;        /* ATKDSP_CFAR_CA / OS / GO / SO */
    int32_t n_ref;       /* reference cells each side, 1..256 */
    int32_t n_guard;     /* guard cells each side, 0..256 */
    int32_t k_frames;    /* frames averaged before testing, 1..64 */
    double  pfa;         /* per cell per frame, 1e-15..0.5 */
    double  looks;       /* periodograms averaged into one frame, 0.25..1e4 */
    int32_t os_k;        /* OS rank 1..2 n_ref; 0 = ceil(3N/4) */
    int32_t open_m;
    int32_t open_n;
    int32_t close_k;
    int32_t merge_gap;
    int32_t min_bins;
    int32_t max_tracks;
    int32_t reserved;
    double  rho2[4];     /* power correlation of bins 1..4 apart, each in [0, 1) */
} atkdsp_cfar_config;    /* 96 bytes, as the reference header compiles (its comment says 104) */
ATKDSP_API void   atkdsp_cfar_defaults(atkdsp_cfar_config *cfg);
ATKDSP_API int    atkdsp_cfar_alpha(int mode, int n_ref, double pfa, double looks, int os_k, const double *rho2,
                                    double *alpha, double *et);
ATKDSP_API double atkdsp_cfar_n_eff(int n_ref, const double *rho2);
ATKDSP_API int    atkdsp_cfar_window_rho2(const float *window, size_t n, double *rho2);

/* ---- CFAR detection over spectrum frames and the burst tracker: reference
 * include/atkdsp.h:1733-1925, src/cfar.c:270-819. Frames of nbins dB values
 * go to the GPU, which forms power, the k_frames average and the CA / OS /
 * GO / SO test of every bin against its reference cells (FP64); the tracker
 * that turns hit bins into OPEN / UPDATE / CLOSE events runs on the host,
 * frame by frame. ---- */
#define ATKDSP_CFAR_OPEN   1
#define ATKDSP_CFAR_UPDATE 2
#define ATKDSP_CFAR_CLOSE  3
#define ATKDSP_CFAR_T0_COARSE 1u   /* flags */
#define ATKDSP_CFAR_T1_COARSE 2u
#define ATKDSP_CFAR_EDGE      4u   /* a bin of the span borrowed or lacked reference cells */
#define ATKDSP_CFAR_MASK_ADJ  8u   /* a masked bin within merge_gap + 1 of the span */
#define ATKDSP_CFAR_FLUSHED  16u
#define ATKDSP_CFAR_EDGE_PFA 1e-3  /* the single-frame band test behind t0 / t1 */
typedef struct {
    double   t0, t1;     /* seconds; t1 NaN before CLOSE */
    double   fc_bin;     /* excess-weighted centroid */
    int32_t  type;
    uint32_t id;         /* from 1, per handle */
    uint32_t flags;
    int32_t  lo_bin, hi_bin;
    int32_t  n_hits;     /* frames with a hit */
    float    level_db, noise_db, snr_db, band_snr_db;
} atkdsp_cfar_event;     /* 64 bytes */
typedef struct {
    uint64_t frames, warmup;
    uint64_t cells_tested, cell_hits, untestable;
    uint64_t opened, closed, updates;
    uint64_t dropped;    /* blobs refused at max_tracks */
    uint64_t resets;
    uint64_t tracks;     /* pending + open now */
} atkdsp_cfar_counters;  /* 88 bytes */
typedef struct atkdsp_cfar atkdsp_cfar;
/* NULL on an invalid configuration or without a GPU */
ATKDSP_API atkdsp_cfar *atkdsp_cfar_create(size_t nbins, const atkdsp_cfar_config *cfg);
ATKDSP_API void         atkdsp_cfar_destroy(atkdsp_cfar *c);
ATKDSP_API void         atkdsp_cfar_reset(atkdsp_cfar *c);         /* frames and tracks forgotten; mask kept */
ATKDSP_API int          atkdsp_cfar_set_mask(atkdsp_cfar *c, const uint8_t *mask);   /* NULL clears */
ATKDSP_API size_t       atkdsp_cfar_events_max(const atkdsp_cfar *c, size_t nframes);
/* E_CAP, before any frame is taken, when cap < events_max(nframes) */
ATKDSP_API ptrdiff_t    atkdsp_cfar_process(atkdsp_cfar *c, const float *frames_db, size_t nframes,
                                            double frame_t0, double frame_dt, atkdsp_cfar_event *ev, size_t cap);
ATKDSP_API ptrdiff_t    atkdsp_cfar_flush(atkdsp_cfar *c, atkdsp_cfar_event *ev, size_t cap);
/* the last tested frame's average and threshold (dB) and hits; E_EMPTY before one */
ATKDSP_API int          atkdsp_cfar_last(const atkdsp_cfar *c, float *avg_db, float *thr_db, uint8_t *hits);
ATKDSP_API int          atkdsp_cfar_info(const atkdsp_cfar *c, double *alpha, double *alpha_ca, double *et,
                                         atkdsp_cfar_config *cfg);
ATKDSP_API int          atkdsp_cfar_describe(const atkdsp_cfar *c, char *buf, size_t cap);
ATKDSP_API int          atkdsp_cfar_stats(const atkdsp_cfar *c, atkdsp_cfar_counters *st);

/* ---- NCO: reference include/atkdsp.h:309-319, src/nco.c:27-98.
 * out[n] = in[n] * exp(j*(phase + n*step)); the phase advances by n*step,
 * wrapped, in FP64 in the caller's struct; step = -2*pi*f/fs. in and out may
 * alias. The rotation runs on the GPU. */
typedef struct { double phase; double step; } atkdsp_nco;
ATKDSP_API void atkdsp_nco_init(atkdsp_nco *nco, double freq_hz, double sample_rate);
ATKDSP_API void atkdsp_nco_set_freq(atkdsp_nco *nco, double freq_hz, double sample_rate);
ATKDSP_API void atkdsp_nco_mix(atkdsp_nco *nco, const atkdsp_cf32 *in, atkdsp_cf32 *out, size_t n);

/* ---- the block ring: reference include/atkdsp.h:2020-2107, src/ring.c.
 * A block ring in caller memory (a private buffer, or a shared mapping another
 * process attaches to): one producer that never waits, up to 16 consumers
 * with their own cursors, fixed-size blocks each with a header, a seqlock per
 * block (lock word 2 seq + 1 while written, 2 seq + 2 once committed). The
 * memory layout is the reference's, offset for offset, so a ring made by
 * either library is attached by the other. Host code: it allocates nothing
 * and starts no thread. atkgpu_server_submit_ring (include/atkgpu.h) feeds
 * a ring's blocks into the serving loop. */
#define ATKDSP_RING_MAX_CONSUMERS 16
#define ATKDSP_RING_RETUNE   1u
#define ATKDSP_RING_OVERFLOW 2u
#define ATKDSP_RING_CAL      4u
typedef struct {
    uint64_t seq, first_sample;
    int64_t  t_utc_ns;
    double   fc_hz, fs_hz;
    float    gain_db;
    int32_t  format;
    uint32_t flags, nbytes;
} atkdsp_ring_meta;                /* 56 bytes */
typedef struct {
    uint64_t committed, oldest;
    uint64_t block_bytes, nblocks, total_bytes;
    uint64_t torn;
    uint32_t max_consumers, active, active_mask, reserved;
    uint64_t lost[ATKDSP_RING_MAX_CONSUMERS];
    uint64_t read[ATKDSP_RING_MAX_CONSUMERS];
    uint64_t lag[ATKDSP_RING_MAX_CONSUMERS];
} atkdsp_ring_counters;            /* 448 bytes */
typedef struct atkdsp_ring atkdsp_ring;
ATKDSP_API size_t       atkdsp_ring_bytes(size_t block_bytes, size_t nblocks, int max_consumers);
ATKDSP_API atkdsp_ring *atkdsp_ring_init(void *mem, size_t mem_bytes, size_t block_bytes, size_t nblocks,
                                         int max_consumers);
ATKDSP_API atkdsp_ring *atkdsp_ring_attach(void *mem, size_t mem_bytes);
ATKDSP_API void        *atkdsp_ring_claim(atkdsp_ring *r);
ATKDSP_API int          atkdsp_ring_commit(atkdsp_ring *r, const atkdsp_ring_meta *m);
ATKDSP_API int          atkdsp_ring_open(atkdsp_ring *r, int from_oldest);
ATKDSP_API int          atkdsp_ring_close(atkdsp_ring *r, int id);
ATKDSP_API ptrdiff_t    atkdsp_ring_read(atkdsp_ring *r, int id, void *out, size_t cap, atkdsp_ring_meta *m,
                                         uint64_t *lost);
ATKDSP_API const void  *atkdsp_ring_peek(atkdsp_ring *r, int id, atkdsp_ring_meta *m, uint64_t *lost);
ATKDSP_API int          atkdsp_ring_release(atkdsp_ring *r, int id);
ATKDSP_API ptrdiff_t    atkdsp_ring_copy_range(const atkdsp_ring *r, uint64_t seq0, uint64_t seq1, void *out,
                                               size_t cap, atkdsp_ring_meta *metas, size_t mcap);
ATKDSP_API int          atkdsp_ring_stats(const atkdsp_ring *r, atkdsp_ring_counters *st);
ATKDSP_API int          atkdsp_ring_describe(const atkdsp_ring *r, char *buf, size_t cap);

/* ---- arbitrary-ratio resampler: reference include/atkdsp.h:368-377,
 * src/resample.c:100-252. A 64-phase (nphase) bank of the Kaiser lowpass below
 * with linear interpolation toward the next phase; output K sits at input
 * position base + K*in/out, carried across calls, so block boundaries are
 * invisible. The FIR runs on the GPU; a call whose outputs exceed out_cap
 * returns ATKDSP_E_CAP with the state unchanged. */
typedef struct atkdsp_arb_resampler atkdsp_arb_resampler;
ATKDSP_API atkdsp_arb_resampler *atkdsp_arb_resampler_create(double in_rate, double out_rate,
                                                             double atten_db, unsigned nphase);
ATKDSP_API void      atkdsp_arb_resampler_destroy(atkdsp_arb_resampler *r);
ATKDSP_API void      atkdsp_arb_resampler_reset(atkdsp_arb_resampler *r);
ATKDSP_API void      atkdsp_arb_resampler_set_ratio(atkdsp_arb_resampler *r, double ratio);
ATKDSP_API double    atkdsp_arb_resampler_ratio(const atkdsp_arb_resampler *r);
ATKDSP_API size_t    atkdsp_arb_resampler_out_max(const atkdsp_arb_resampler *r, size_t n_in);
ATKDSP_API ptrdiff_t atkdsp_arb_resampler_process(atkdsp_arb_resampler *r, const atkdsp_cf32 *in,
                                                  size_t n, atkdsp_cf32 *out, size_t out_cap);

/* ---- Kaiser lowpass design: reference include/atkdsp.h:505-506,
 * src/design.c:5-62. Odd length, unit DC gain, FP64 design cast once to
 * float; out == NULL returns the length. Host code (a design-time helper). */
ATKDSP_API ptrdiff_t atkdsp_design_lowpass(double fp_hz, double fs_hz, double atten_db,
                                           double rate, float *out, size_t out_cap);

/* ---- polyphase filterbank: reference include/atkdsp.h:3191-3250,
 * src/channelizer.c:79-471. Output layouts, bookkeeping, cal and W1/W2 are
 * the reference's; the streaming history (ov samples per element) stays in
 * device memory between calls. */
typedef struct atkdsp_pfb atkdsp_pfb;
ATKDSP_API atkdsp_pfb *atkdsp_pfb_create(double fs, unsigned nchan, unsigned oversample,
                                         unsigned taps_per_poly, double atten_db,
                                         unsigned max_nel);
ATKDSP_API void        atkdsp_pfb_destroy(atkdsp_pfb *p);
ATKDSP_API void        atkdsp_pfb_reset(atkdsp_pfb *p);
ATKDSP_API unsigned    atkdsp_pfb_nchan(const atkdsp_pfb *p);
ATKDSP_API double      atkdsp_pfb_out_rate(const atkdsp_pfb *p);
ATKDSP_API size_t      atkdsp_pfb_out_hops(const atkdsp_pfb *p, size_t n);
ATKDSP_API ptrdiff_t   atkdsp_pfb_process(atkdsp_pfb *p, const atkdsp_cf32 *in, size_t n,
                                          atkdsp_cf32 *out, size_t cap);
ATKDSP_API ptrdiff_t   atkdsp_pfb_process_n(atkdsp_pfb *p, const atkdsp_cf32 *const *ins,
                                            unsigned nel, size_t n,
                                            atkdsp_cf32 *out, size_t cap);
ATKDSP_API int64_t     atkdsp_pfb_block_origin(const atkdsp_pfb *p, int64_t hop_index);
ATKDSP_API double      atkdsp_pfb_group_delay(const atkdsp_pfb *p);
ATKDSP_API int         atkdsp_pfb_set_cal(atkdsp_pfb *p, const atkdsp_cf32 *table, unsigned nel);
ATKDSP_API int         atkdsp_pfb_describe(const atkdsp_pfb *p, char *buf, size_t cap);
ATKDSP_API void        atkdsp_pfb_cov_accumulate(const atkdsp_cf32 *out, unsigned nel,
                                                 unsigned nchan, size_t m,
                                                 const uint8_t *gate,
                                                 atkdsp_cf32 *R, uint32_t *counts);
ATKDSP_API void        atkdsp_pfb_cov_reset(atkdsp_cf32 *R, uint32_t *counts, unsigned nel,
                                            unsigned nchan, const uint8_t *gate);
ATKDSP_API void        atkdsp_pfb_channel_power(const atkdsp_cf32 *out, unsigned nel,
                                                unsigned nchan, size_t m,
                                                float *pwr, float *mean_pwr);

/* ---- overlap-save extractor: reference include/atkdsp.h:3252-3289,
 * src/channelizer.c:473-788. One nfft-point forward FFT per block (overlap
 * nfft/2); per channel the roll to its integer bin, its lowpass's frequency
 * response, the fold to M = nfft/decim and the M-point inverse FFT (the
 * decimation), the ov/decim transient dropped, the mix to absolute time in
 * FP64, then the arbitrary resampler to out_rate when the channel rate
 * differs. All of it runs on the GPU; the history stays in device memory.
 * Channels are independent: add, retune and remove leave the others
 * bit-identical. outs is [nel][nslots] (NULL for a free slot); a call whose
 * outputs exceed cap_each returns ATKDSP_E_CAP with the state unchanged. */
typedef struct atkdsp_osx atkdsp_osx;
ATKDSP_API atkdsp_osx *atkdsp_osx_create(double fs, size_t nfft, double atten_db, unsigned max_nel);
ATKDSP_API void        atkdsp_osx_destroy(atkdsp_osx *x);
ATKDSP_API void        atkdsp_osx_reset(atkdsp_osx *x);
ATKDSP_API int         atkdsp_osx_add(atkdsp_osx *x, double offset_hz, double bw_hz, double out_rate);
ATKDSP_API int         atkdsp_osx_retune(atkdsp_osx *x, int ch, double offset_hz);
ATKDSP_API int         atkdsp_osx_remove(atkdsp_osx *x, int ch);
ATKDSP_API int         atkdsp_osx_channels(const atkdsp_osx *x);
ATKDSP_API double      atkdsp_osx_chan_rate(const atkdsp_osx *x, int ch);
ATKDSP_API size_t      atkdsp_osx_out_max(const atkdsp_osx *x, int ch, size_t n);
ATKDSP_API ptrdiff_t   atkdsp_osx_process(atkdsp_osx *x, const atkdsp_cf32 *in, size_t n,
                                          atkdsp_cf32 *const *outs, size_t *out_counts, size_t cap_each);
ATKDSP_API ptrdiff_t   atkdsp_osx_process_n(atkdsp_osx *x, const atkdsp_cf32 *const *ins, unsigned nel,
                                            size_t n, atkdsp_cf32 *const *outs, size_t *out_counts,
                                            size_t cap_each);
ATKDSP_API int         atkdsp_osx_describe(const atkdsp_osx *x, char *buf, size_t cap);

#ifdef __cplusplus
}
#endif
#endif /* ATKDSP_H */
