/*
 * atkgpu.h — the device-resident side of the wideband front end on one
 * MI355X, beside the drop-in atkdsp ABI (include/atkdsp.h).
 *
 * The reference ABI takes host pointers and returns host results, so every
 * call pays PCIe both ways. These entry points keep samples, filterbank
 * history, outputs and display lines in HBM across blocks, for device
 * consumers and for throughput. None of them exists in the reference; their
 * arithmetic is the reference's (unpack src/convert.c:33-111, filterbank
 * src/channelizer.c:218-402, display reduce src/fft.c:177-295,
 * src/display.c:19-37, stitch src/detect.c:116-134).
 *
 * Plain C. Nothing in a *_run / *_submit call allocates; everything is made
 * at *_create. Return codes are the ATKDSP_E_* values of atkdsp.h.
 */
#ifndef ATKGPU_H
#define ATKGPU_H

#include <stddef.h>
#include <stdint.h>
#include "atkdsp.h"

#define ATKGPU_API __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif

#define ATKGPU_VERSION 1
ATKGPU_API int atkgpu_version(void);

/* ---- the chain: S streams of one shape, one block per stream per run --------
 * Per stream, per block of n raw samples:
 *   unpack (bit-exact, block DC removal with the previous block's mean when
 *   dc_mode = 1) -> filterbank (M channels, os x, K taps/branch, history in
 *   HBM) -> display line over every nfft frame at fft_hop (the < nfft tail is
 *   carried to the next block) -> decimate_max to line_cols -> NaN-aware
 *   max stitch into the stream's trace.
 * The display reduction reads each raw block directly and unpacks it into the
 * stream's row as it goes (with the block mean); the filterbank then reads
 * the row. Stream s covers [fc0 + s*fc_step - fs/2, fc0 + s*fc_step + fs/2]. */
typedef struct {
    int fmt;                    /* enum atkdsp_format of the raw samples   */
    double fs;
    unsigned nchan, oversample, taps_per_poly;
    double atten_db;
    unsigned nfft, fft_hop;     /* nfft in {1024, 2048, 4096}              */
    int window;                 /* enum atkdsp_window                      */
    int detector;               /* enum atkdsp_detector                    */
    unsigned line_cols;         /* decimated columns                       */
    unsigned trace_bins;        /* stitched trace bins (0: no trace)       */
    double trace_lo_hz, trace_hz_per_bin;
    double fc0_hz, fc_step_hz;
    int dc_mode;                /* 0: no offset; 1: previous block's mean  */
    size_t block_max;           /* most samples in one block               */
} atkgpu_chain_cfg;

typedef struct atkgpu_chain atkgpu_chain;
ATKGPU_API atkgpu_chain *atkgpu_chain_create(const atkgpu_chain_cfg *cfg, unsigned nstreams);
ATKGPU_API void          atkgpu_chain_destroy(atkgpu_chain *c);
ATKGPU_API void          atkgpu_chain_reset(atkgpu_chain *c);
/* stream s's raw input slot in device memory (block_max * bytes_per_sample) */
ATKGPU_API void         *atkgpu_chain_input(atkgpu_chain *c, unsigned s);
/* copy host bytes into stream s's input slot (ordered before the next run) */
ATKGPU_API int           atkgpu_chain_upload(atkgpu_chain *c, unsigned s, const void *raw, size_t nbytes);
/* enqueue one block of n samples for every stream (asynchronous);
 * returns the filterbank hops m the block produces per stream */
ATKGPU_API ptrdiff_t     atkgpu_chain_run(atkgpu_chain *c, size_t n);
ATKGPU_API int           atkgpu_chain_sync(atkgpu_chain *c);
/* device times of the last run, ms: whole run, the filterbank kernel alone,
 * the display reduction with the unpack and block mean (ATKDSP_E_ARG when
 * the last run was not timed) */
ATKGPU_API int           atkgpu_chain_timing(atkgpu_chain *c, float *ms_run, float *ms_pfb, float *ms_disp);
/* on != 0 (the default): each run records the four timing events that
 * atkgpu_chain_timing reads; each costs the stream about 5 us, so a caller
 * that does not read them may turn them off */
ATKGPU_API int           atkgpu_chain_set_timing(atkgpu_chain *c, int on);
/* results of the last run, copied to host (synchronising) */
ATKGPU_API ptrdiff_t     atkgpu_chain_read_pfb(atkgpu_chain *c, unsigned s, atkdsp_cf32 *out, size_t cap);
ATKGPU_API int           atkgpu_chain_read_line(atkgpu_chain *c, unsigned s, float *line_db, float *cols,
                                                float *trace);
ATKGPU_API int           atkgpu_chain_read_mean(atkgpu_chain *c, unsigned s, double *mean2);
ATKGPU_API long long     atkgpu_chain_frames(const atkgpu_chain *c);   /* frames of the last run */

/* ---- the serving loop: many independent streams, chunks of any size ----------
 * A stream is admitted with its own format and filterbank shape; its device
 * state (history, taps, an output ring of out_slots chunk results) is
 * allocated at admission and charged against the server's HBM budget, and a
 * stream that does not fit is refused with ATKDSP_E_NOMEM. Chunks are taken
 * by a dispatcher thread in batches (one chunk per stream per batch, streams
 * round-robin) while earlier batches run; every stream's chunks execute in
 * order. The offset (off_re, off_im) is fixed per stream, so the output does
 * not depend on how the stream is chunked or what else is in a batch.
 * Latency = completion of the chunk's batch minus its submit call, both on
 * the host monotonic clock.
 *
 * A stream may be an array of nel phase-coherent elements sharing the
 * prototype: a chunk then holds nel element blocks of n samples each,
 * element-major, and its result is [nel][nchan][m]. A coherent stream also
 * accumulates W1 (per-channel nel x nel cross-spectra,
 * atkdsp_pfb_cov_accumulate's arithmetic) over all its chunks and keeps W2
 * (atkdsp_pfb_channel_power) of its latest chunk, in device memory.
 *
 * A stream may also carry a display: every disp_frames frames of its samples
 * (frame f = samples [f*disp_hop, f*disp_hop + disp_nfft) of the stream from
 * its first sample) make one line, computed as atkdsp_spectrum_reduce over
 * those frames (window, FFT, |X|^2/n^2, detector, dB, fftshifted) followed by
 * atkdsp_decimate_max to disp_cols columns. Frames are transformed as the
 * chunks that complete them arrive; the running detector line stays on the
 * device, and a completed line goes into a ring of disp_slots lines. A chunk's
 * latency covers the lines it completes. */
typedef struct {
    int fmt;                    /* enum atkdsp_format */
    double fs;
    unsigned nchan, oversample, taps_per_poly;  /* nchan 64..2048 (power of 2), os 1 or 2, K*os <= 32 */
    double atten_db;
    float off_re, off_im;       /* offset removed at unpack */
    size_t chunk_max;           /* largest chunk, samples (per element) */
    unsigned out_slots;         /* chunk results kept per stream */
    unsigned nel;               /* elements per chunk, 1..16 (0 means 1) */
    int coherent;               /* W1 / W2 for an array stream */
    unsigned disp_nfft;         /* display FFT length: 0 (no display), 1024, 2048 or 4096 */
    unsigned disp_hop;          /* frame hop, 1..disp_nfft (0 means disp_nfft) */
    int disp_window;            /* enum atkdsp_window */
    int disp_detector;          /* enum atkdsp_detector */
    unsigned disp_frames;       /* frames per line (>= 1) */
    unsigned disp_cols;         /* columns per line, 1..disp_nfft (0 means disp_nfft) */
    unsigned disp_slots;        /* lines kept per stream (0 means 4) */
} atkgpu_stream_cfg;

typedef struct {
    long long chunks, samples, batches;
    double p50_ms, p99_ms, p999_ms, max_ms, mean_ms;
    double busy_s;              /* first submit to last completion */
    size_t hbm_used, hbm_budget;
    unsigned streams;
    long long lat_count;        /* chunks the latency figures cover: every completed chunk */
    double lat_rel_res;         /* relative resolution of p50/p99/p99.9 (log-spaced histogram);
                                   mean and max are exact */
} atkgpu_serve_stats;

typedef struct atkgpu_server atkgpu_server;
/* staging_bytes: pinned host ring for chunks in flight (>= 1 MiB) */
ATKGPU_API atkgpu_server *atkgpu_server_create(size_t hbm_budget_bytes, unsigned max_streams, unsigned max_batch,
                                               size_t staging_bytes);
ATKGPU_API void      atkgpu_server_destroy(atkgpu_server *v);
/* returns the stream id (>= 0), ATKDSP_E_NOMEM (over budget / no slot) or ATKDSP_E_ARG */
ATKGPU_API int       atkgpu_server_add_stream(atkgpu_server *v, const atkgpu_stream_cfg *cfg);
/* waits for the stream's chunks, then returns its memory to the budget */
ATKGPU_API int       atkgpu_server_remove_stream(atkgpu_server *v, int sid);
/* copies the chunk (whole samples of nbytes) and returns its ticket (>= 0);
 * blocks while the stream's queue or the staging ring is full */
ATKGPU_API long long atkgpu_server_submit(atkgpu_server *v, int sid, const void *raw, size_t nbytes);
/* submit from a block ring (atkdsp_ring_*): the next intact block of the
 * ring consumer becomes one chunk of the stream, copied once, from the ring
 * slot straight into staging. Returns the ticket, or ATKGPU_RING_EMPTY when
 * the consumer has nothing new; *meta gets the block's header (first_sample
 * and time tag, for the caller's continuity checks) and *lost the blocks the
 * consumer lost before it (lapped by the producer, or rewritten while being
 * copied; those are skipped, as atkdsp_ring_read skips them). A block whose
 * format is not the stream's, or that is larger than chunk_max, is not
 * consumed: ATKDSP_E_ARG. meta and lost may be NULL. */
#define ATKGPU_RING_EMPTY (-64)
ATKGPU_API long long atkgpu_server_submit_ring(atkgpu_server *v, int sid, atkdsp_ring *ring, int consumer,
                                               atkdsp_ring_meta *meta, uint64_t *lost);
ATKGPU_API int       atkgpu_server_wait(atkgpu_server *v, int sid, long long ticket);
/* the chunk's filterbank output, [nchan][m] channel-major; returns m. Valid
 * until out_slots later chunks of the stream have been dispatched:
 * ATKDSP_E_ARG when that happened before the call, ATKDSP_E_OVERRUN (the copy
 * may be torn) when it happened during the copy. */
ATKGPU_API ptrdiff_t atkgpu_server_read(atkgpu_server *v, int sid, long long ticket, atkdsp_cf32 *out, size_t cap);
ATKGPU_API int       atkgpu_server_drain(atkgpu_server *v);
/* a coherent stream's W1 over every chunk submitted so far (waits for them):
 * R [nchan][nel][nel], counts [nchan]; and W2 of its latest completed chunk,
 * pwr [nel][nchan], mean_pwr [nchan]. Either output may be NULL. */
ATKGPU_API int       atkgpu_server_read_cov(atkgpu_server *v, int sid, atkdsp_cf32 *R, uint32_t *counts);
ATKGPU_API int       atkgpu_server_read_power(atkgpu_server *v, int sid, float *pwr, float *mean_pwr);
ATKGPU_API int       atkgpu_server_stats(atkgpu_server *v, atkgpu_serve_stats *st, int reset);
/* display lines: how many are complete, and line L as [nel][disp_cols] dB
 * values (waits for the chunk that completes it; L must be completed by the
 * chunks submitted so far). Valid until disp_slots later lines of the stream
 * have been dispatched: ATKDSP_E_ARG when that happened before the call,
 * ATKDSP_E_OVERRUN (the copy may be torn) when it happened during it. */
ATKGPU_API long long atkgpu_server_lines(atkgpu_server *v, int sid);
ATKGPU_API int       atkgpu_server_read_line(atkgpu_server *v, int sid, long long line, float *out, size_t cap);
/* the line's consumer on the device: atkdsp_detect_channels over element
 * `element` of line L (waits as read_line does), only the records returned,
 * strongest first. floor_db NaN = the median of the 5-bin-smoothed line
 * (signal_id's floor), computed on the device; *floor_used (may be NULL)
 * gets the floor applied. ATKDSP_E_CAP when more than out_cap are found (out
 * then holds the first out_cap in bin order); ATKDSP_E_OVERRUN when the
 * line's ring slot was reused while the query ran. */
ATKGPU_API ptrdiff_t atkgpu_server_detect(atkgpu_server *v, int sid, long long line, unsigned element,
                                          float floor_db, float threshold_db, int gap_bins, int min_run,
                                          atkdsp_channel *out, size_t out_cap, float *floor_used);
/* CFAR (atkdsp_cfar, include/atkdsp.h) over served lines where they live in
 * HBM: lines [first, first + n) of one element are the handle's next n frames
 * (frame times frame_t0 + i frame_dt), tested on the device and tracked as
 * atkdsp_cfar_process would take them from host memory; events as there.
 * The handle's nbins must equal the stream's display columns; n at most 256
 * and at most the lines the ring keeps. Waits for the lines to complete.
 * ATKDSP_E_ARG when a line is not completed by the chunks submitted or has
 * left the ring; ATKDSP_E_OVERRUN (handle unchanged) when the ring overwrote
 * one while it was being copied. Calls on one handle must not overlap. */
ATKGPU_API ptrdiff_t atkgpu_server_cfar(atkgpu_server *v, int sid, unsigned element, atkdsp_cfar *c,
                                        long long first, size_t n, double frame_t0, double frame_dt,
                                        atkdsp_cfar_event *ev, size_t cap);
/* the same over the newest completed line (element 0) of each of n streams
 * (one display line length among them), in one launch, the floor of each
 * the median of its smoothed line: for stream i, counts[i] channels at
 * out + i*per_cap, strongest first (ATKDSP_E_CAP past min(per_cap, 256),
 * the first ones then in bin order), lines[i] the line used (-1: none yet,
 * count 0) and floors[i] its floor; lines and floors may be NULL */
ATKGPU_API int       atkgpu_server_detect_latest(atkgpu_server *v, const int *sids, unsigned n, float threshold_db,
                                                 int gap_bins, int min_run, atkdsp_channel *out, size_t per_cap,
                                                 ptrdiff_t *counts, long long *lines, float *floors);

/* ---- the filterbank in matrix-core form: S streams of one shape ----------------
 * The polyphase fold as a matrix product on the MFMA units, from raw 8-bit
 * (cu8, ci8) or 12-bit (ci16q11) samples held exactly in BF16 or FP16, taps
 * split into hi + lo terms (or one term), FP32 accumulation; the FFT,
 * rotation and channel-major store are the FP32 path's. The offset is fixed
 * per stream (unpack_dc's offset argument). form ATKGPU_FORM_FP32 runs the
 * plain FP32 path (unpack, then the fused FP32 filterbank) on the same
 * plumbing, the form the others are measured against; ATKGPU_FORM_FP32MM is
 * that path with the FP32 fold on the matrix cores (4x4x1 FP32 products,
 * one rounding a tap, taps in ascending order: bit-identical to the FP32
 * form for finite input; 256, 512 or 1024 channels at 2x, 256 or 512 at 1x,
 * up to 16 taps/branch). ATKGPU_FORM_FP32RAW is the FP32 form reading the raw
 * rows the matrix-core forms read, each sample converted as it is loaded
 * (bit-identical to the FP32 form; 64 or 128 channels, 1x or 2x, up to 32 tap
 * rows, cu8/ci8/ci16/ci16q11): the like-for-like baseline of those forms.
 * Supported shapes: 64 channels 1x with up to 8 taps/branch (every form and
 * format), 64 channels 1x up to 16 or 32 taps, 64 channels 2x up to 8 taps,
 * 128 channels 2x up to 8 taps (BF16 hi+lo for 8-bit, FP16 hi+lo for 12-bit). */
enum { ATKGPU_FORM_FP32 = 0, ATKGPU_FORM_BF16X2 = 1, ATKGPU_FORM_BF16 = 2, ATKGPU_FORM_FP16X2 = 3,
       ATKGPU_FORM_FP16 = 4, ATKGPU_FORM_FP32MM = 5, ATKGPU_FORM_FP32RAW = 6 };
typedef struct {
    int fmt;                   /* ATKDSP_FMT_CU8, _CI8 or _CI16Q11 (FP32 form: any) */
    double fs;
    unsigned nchan, oversample, taps_per_poly;
    double atten_db;
    int form;                  /* ATKGPU_FORM_* */
    size_t block_max;          /* samples per stream per run, a multiple of 8 */
    float off_re, off_im;      /* the fixed offset subtracted by the unpack */
    int coherent;              /* 1: the streams are phase-coherent elements of one array; every
                                  run also accumulates W1 (per-channel cross-spectra over all
                                  elements, atkdsp_pfb_cov_accumulate's arithmetic) and computes
                                  W2 (atkdsp_pfb_channel_power) from the device output */
} atkgpu_pfbx_cfg;
typedef struct atkgpu_pfbx atkgpu_pfbx;
ATKGPU_API atkgpu_pfbx *atkgpu_pfbx_create(const atkgpu_pfbx_cfg *cfg, unsigned nstreams);
ATKGPU_API void         atkgpu_pfbx_destroy(atkgpu_pfbx *v);
ATKGPU_API int          atkgpu_pfbx_reset(atkgpu_pfbx *v);
/* device address of stream s's next raw block (valid until the next run) */
ATKGPU_API void        *atkgpu_pfbx_input(atkgpu_pfbx *v, unsigned s);
ATKGPU_API int          atkgpu_pfbx_upload(atkgpu_pfbx *v, unsigned s, const void *raw, size_t nbytes);
/* one block of n samples for every stream; returns the hops produced */
ATKGPU_API ptrdiff_t    atkgpu_pfbx_run(atkgpu_pfbx *v, size_t n);
ATKGPU_API int          atkgpu_pfbx_sync(atkgpu_pfbx *v);
/* device times of the last run: whole run, the unpack (FP32 form only), the filterbank kernel,
 * and W1 + W2 (coherent only) */
ATKGPU_API int          atkgpu_pfbx_timing(atkgpu_pfbx *v, float *ms_run, float *ms_unpack, float *ms_pfb,
                                           float *ms_w);
/* stream s's last output, [nchan][m] */
ATKGPU_API ptrdiff_t    atkgpu_pfbx_read(atkgpu_pfbx *v, unsigned s, atkdsp_cf32 *out, size_t cap);
/* coherent arrays: W1 accumulated since create / cov_reset, R [nchan][nel][nel] and counts
 * [nchan]; W2 of the last run, pwr [nel][nchan] and mean_pwr [nchan] (either may be NULL) */
ATKGPU_API int          atkgpu_pfbx_read_cov(atkgpu_pfbx *v, atkdsp_cf32 *R, uint32_t *counts);
ATKGPU_API int          atkgpu_pfbx_read_power(atkgpu_pfbx *v, float *pwr, float *mean_pwr);
ATKGPU_API int          atkgpu_pfbx_cov_reset(atkgpu_pfbx *v);

/* ---- the device-resident NCO: a long run fed from device buffers ---------------
 * The FP64 phase lives in device memory across runs (two slots used in turn)
 * and is logged after each run; the rotation is the drop-in's (FP64 phase
 * per sample, FP32 sincos of the reduced angle). Nothing in a run allocates. */
typedef struct atkgpu_nco atkgpu_nco;
ATKGPU_API atkgpu_nco *atkgpu_nco_create(double freq_hz, double sample_rate, size_t chunk_max, size_t log_max);
ATKGPU_API void        atkgpu_nco_destroy(atkgpu_nco *v);
ATKGPU_API void       *atkgpu_nco_input(atkgpu_nco *v);     /* chunk_max cf32, device */
ATKGPU_API void       *atkgpu_nco_output(atkgpu_nco *v);    /* chunk_max cf32, device */
ATKGPU_API int         atkgpu_nco_upload(atkgpu_nco *v, const atkdsp_cf32 *in, size_t n);
/* mix input[0, n) into output[0, n) and advance the phase by n samples */
ATKGPU_API int         atkgpu_nco_run(atkgpu_nco *v, size_t n);
ATKGPU_API int         atkgpu_nco_sync(atkgpu_nco *v);
ATKGPU_API ptrdiff_t   atkgpu_nco_read(atkgpu_nco *v, atkdsp_cf32 *out, size_t n);
ATKGPU_API int         atkgpu_nco_phase(atkgpu_nco *v, double *phase, long long *samples);
/* the phase after each of the first log_max runs */
ATKGPU_API ptrdiff_t   atkgpu_nco_log(atkgpu_nco *v, double *phases, size_t cap);

/* ---- unpack with the block mean, allocating nothing ---------------------------
 * No counterpart in the reference: the device-buffer form of atkdsp_unpack_dc
 * (include/atkdsp.h) for callers that cannot allocate or synchronize, such as
 * code under HIP graph capture and tight serving loops. raw, out and mean are
 * device pointers: out receives min(nbytes / bytes_per_sample, out_cap)
 * samples, bit-identical to atkdsp_unpack_dc's, and mean (2 doubles, or NULL)
 * the mean of the samples before the offset came out, bit-identical to its
 * mean_re / mean_im. Work is queued on `stream` (NULL: the null stream) and
 * not waited for; the return value is the sample count.
 *
 * The mean's scratch (the FP64 block partials of one pass and the sums of up
 * to 512 passes of 2^19 samples, 24 KiB) is one static device array per GPU,
 * held until the process ends: calls that overlap on one GPU race on it, so
 * serialize them (one stream, or an event between). More than
 * ATKGPU_NOALLOC_MAX_SAMPLES samples: ATKDSP_E_CAP, nothing queued. No
 * hipMalloc, hipFree, memset or host copy is made. */
#define ATKGPU_NOALLOC_MAX_SAMPLES ((size_t)512 << 19)
ATKGPU_API ptrdiff_t atkgpu_unpack_dc_noalloc(void *stream, const void *raw, size_t nbytes, int fmt,
                                               atkdsp_cf32 *out, size_t out_cap, float off_re, float off_im,
                                               double *mean);

#ifdef __cplusplus
}
#endif
#endif /* ATKGPU_H */
