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
 * Stream s covers [fc0 + s*fc_step - fs/2, fc0 + s*fc_step + fs/2]. */
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
 * the display kernels alone */
ATKGPU_API int           atkgpu_chain_timing(atkgpu_chain *c, float *ms_run, float *ms_pfb, float *ms_disp);
/* results of the last run, copied to host (synchronising) */
ATKGPU_API ptrdiff_t     atkgpu_chain_read_pfb(atkgpu_chain *c, unsigned s, atkdsp_cf32 *out, size_t cap);
ATKGPU_API int           atkgpu_chain_read_line(atkgpu_chain *c, unsigned s, float *line_db, float *cols,
                                                float *trace);
ATKGPU_API int           atkgpu_chain_read_mean(atkgpu_chain *c, unsigned s, double *mean2);
ATKGPU_API long long     atkgpu_chain_frames(const atkgpu_chain *c);   /* frames of the last run */

#ifdef __cplusplus
}
#endif
#endif /* ATKGPU_H */
