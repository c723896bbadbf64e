/* Host-side internals of the drop-in atkdsp library (C11). */
#ifndef ATK_HOST_INTERNAL_H
#define ATK_HOST_INTERNAL_H

#include <stdatomic.h>
#include <stddef.h>
#include <stdint.h>
#include "../../../include/atkdsp.h"
#include "../dev/atkgpu_dev.h"

#define ATK_PI 3.14159265358979323846

/* A device slot: one stream plus fixed device buffers, claimed by CAS so that
 * concurrent callers of a shared object never share a stream or a buffer
 * (the reference claims CPU scratch slots the same way, src/internal.h:51-75,
 * src/fft.c:74-88). Slots are made at create / one-time init, never per call. */
typedef struct {
    atomic_int busy;
    void *stream;
    void *stream2;             /* the library pool's slots: the return copies of atk_pipe */
    void *d_a, *d_b, *d_c;     /* general device buffers */
    size_t bytes_a, bytes_b, bytes_c;
    void *d_small;             /* 1 MiB of small device scratch (lines, sums, luts) */
} atk_slot;

#define ATK_SMALL_BYTES (1u << 20)

int  atk_slot_init(atk_slot *s, size_t a, size_t b, size_t c);
void atk_slot_free(atk_slot *s);
int  atk_claim(atk_slot *slots, int n);   /* spins until one is free; returns its index */
void atk_release(atk_slot *slots, int i);

/* The library-wide pool used by the stateless entry points (unpack, display,
 * stitch, W1/W2). Made once, on first use. Returns NULL when there is no GPU. */
typedef struct {
    atk_slot slot[4];
    int nslot;
    size_t chunk;              /* samples a slot converts per pass */
} atk_pool;
atk_pool *atk_pool_get(void);

/* Copies between caller-owned host memory and the device on `stream`
 * (xfer_host.c): large ones through pinned staging. atk_put returns once src
 * has been read (the device copy lands in stream order); atk_get and
 * atk_get_rows return once dst is written (after the stream's earlier work).
 * atk_get_rows reads `rows` packed rows of `width` bytes and writes them
 * `dpitch` bytes apart. Nonzero on a device error. */
int atk_put(void *stream, void *dst, const void *src, size_t bytes);
int atk_get(void *stream, void *dst, const void *src, size_t bytes);
int atk_get_rows(void *stream, void *dst, size_t dpitch, const void *src, size_t width, size_t rows);

/* Elementwise work over n items of caller memory, pipelined: piece by piece
 * (at most `piece` items), the input bytes go to d_in on `s`, run(ctx, s, i0,
 * cnt) queues the piece's kernels on `s` (items [i0, i0 + cnt): d_in + i0 *
 * in_bpi in, d_out + i0 * out_bpi out), and the output comes back on `s2`, so
 * one piece's copy out overlaps the next piece's copy in. d_in and d_out hold
 * all n items. Returns once out holds every output byte; nonzero on an error
 * (from run, or a device error). */
typedef int (*atk_run_fn)(void *ctx, void *s, size_t i0, size_t cnt);
int atk_pipe(void *s, void *s2, const void *in, size_t in_bpi, void *d_in, void *out, size_t out_bpi,
             const void *d_out, size_t n, size_t piece, atk_run_fn run, void *ctx);

/* Work whose input is one run of caller bytes and whose output comes back in
 * ranges, each its own set of host rows (a filterbank call: every range of
 * hops is a column block of the channel-major output). For r = 0 .. nr-1,
 * desc(ctx, r, &g) says how many input bytes range r needs on the device
 * (g.need, non-decreasing), where its packed output sits (g.src, at most one
 * staging piece, ATK_PIPE_PIECE bytes; ranges r and r+2 may share a buffer,
 * never r and r+1) and where its rows go (g.dst, g.rows rows of g.width
 * bytes, g.dpitch apart); run(ctx, s, r) queues its kernels on `s`. The input
 * goes to d_in on `s` piece by piece as the ranges need it, range r's output
 * comes back on `s2` while range r+1's input goes in and its kernels run, and
 * the host writes range r-1's rows meanwhile. Returns once every row is
 * written and all in_bytes are on the device (in `s` order); nonzero on an
 * error (from desc or run, or a device error). Without staging, it falls back
 * to one copy in, every range's kernels, and each range's rows out in turn. */
#define ATK_PIPE_PIECE ((size_t)4 << 20)
typedef struct {
    size_t need;
    const void *src;
    char *dst;
    size_t dpitch, width, rows;
} atk_range;
typedef int (*atk_desc_fn)(void *ctx, size_t r, atk_range *g);
typedef int (*atk_range_fn)(void *ctx, void *s, size_t r);
int atk_pipe_ranges(void *s, void *s2, void *d_in, const void *in, size_t in_bytes, size_t nr, atk_desc_fn desc,
                    atk_range_fn run, void *ctx);

/* channel records strongest first, with the reference's comparator and qsort */
void atk_sort_channels(atkdsp_channel *c, size_t n);

/* CFAR over frames that live on the device (cfar_run.c): at most atk_cfar_slab
 * frames, each a row of atk_cfar_nbins dB values, copied into the handle's
 * own staging by atk_cfar_stage_rows, then tested and tracked by
 * atk_cfar_process_staged as atkdsp_cfar_process would */
size_t    atk_cfar_nbins(const atkdsp_cfar *c);
size_t    atk_cfar_slab(const atkdsp_cfar *c);
int       atk_cfar_stage_rows(atkdsp_cfar *c, const float *const *rows, size_t n);
ptrdiff_t atk_cfar_process_staged(atkdsp_cfar *c, size_t n, double frame_t0, double frame_dt,
                                  atkdsp_cfar_event *ev, size_t cap);

/* ln P(a, x) and ln Q(a, x), the regularized incomplete gamma (cfar_host.c) */
void atk_cfar_lgam_pq(double a, double x, double *lp, double *lq);

#endif
