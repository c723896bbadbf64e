/* The block ring (reference include/atkdsp.h:2020-2107, src/ring.c): one
 * producer that never waits, up to 16 consumers with their own cursors,
 * fixed-size blocks each with a header, a seqlock per block, all in caller
 * memory so that another process can attach to it.
 *
 * What is shared with the reference is the memory layout and the lock
 * protocol, the contract between processes and languages:
 *   0            header, 256 bytes: magic "ATKRING1", version 1, the header
 *                size, block/slot/total sizes, the consumer count, the next
 *                seq to commit (offset 64), the producer's claimed flag (72)
 *                and the torn-read counter (128)
 *   256          max_consumers records of 64 bytes: active, cursor, lost,
 *                reads, the seq a peek holds (+1; 0 = none)
 *   256 + 64 c   nblocks slots: the 8-byte lock word, the 56-byte block
 *                header, then block_bytes (a multiple of 64) of data
 * A slot's lock word is 2 seq + 1 while block seq is written into it and
 * 2 seq + 2 once that block is committed.
 *
 * How this library uses it. Every call first takes a view of the ring (the
 * header and where the slots start). A consumer's next block is found by
 * locate(): blocks that left the ring before the consumer got to them are
 * counted as lost and stepped over. A block is then taken by take(), which
 * reads the block's header and (for a copy) its bytes between two reads of
 * its lock word: equal to 2 seq + 2 both times means nothing was rewritten
 * meanwhile; otherwise the producer has lapped the reader mid-copy, the block
 * counts as lost (and torn) and the next one is tried. Peek hands out the
 * slot itself and release() makes the same check after the caller's use.
 * Shared words are accessed with the __atomic builtins on plain fields, so
 * the layout stays a C struct. */
#include <stddef.h>
#include <stdio.h>
#include <string.h>
#include "internal.h"

#define RING_MAGIC 0x31474E49524B5441ULL       /* "ATKRING1" read as a little-endian word */
#define RING_VERSION 1u
#define RING_HDR 256u
#define RING_CONS 64u
#define RING_SLOT_HDR 64u

typedef struct {
    uint64_t magic;
    uint32_t version, header_bytes;
    uint64_t block_bytes, nblocks, slot_bytes, total_bytes;
    uint32_t max_consumers, pad0;
    uint64_t pad1;
    uint64_t write_seq;         /* the next seq to commit = blocks committed so far */
    uint32_t claimed;
    uint32_t pad2;
    uint64_t pad3[6];
    uint64_t torn;
    uint64_t pad4[15];
} ring_header;

typedef struct {
    uint32_t active, pad;
    uint64_t cursor, lost, nread, peek;
    uint64_t pad2[3];
} ring_consumer;

typedef struct {
    uint64_t lock;
    atkdsp_ring_meta meta;
} ring_slot_hdr;

_Static_assert(sizeof(ring_header) == RING_HDR, "ring header layout");
_Static_assert(offsetof(ring_header, write_seq) == 64 && offsetof(ring_header, claimed) == 72 &&
                   offsetof(ring_header, torn) == 128, "ring header offsets");
_Static_assert(sizeof(ring_consumer) == RING_CONS && offsetof(ring_consumer, cursor) == 8 &&
                   offsetof(ring_consumer, peek) == 32, "ring consumer layout");
_Static_assert(sizeof(ring_slot_hdr) == RING_SLOT_HDR && sizeof(atkdsp_ring_meta) == 56, "ring slot layout");

#define LOAD(p, o) __atomic_load_n((p), (o))
#define STORE(p, v, o) __atomic_store_n((p), (v), (o))
#define ACQ __ATOMIC_ACQUIRE
#define REL __ATOMIC_RELEASE
#define RLX __ATOMIC_RELAXED

/* the lock word of a slot that holds committed block s */
static inline uint64_t committed_lock(uint64_t s) { return 2 * s + 2; }

typedef struct {
    ring_header *h;
    ring_consumer *cons;        /* max_consumers records */
    unsigned char *slots;       /* slot 0 */
    uint64_t nblk, stride;      /* slot count and bytes per slot */
} ring_view;

static ring_view view_of(const atkdsp_ring *r) {
    ring_view v;
    v.h = (ring_header *)(void *)r;
    v.cons = (ring_consumer *)(void *)((char *)(void *)r + RING_HDR);
    v.slots = (unsigned char *)(void *)r + RING_HDR + (size_t)v.h->max_consumers * RING_CONS;
    v.nblk = v.h->nblocks;
    v.stride = v.h->slot_bytes;
    return v;
}

static ring_slot_hdr *slot_at(const ring_view *v, uint64_t s) {
    return (ring_slot_hdr *)(void *)(v->slots + (size_t)(s % v->nblk) * (size_t)v->stride);
}

static unsigned char *payload_of(const ring_view *v, uint64_t s) {
    return (unsigned char *)slot_at(v, s) + RING_SLOT_HDR;
}

/* the oldest block still held when `top` blocks have been committed */
static uint64_t oldest_held(uint64_t top, uint64_t nblk) { return top > nblk ? top - nblk : 0; }

/* the record of an open consumer id, or NULL */
static ring_consumer *open_consumer(const ring_view *v, int id) {
    if (id < 0 || (uint32_t)id >= v->h->max_consumers) return NULL;
    ring_consumer *c = &v->cons[id];
    return LOAD(&c->active, ACQ) ? c : NULL;
}

size_t atkdsp_ring_bytes(size_t block_bytes, size_t nblocks, int max_consumers) {
    if (!block_bytes || nblocks < 2 || max_consumers < 1 || max_consumers > ATKDSP_RING_MAX_CONSUMERS) return 0;
    size_t data, slot, slots, total;
    if (__builtin_add_overflow(block_bytes, (size_t)63, &data)) return 0;
    data &= ~(size_t)63;
    if (__builtin_add_overflow(data, (size_t)RING_SLOT_HDR, &slot) || __builtin_mul_overflow(slot, nblocks, &slots) ||
        __builtin_add_overflow(slots, (size_t)RING_HDR + (size_t)RING_CONS * (size_t)max_consumers, &total))
        return 0;
    return total;
}

atkdsp_ring *atkdsp_ring_init(void *mem, size_t mem_bytes, size_t block_bytes, size_t nblocks, int max_consumers) {
    const size_t total = atkdsp_ring_bytes(block_bytes, nblocks, max_consumers);
    if (!total || !mem || mem_bytes < total || ((uintptr_t)mem % 8)) return NULL;
    ring_header *h = (ring_header *)mem;
    /* header and consumer records from zero; the payload is left as it is */
    memset(mem, 0, (size_t)RING_HDR + (size_t)RING_CONS * (size_t)max_consumers);
    const size_t data = (block_bytes + 63) & ~(size_t)63;
    h->version = RING_VERSION;
    h->header_bytes = RING_HDR;
    h->max_consumers = (uint32_t)max_consumers;
    h->block_bytes = data;
    h->slot_bytes = data + RING_SLOT_HDR;
    h->nblocks = nblocks;
    h->total_bytes = total;
    const ring_view v = view_of((atkdsp_ring *)mem);
    for (uint64_t s = 0; s < v.nblk; ++s) memset(slot_at(&v, s), 0, RING_SLOT_HDR);
    /* published last: a process that sees the magic sees a formatted ring */
    STORE(&h->magic, RING_MAGIC, REL);
    return (atkdsp_ring *)mem;
}

atkdsp_ring *atkdsp_ring_attach(void *mem, size_t mem_bytes) {
    if (!mem || ((uintptr_t)mem % 8) || mem_bytes < RING_HDR) return NULL;
    const ring_header *h = (const ring_header *)mem;
    if (LOAD(&h->magic, ACQ) != RING_MAGIC || h->version != RING_VERSION || h->header_bytes != RING_HDR) return NULL;
    /* the sizes must be exactly the ones init would have written for them,
     * and fit the memory given, before anything beyond the header is used */
    const uint64_t bb = h->block_bytes;
    if (!bb || bb % 64 || h->slot_bytes != bb + RING_SLOT_HDR) return NULL;
    const size_t total = atkdsp_ring_bytes((size_t)bb, (size_t)h->nblocks, (int)h->max_consumers);
    if (!total || h->max_consumers > ATKDSP_RING_MAX_CONSUMERS || h->total_bytes != total || total > mem_bytes)
        return NULL;
    return (atkdsp_ring *)mem;
}

/* ---- producer ---------------------------------------------------------------------- */
void *atkdsp_ring_claim(atkdsp_ring *r) {
    if (!r) return NULL;
    const ring_view v = view_of(r);
    const uint64_t s = LOAD(&v.h->write_seq, RLX);
    STORE(&slot_at(&v, s)->lock, committed_lock(s) - 1, RLX);
    /* the odd lock word is visible before any byte the caller writes next */
    __atomic_thread_fence(REL);
    v.h->claimed = 1;
    return payload_of(&v, s);
}

int atkdsp_ring_commit(atkdsp_ring *r, const atkdsp_ring_meta *m) {
    if (!r || !m) return ATKDSP_E_ARG;
    const ring_view v = view_of(r);
    if (!v.h->claimed || !m->nbytes || m->nbytes > v.h->block_bytes) return ATKDSP_E_ARG;
    const uint64_t s = LOAD(&v.h->write_seq, RLX);
    ring_slot_hdr *sl = slot_at(&v, s);
    atkdsp_ring_meta hd = *m;
    hd.seq = s;
    sl->meta = hd;
    STORE(&sl->lock, committed_lock(s), REL);
    STORE(&v.h->write_seq, s + 1, REL);
    v.h->claimed = 0;
    return ATKDSP_OK;
}

/* ---- consumers --------------------------------------------------------------------- */
int atkdsp_ring_open(atkdsp_ring *r, int from_oldest) {
    if (!r) return ATKDSP_E_ARG;
    const ring_view v = view_of(r);
    for (uint32_t id = 0; id < v.h->max_consumers; ++id) {
        ring_consumer *c = &v.cons[id];
        uint32_t expect = 0;
        if (!__atomic_compare_exchange_n(&c->active, &expect, 1u, 0, __ATOMIC_ACQ_REL, ACQ)) continue;
        const uint64_t top = LOAD(&v.h->write_seq, ACQ);
        STORE(&c->peek, 0, RLX);
        STORE(&c->nread, 0, RLX);
        STORE(&c->lost, 0, RLX);
        STORE(&c->cursor, from_oldest ? oldest_held(top, v.nblk) : top, REL);
        return (int)id;
    }
    return ATKDSP_E_CAP;
}

int atkdsp_ring_close(atkdsp_ring *r, int id) {
    if (!r) return ATKDSP_E_ARG;
    const ring_view v = view_of(r);
    if (id < 0 || (uint32_t)id >= v.h->max_consumers) return ATKDSP_E_ARG;
    STORE(&v.cons[id].active, 0u, REL);
    return ATKDSP_OK;
}

/* The next block consumer c can take: 1 with its seq, or 0 when none is
 * waiting. Blocks the ring no longer holds, or whose slot was rewritten before
 * the consumer got there, are added to *skipped and the cursor steps past them. */
static int locate(const ring_view *v, ring_consumer *c, uint64_t *seq, uint64_t *skipped) {
    const uint64_t start = LOAD(&c->cursor, RLX);
    uint64_t cur = start;
    int found = 0;
    for (;;) {
        const uint64_t top = LOAD(&v->h->write_seq, ACQ);
        if (cur >= top) break;
        const uint64_t floor = oldest_held(top, v->nblk);
        if (cur < floor) {
            *skipped += floor - cur;
            cur = floor;
        }
        const uint64_t lk = LOAD(&slot_at(v, cur)->lock, ACQ);
        if (lk == committed_lock(cur)) {
            found = 1;
            break;
        }
        if (lk < committed_lock(cur)) break;     /* not yet there (cannot happen below top) */
        *skipped += 1;                           /* already holds a later block */
        ++cur;
    }
    if (cur != start) STORE(&c->cursor, cur, REL);
    *seq = cur;
    return found;
}

static void note_lost(ring_consumer *c, uint64_t n, uint64_t *out) {
    if (n) STORE(&c->lost, LOAD(&c->lost, RLX) + n, REL);
    if (out) *out = n;
}

static void note_torn(const ring_view *v, ring_consumer *c, uint64_t s, uint64_t *skipped) {
    __atomic_fetch_add(&v->h->torn, 1, RLX);
    *skipped += 1;
    STORE(&c->cursor, s + 1, REL);
}

enum { TAKE_COPY, TAKE_PEEK };

/* Take consumer c's next block: TAKE_COPY copies it to out (up to cap bytes;
 * a longer block is left in place and gives E_CAP), TAKE_PEEK marks it held
 * and returns its slot in *where. Returns the byte count, 0 when nothing is
 * waiting, or E_CAP. */
static ptrdiff_t take(const ring_view *v, ring_consumer *c, int mode, void *out, size_t cap, atkdsp_ring_meta *m,
                      uint64_t *lost, const void **where) {
    uint64_t skipped = 0, s;
    ptrdiff_t ret = 0;
    while (locate(v, c, &s, &skipped)) {
        ring_slot_hdr *sl = slot_at(v, s);
        const atkdsp_ring_meta hd = sl->meta;
        const int fits = mode == TAKE_PEEK || hd.nbytes <= cap;
        if (mode == TAKE_COPY && fits) memcpy(out, payload_of(v, s), hd.nbytes);
        __atomic_thread_fence(ACQ);
        if (LOAD(&sl->lock, RLX) != committed_lock(s)) {
            note_torn(v, c, s, &skipped);        /* lapped during the copy: try the next one */
            continue;
        }
        if (!fits) {
            ret = ATKDSP_E_CAP;                  /* intact but too long: it stays for a bigger buffer */
        } else if (mode == TAKE_PEEK) {
            STORE(&c->peek, s + 1, REL);
            *where = payload_of(v, s);
            ret = (ptrdiff_t)hd.nbytes;
        } else {
            STORE(&c->cursor, s + 1, REL);
            STORE(&c->nread, LOAD(&c->nread, RLX) + 1, REL);
            ret = (ptrdiff_t)hd.nbytes;
        }
        if (m && ret > 0) *m = hd;
        break;
    }
    note_lost(c, skipped, lost);
    return ret;
}

ptrdiff_t atkdsp_ring_read(atkdsp_ring *r, int id, void *out, size_t cap, atkdsp_ring_meta *m, uint64_t *lost) {
    if (!r) return ATKDSP_E_ARG;
    const ring_view v = view_of(r);
    ring_consumer *c = open_consumer(&v, id);
    if (!c || (cap && !out)) return ATKDSP_E_ARG;
    return take(&v, c, TAKE_COPY, out, cap, m, lost, NULL);
}

const void *atkdsp_ring_peek(atkdsp_ring *r, int id, atkdsp_ring_meta *m, uint64_t *lost) {
    if (!r) return NULL;
    const ring_view v = view_of(r);
    ring_consumer *c = open_consumer(&v, id);
    if (!c) return NULL;
    const void *where = NULL;
    return take(&v, c, TAKE_PEEK, NULL, 0, m, lost, &where) > 0 ? where : NULL;
}

int atkdsp_ring_release(atkdsp_ring *r, int id) {
    if (!r) return ATKDSP_E_ARG;
    const ring_view v = view_of(r);
    ring_consumer *c = open_consumer(&v, id);
    const uint64_t held = c ? LOAD(&c->peek, RLX) : 0;
    if (!held) return ATKDSP_E_ARG;
    const uint64_t s = held - 1;
    STORE(&c->peek, 0, RLX);
    /* the caller's reads of the slot come before this check */
    __atomic_thread_fence(ACQ);
    const int intact = LOAD(&slot_at(&v, s)->lock, RLX) == committed_lock(s);
    STORE(&c->cursor, s + 1, REL);
    if (!intact) {
        __atomic_fetch_add(&v.h->torn, 1, RLX);
        note_lost(c, 1, NULL);
        return ATKDSP_E_OVERRUN;
    }
    STORE(&c->nread, LOAD(&c->nread, RLX) + 1, REL);
    return ATKDSP_OK;
}

ptrdiff_t atkdsp_ring_copy_range(const atkdsp_ring *r, uint64_t seq0, uint64_t seq1, void *out, size_t cap,
                                 atkdsp_ring_meta *metas, size_t mcap) {
    if (!r || (cap && !out) || (mcap && !metas)) return ATKDSP_E_ARG;
    const ring_view v = view_of(r);
    const uint64_t top = LOAD(&v.h->write_seq, ACQ);
    const uint64_t lo = seq0 > oldest_held(top, v.nblk) ? seq0 : oldest_held(top, v.nblk);
    const uint64_t hi = seq1 < top ? seq1 : top;
    unsigned char *dst = (unsigned char *)out;
    size_t used = 0, count = 0;
    for (uint64_t s = lo; s < hi && count < mcap; ++s) {
        ring_slot_hdr *sl = slot_at(&v, s);
        if (LOAD(&sl->lock, ACQ) != committed_lock(s)) continue;    /* already rewritten */
        const atkdsp_ring_meta hd = sl->meta;
        if (hd.nbytes > cap - used) break;                          /* the byte budget is spent */
        memcpy(dst + used, payload_of(&v, s), hd.nbytes);
        __atomic_thread_fence(ACQ);
        if (LOAD(&sl->lock, RLX) != committed_lock(s)) {            /* rewritten during the copy */
            __atomic_fetch_add(&v.h->torn, 1, RLX);
            continue;
        }
        metas[count++] = hd;
        used += hd.nbytes;
    }
    return (ptrdiff_t)count;
}

int atkdsp_ring_stats(const atkdsp_ring *r, atkdsp_ring_counters *st) {
    if (!r || !st) return ATKDSP_E_ARG;
    const ring_view v = view_of(r);
    const uint64_t top = LOAD(&v.h->write_seq, ACQ);
    memset(st, 0, sizeof *st);
    st->committed = top;
    st->oldest = oldest_held(top, v.nblk);
    st->block_bytes = v.h->block_bytes;
    st->nblocks = v.nblk;
    st->total_bytes = v.h->total_bytes;
    st->torn = LOAD(&v.h->torn, RLX);
    st->max_consumers = v.h->max_consumers;
    for (uint32_t id = 0; id < v.h->max_consumers; ++id) {
        const ring_consumer *c = &v.cons[id];
        if (!LOAD(&c->active, ACQ)) continue;
        const uint64_t cur = LOAD(&c->cursor, RLX);
        st->active += 1;
        st->active_mask |= 1u << id;
        st->read[id] = LOAD(&c->nread, RLX);
        st->lost[id] = LOAD(&c->lost, RLX);
        st->lag[id] = top > cur ? top - cur : 0;
    }
    return ATKDSP_OK;
}

int atkdsp_ring_describe(const atkdsp_ring *r, char *buf, size_t cap) {
    if (!r || !buf || !cap) return ATKDSP_E_ARG;
    atkdsp_ring_counters st;
    atkdsp_ring_stats(r, &st);
    const double mib = (double)st.nblocks * (double)st.block_bytes / (1024.0 * 1024.0);
    snprintf(buf, cap, "block ring %llu x %llu B = %.1f MiB; consumers open %u/%u; blocks committed %llu, "
             "held from %llu; torn reads %llu",
             (unsigned long long)st.nblocks, (unsigned long long)st.block_bytes, mib, st.active, st.max_consumers,
             (unsigned long long)st.committed, (unsigned long long)st.oldest, (unsigned long long)st.torn);
    return ATKDSP_OK;
}
