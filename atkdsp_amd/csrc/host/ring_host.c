/* The block ring (reference include/atkdsp.h:2020-2107, src/ring.c): one
 * producer that never waits, up to 16 consumers with their own cursors,
 * fixed-size blocks each with a header, a seqlock per block, all in caller
 * memory so that another process can attach to it.
 *
 * The memory layout below is the reference's, offset for offset (it is the
 * contract between processes and languages):
 *   0            header, 256 bytes: magic "ATKRING1", version 1, the header
 *                size, block/slot/total sizes, the consumer count, the next
 *                seq to commit (offset 64), the producer's claimed flag (72)
 *                and the torn-read counter (128)
 *   256          max_consumers records of 64 bytes: active, cursor, lost,
 *                reads, the seq a peek holds (+1; 0 = none)
 *   256 + 64 c   nblocks slots: the 8-byte lock word, the 56-byte block
 *                header, then block_bytes (a multiple of 64) of data
 * A slot's lock is 2 seq + 1 while block seq is being written and 2 seq + 2
 * once committed. A reader takes the lock, copies, and takes it again: the
 * copy is good only when both reads are 2 s + 2; anything larger means the
 * producer lapped the reader, and the block counts as lost (torn).
 * Accesses to shared words are GCC/Clang __atomic builtins on the plain
 * fields, so the layout stays a C struct. */
#include <stddef.h>
#include <stdio.h>
#include <string.h>
#include "internal.h"

#define RING_MAGIC 0x31474E49524B5441ULL       /* "ATKRING1", little-endian */
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
    uint64_t write_seq;         /* the next seq to commit */
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

#define LD_ACQ(p) __atomic_load_n((p), __ATOMIC_ACQUIRE)
#define LD_RLX(p) __atomic_load_n((p), __ATOMIC_RELAXED)
#define ST_REL(p, v) __atomic_store_n((p), (v), __ATOMIC_RELEASE)
#define ST_RLX(p, v) __atomic_store_n((p), (v), __ATOMIC_RELAXED)

static ring_header *hdr_of(const atkdsp_ring *r) { return (ring_header *)(void *)r; }

static ring_consumer *cons_of(const atkdsp_ring *r, int i) {
    return (ring_consumer *)(void *)((char *)(void *)r + RING_HDR + (size_t)i * RING_CONS);
}

static ring_slot_hdr *slot_of(const atkdsp_ring *r, uint64_t seq) {
    const ring_header *h = hdr_of(r);
    const size_t first = RING_HDR + (size_t)h->max_consumers * RING_CONS;
    return (ring_slot_hdr *)(void *)((char *)(void *)r + first + (size_t)(seq % h->nblocks) * (size_t)h->slot_bytes);
}

static unsigned char *data_of(const atkdsp_ring *r, uint64_t seq) {
    return (unsigned char *)slot_of(r, seq) + RING_SLOT_HDR;
}

static size_t up64(size_t v) { return (v + 63) & ~(size_t)63; }

size_t atkdsp_ring_bytes(size_t block_bytes, size_t nblocks, int max_consumers) {
    if (block_bytes < 1 || nblocks < 2 || max_consumers < 1 || max_consumers > ATKDSP_RING_MAX_CONSUMERS) return 0;
    const size_t slot = RING_SLOT_HDR + up64(block_bytes);
    if (nblocks > (SIZE_MAX - 8192) / slot) return 0;
    return RING_HDR + RING_CONS * (size_t)max_consumers + nblocks * slot;
}

atkdsp_ring *atkdsp_ring_init(void *mem, size_t mem_bytes, size_t block_bytes, size_t nblocks, int max_consumers) {
    const size_t need = atkdsp_ring_bytes(block_bytes, nblocks, max_consumers);
    if (!mem || !need || mem_bytes < need || ((uintptr_t)mem & 7u)) return NULL;
    memset(mem, 0, RING_HDR + RING_CONS * (size_t)max_consumers);
    ring_header *h = (ring_header *)mem;
    h->version = RING_VERSION;
    h->header_bytes = RING_HDR;
    h->block_bytes = up64(block_bytes);
    h->nblocks = nblocks;
    h->slot_bytes = RING_SLOT_HDR + up64(block_bytes);
    h->total_bytes = need;
    h->max_consumers = (uint32_t)max_consumers;
    atkdsp_ring *r = (atkdsp_ring *)mem;
    for (size_t i = 0; i < nblocks; ++i) memset(slot_of(r, i), 0, RING_SLOT_HDR);
    /* the magic goes last: an attach that sees it sees a formatted ring */
    __atomic_thread_fence(__ATOMIC_RELEASE);
    ST_REL(&h->magic, RING_MAGIC);
    return r;
}

atkdsp_ring *atkdsp_ring_attach(void *mem, size_t mem_bytes) {
    if (!mem || mem_bytes < RING_HDR || ((uintptr_t)mem & 7u)) return NULL;
    const ring_header *h = (const ring_header *)mem;
    /* every size is checked before anything past the header is used */
    if (LD_ACQ(&h->magic) != RING_MAGIC || h->version != RING_VERSION || h->header_bytes != RING_HDR ||
        h->max_consumers < 1 || h->max_consumers > ATKDSP_RING_MAX_CONSUMERS || h->nblocks < 2 ||
        h->block_bytes < 64 || (h->block_bytes & 63u) || h->slot_bytes != RING_SLOT_HDR + h->block_bytes ||
        !h->total_bytes || h->total_bytes > mem_bytes ||
        h->total_bytes != atkdsp_ring_bytes((size_t)h->block_bytes, (size_t)h->nblocks, (int)h->max_consumers))
        return NULL;
    return (atkdsp_ring *)mem;
}

void *atkdsp_ring_claim(atkdsp_ring *r) {
    if (!r) return NULL;
    ring_header *h = hdr_of(r);
    const uint64_t s = LD_RLX(&h->write_seq);
    ST_RLX(&slot_of(r, s)->lock, 2 * s + 1);
    __atomic_thread_fence(__ATOMIC_RELEASE);      /* the odd lock before any data store */
    h->claimed = 1;
    return data_of(r, s);
}

int atkdsp_ring_commit(atkdsp_ring *r, const atkdsp_ring_meta *m) {
    if (!r || !m) return ATKDSP_E_ARG;
    ring_header *h = hdr_of(r);
    if (!h->claimed || m->nbytes < 1 || m->nbytes > h->block_bytes) return ATKDSP_E_ARG;
    const uint64_t s = LD_RLX(&h->write_seq);
    ring_slot_hdr *sl = slot_of(r, s);
    sl->meta = *m;
    sl->meta.seq = s;
    ST_REL(&sl->lock, 2 * s + 2);
    ST_REL(&h->write_seq, s + 1);
    h->claimed = 0;
    return ATKDSP_OK;
}

int atkdsp_ring_open(atkdsp_ring *r, int from_oldest) {
    if (!r) return ATKDSP_E_ARG;
    ring_header *h = hdr_of(r);
    for (int i = 0; i < (int)h->max_consumers; ++i) {
        ring_consumer *c = cons_of(r, i);
        uint32_t idle = 0;
        if (__atomic_compare_exchange_n(&c->active, &idle, 1u, 0, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE)) {
            const uint64_t w = LD_ACQ(&h->write_seq);
            const uint64_t oldest = w > h->nblocks ? w - h->nblocks : 0;
            ST_RLX(&c->lost, 0);
            ST_RLX(&c->nread, 0);
            ST_RLX(&c->peek, 0);
            ST_REL(&c->cursor, from_oldest ? oldest : w);
            return i;
        }
    }
    return ATKDSP_E_CAP;
}

int atkdsp_ring_close(atkdsp_ring *r, int id) {
    if (!r || id < 0 || id >= (int)hdr_of(r)->max_consumers) return ATKDSP_E_ARG;
    ST_REL(&cons_of(r, id)->active, 0u);
    return ATKDSP_OK;
}

static ring_consumer *active_consumer(atkdsp_ring *r, int id) {
    if (!r || id < 0 || id >= (int)hdr_of(r)->max_consumers) return NULL;
    ring_consumer *c = cons_of(r, id);
    return LD_ACQ(&c->active) ? c : NULL;
}

/* Advance c past the blocks that are gone; 1 with the next committed block's
 * seq and lock value, or 0 when there is nothing new. *lost counts the
 * blocks skipped (lapped, or caught being rewritten). */
static int next_committed(atkdsp_ring *r, ring_consumer *c, uint64_t *seq, uint64_t *lockv, uint64_t *lost) {
    ring_header *h = hdr_of(r);
    for (;;) {
        uint64_t s = LD_RLX(&c->cursor);
        const uint64_t w = LD_ACQ(&h->write_seq);
        if (s >= w) return 0;
        if (w - s > h->nblocks) {                  /* lapped: jump to the oldest held */
            *lost += w - h->nblocks - s;
            s = w - h->nblocks;
            ST_REL(&c->cursor, s);
        }
        const uint64_t v = LD_ACQ(&slot_of(r, s)->lock);
        if (v == 2 * s + 2) {
            *seq = s;
            *lockv = v;
            return 1;
        }
        if (v < 2 * s + 2) return 0;               /* still being written */
        *lost += 1;                                /* already rewritten */
        ST_REL(&c->cursor, s + 1);
    }
}

static void add_lost(ring_consumer *c, uint64_t n, uint64_t *out) {
    if (n) ST_REL(&c->lost, LD_RLX(&c->lost) + n);
    if (out) *out = n;
}

ptrdiff_t atkdsp_ring_read(atkdsp_ring *r, int id, void *out, size_t cap, atkdsp_ring_meta *m, uint64_t *lost) {
    ring_consumer *c = active_consumer(r, id);
    if (!c || (!out && cap)) return ATKDSP_E_ARG;
    uint64_t nlost = 0;
    for (;;) {
        uint64_t s, v;
        if (!next_committed(r, c, &s, &v, &nlost)) {
            add_lost(c, nlost, lost);
            return 0;
        }
        ring_slot_hdr *sl = slot_of(r, s);
        const atkdsp_ring_meta mm = sl->meta;
        const size_t nb = mm.nbytes;
        if (nb > cap) {
            __atomic_thread_fence(__ATOMIC_ACQUIRE);
            if (LD_RLX(&sl->lock) == v) {          /* a real block, too big: it stays */
                add_lost(c, nlost, lost);
                return ATKDSP_E_CAP;
            }
        } else {
            memcpy(out, data_of(r, s), nb);
            __atomic_thread_fence(__ATOMIC_ACQUIRE);
            if (LD_RLX(&sl->lock) == v) {          /* not overwritten during the copy */
                ST_REL(&c->cursor, s + 1);
                ST_REL(&c->nread, LD_RLX(&c->nread) + 1);
                add_lost(c, nlost, lost);
                if (m) *m = mm;
                return (ptrdiff_t)nb;
            }
        }
        __atomic_fetch_add(&hdr_of(r)->torn, 1, __ATOMIC_RELAXED);
        nlost += 1;
        ST_REL(&c->cursor, s + 1);
    }
}

const void *atkdsp_ring_peek(atkdsp_ring *r, int id, atkdsp_ring_meta *m, uint64_t *lost) {
    ring_consumer *c = active_consumer(r, id);
    if (!c) return NULL;
    uint64_t nlost = 0;
    for (;;) {
        uint64_t s, v;
        if (!next_committed(r, c, &s, &v, &nlost)) {
            add_lost(c, nlost, lost);
            return NULL;
        }
        ring_slot_hdr *sl = slot_of(r, s);
        const atkdsp_ring_meta mm = sl->meta;
        __atomic_thread_fence(__ATOMIC_ACQUIRE);
        if (LD_RLX(&sl->lock) == v) {
            ST_REL(&c->peek, s + 1);
            add_lost(c, nlost, lost);
            if (m) *m = mm;
            return data_of(r, s);
        }
        __atomic_fetch_add(&hdr_of(r)->torn, 1, __ATOMIC_RELAXED);
        nlost += 1;
        ST_REL(&c->cursor, s + 1);
    }
}

int atkdsp_ring_release(atkdsp_ring *r, int id) {
    ring_consumer *c = active_consumer(r, id);
    if (!c) return ATKDSP_E_ARG;
    const uint64_t held = LD_RLX(&c->peek);
    if (!held) return ATKDSP_E_ARG;
    const uint64_t s = held - 1;
    ST_RLX(&c->peek, 0);
    __atomic_thread_fence(__ATOMIC_ACQUIRE);
    const uint64_t v = LD_RLX(&slot_of(r, s)->lock);
    ST_REL(&c->cursor, s + 1);
    if (v == 2 * s + 2) {
        ST_REL(&c->nread, LD_RLX(&c->nread) + 1);
        return ATKDSP_OK;
    }
    __atomic_fetch_add(&hdr_of(r)->torn, 1, __ATOMIC_RELAXED);
    ST_REL(&c->lost, LD_RLX(&c->lost) + 1);
    return ATKDSP_E_OVERRUN;
}

ptrdiff_t atkdsp_ring_copy_range(const atkdsp_ring *r, uint64_t seq0, uint64_t seq1, void *out, size_t cap,
                                 atkdsp_ring_meta *metas, size_t mcap) {
    if (!r || (!out && cap) || (!metas && mcap)) return ATKDSP_E_ARG;
    ring_header *h = hdr_of(r);
    const uint64_t w = LD_ACQ(&h->write_seq);
    const uint64_t oldest = w > h->nblocks ? w - h->nblocks : 0;
    const uint64_t end = seq1 < w ? seq1 : w;
    size_t used = 0, nm = 0;
    for (uint64_t s = seq0 > oldest ? seq0 : oldest; s < end; ++s) {
        ring_slot_hdr *sl = slot_of(r, s);
        const uint64_t v = LD_ACQ(&sl->lock);
        if (v != 2 * s + 2) continue;              /* gone, or being rewritten */
        const atkdsp_ring_meta mm = sl->meta;
        const size_t nb = mm.nbytes;
        if (used + nb > cap || nm >= mcap) break;
        memcpy((unsigned char *)out + used, data_of(r, s), nb);
        __atomic_thread_fence(__ATOMIC_ACQUIRE);
        if (LD_RLX(&sl->lock) != v) {
            __atomic_fetch_add(&h->torn, 1, __ATOMIC_RELAXED);
            continue;
        }
        metas[nm++] = mm;
        used += nb;
    }
    return (ptrdiff_t)nm;
}

int atkdsp_ring_stats(const atkdsp_ring *r, atkdsp_ring_counters *st) {
    if (!r || !st) return ATKDSP_E_ARG;
    ring_header *h = hdr_of(r);
    memset(st, 0, sizeof *st);
    const uint64_t w = LD_ACQ(&h->write_seq);
    st->committed = w;
    st->oldest = w > h->nblocks ? w - h->nblocks : 0;
    st->block_bytes = h->block_bytes;
    st->nblocks = h->nblocks;
    st->total_bytes = h->total_bytes;
    st->max_consumers = h->max_consumers;
    st->torn = LD_RLX(&h->torn);
    for (int i = 0; i < (int)h->max_consumers; ++i) {
        ring_consumer *c = cons_of(r, i);
        if (!LD_ACQ(&c->active)) continue;
        st->active++;
        st->active_mask |= 1u << i;
        st->lost[i] = LD_RLX(&c->lost);
        st->read[i] = LD_RLX(&c->nread);
        const uint64_t cur = LD_RLX(&c->cursor);
        st->lag[i] = w > cur ? w - cur : 0;
    }
    return ATKDSP_OK;
}

int atkdsp_ring_describe(const atkdsp_ring *r, char *buf, size_t cap) {
    if (!r || !buf || !cap) return ATKDSP_E_ARG;
    atkdsp_ring_counters st;
    atkdsp_ring_stats(r, &st);
    snprintf(buf, cap,
             "ring: %llu blocks x %llu bytes (%.1f MiB of data), %u of %u consumers, %llu committed, %llu torn reads",
             (unsigned long long)st.nblocks, (unsigned long long)st.block_bytes,
             (double)st.nblocks * (double)st.block_bytes / 1048576.0, st.active, st.max_consumers,
             (unsigned long long)st.committed, (unsigned long long)st.torn);
    return ATKDSP_OK;
}
