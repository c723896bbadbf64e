"""The block ring (reference include/atkdsp.h:2020-2107, src/ring.c), host code:
the reference's own stress test (tests/test_ring_stress.c, compiled unmodified
with the reference header) linked against this library; layout parity with
the reference build (a ring made by either library, after the same calls, is
the same bytes, and each attaches to the other's), the
producer that never waits (a lapped consumer is told exactly how many blocks
it lost), E_CAP on a short buffer, zero-copy peek/release with overrun
detection, copy_range, and a threaded stress run in which no torn block is
ever returned and every block is either read or counted lost."""
import ctypes as C
import subprocess
import threading
from pathlib import Path

import numpy as np
import pytest

import atkdsp_amd as A
from oracle import ref as R

needs_ref = pytest.mark.skipif(not R.available(), reason="oracle/_ref not built")
ROOT = Path(__file__).resolve().parents[1]
STRESS = ROOT / "oracle" / "_ref" / "ring_stress_ours"


@pytest.mark.skipif(not STRESS.exists(), reason="the reference stress test is not built (oracle/build_ref.sh)")
@pytest.mark.parametrize("args", [[], ["3", "160", "1024", "128"]], ids=["full_speed_2s", "160MBps_1MiB_blocks"])
def test_reference_stress_test_passes_on_this_library(args):
    """one producer, a fast, a slow (3x work) and a stalled consumer: every
    block read is whole and in order, and read + lost is exact per consumer
    (reference tests/test_ring_stress.c:1-18; the 160 MB/s case is the shape
    of bench/ring_bench_result.txt, 3 s instead of 600 s)"""
    ldd = subprocess.run(["ldd", str(STRESS)], capture_output=True, text=True, check=True).stdout
    assert str(ROOT / "atkdsp_amd" / "atkdsp.so") in ldd           # this library, not the reference's
    r = subprocess.run([str(STRESS), *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "RING STRESS: ok" in r.stdout, r.stdout + r.stderr


def _ref_lib():
    L = C.CDLL(str(R.REF_SO))
    vp, sz, mp = C.c_void_p, C.c_size_t, C.POINTER(A._RingMeta)
    sigs = {
        "atkdsp_ring_bytes": (sz, [sz, sz, C.c_int]),
        "atkdsp_ring_init": (vp, [vp, sz, sz, sz, C.c_int]),
        "atkdsp_ring_attach": (vp, [vp, sz]),
        "atkdsp_ring_claim": (vp, [vp]),
        "atkdsp_ring_commit": (C.c_int, [vp, mp]),
        "atkdsp_ring_open": (C.c_int, [vp, C.c_int]),
        "atkdsp_ring_read": (C.c_ssize_t, [vp, C.c_int, vp, sz, mp, C.POINTER(C.c_uint64)]),
    }
    for k, (res, args) in sigs.items():
        f = getattr(L, k)
        f.restype, f.argtypes = res, args
    return L


def _meta(nbytes, **kw):
    m = A._RingMeta(**kw)
    m.nbytes = nbytes
    return m


def _aligned(n):
    raw = np.zeros(n + 64, np.uint8)
    off = (-raw.ctypes.data) % 64
    return raw[off:off + n]


@needs_ref
def test_layout_is_the_references_byte_for_byte():
    L = _ref_lib()
    lib = A.load()
    need = lib.atkdsp_ring_bytes(300, 5, 3)
    assert need == L.atkdsp_ring_bytes(300, 5, 3) and need == 256 + 3 * 64 + 5 * (64 + 320)
    mine, theirs = _aligned(need), _aligned(need)
    rm = lib.atkdsp_ring_init(mine.ctypes.data, need, 300, 5, 3)
    rt = L.atkdsp_ring_init(theirs.ctypes.data, need, 300, 5, 3)
    assert rm and rt and np.array_equal(mine, theirs)
    # the same producer and consumer calls on both
    for lb, r in ((lib, rm), (L, rt)):
        lb.atkdsp_ring_open(r, 1)
        for k in range(7):                                   # laps the 5 slots
            p = lb.atkdsp_ring_claim(r)
            C.memmove(p, bytes([k]) * (100 + k), 100 + k)
            m = _meta(100 + k, first_sample=1000 * k, t_utc_ns=-5, fc_hz=1e8, fs_hz=2.4e6, gain_db=3.5,
                      format=2, flags=k & 7)
            assert lb.atkdsp_ring_commit(r, C.byref(m)) == 0
    assert np.array_equal(mine, theirs)


@needs_ref
def test_each_library_attaches_to_the_others_ring():
    L = _ref_lib()
    lib = A.load()
    # the reference makes and fills a ring; this library attaches and reads it
    need = L.atkdsp_ring_bytes(64, 4, 2)
    mem = _aligned(need)
    rt = L.atkdsp_ring_init(mem.ctypes.data, need, 64, 4, 2)
    p = L.atkdsp_ring_claim(rt)
    C.memmove(p, b"reference block", 15)
    m = _meta(15, first_sample=77, format=3)
    assert L.atkdsp_ring_commit(rt, C.byref(m)) == 0
    ring = A.Ring(mem=mem)
    cid = ring.open(from_oldest=True)
    data, meta, lost = ring.read(cid)
    assert bytes(data) == b"reference block" and meta["first_sample"] == 77 and meta["seq"] == 0 and lost == 0
    # and the other way round
    mine = A.Ring(64, 4, 2)
    mine.write(b"this library's block", first_sample=9, format=4)
    rr = L.atkdsp_ring_attach(mine.memory.ctypes.data, mine.memory.size)
    assert rr
    c = L.atkdsp_ring_open(rr, 1)
    out = np.zeros(64, np.uint8)
    got_meta = A._RingMeta()
    lost = C.c_uint64()
    n = L.atkdsp_ring_read(rr, c, out.ctypes.data, 64, C.byref(got_meta), C.byref(lost))
    assert n == 20 and bytes(out[:n]) == b"this library's block" and got_meta.format == 4


def test_attach_refuses_what_is_not_a_ring():
    ring = A.Ring(128, 4, 2)
    bad = ring.memory.copy()
    bad[0] ^= 1                                              # the magic
    with pytest.raises(A.AtkDspError):
        A.Ring(mem=bad)
    with pytest.raises(A.AtkDspError):
        A.Ring(mem=ring.memory[:256])                        # shorter than the ring
    with pytest.raises(A.AtkDspError):
        A.Ring(0, 4, 2)
    with pytest.raises(A.AtkDspError):                         # read-only memory
        A.Ring(mem=ring.memory.tobytes())
    assert A.ring_bytes(128, 1, 2) == 0 and A.ring_bytes(128, 4, 17) == 0


def test_a_lapped_consumer_is_told_what_it_lost():
    ring = A.Ring(64, 4, 3)
    slow = ring.open(from_oldest=False)
    for k in range(10):
        ring.write(bytes([k]) * 8, first_sample=k)
    got, lost_total = [], 0
    while True:
        data, meta, lost = ring.read(slow)
        lost_total += lost
        if data is None:
            break
        got.append(meta["seq"])
        assert bytes(data) == bytes([meta["seq"]]) * 8
    assert got == [6, 7, 8, 9] and lost_total == 6            # the ring holds the last 4
    st = ring.stats()
    assert st["committed"] == 10 and st["oldest"] == 6 and st["lost"][slow] == 6 and st["read"][slow] == 4
    late = ring.open(from_oldest=True)
    assert ring.read(late)[1]["seq"] == 6
    fresh = ring.open(from_oldest=False)
    assert ring.read(fresh) == (None, None, 0)
    with pytest.raises(A.AtkDspError):                        # all 3 consumers taken
        ring.open()
    ring.close_consumer(fresh)
    assert ring.open() == fresh


def test_short_buffer_is_e_cap_and_the_block_stays():
    lib = A.load()
    ring = A.Ring(256, 4, 1)
    c = ring.open()
    ring.write(np.arange(200, dtype=np.uint8))
    small = np.zeros(100, np.uint8)
    m = A._RingMeta()
    lost = C.c_uint64()
    assert lib.atkdsp_ring_read(ring.handle, c, small.ctypes.data, 100, C.byref(m), C.byref(lost)) == -3
    data, meta, _ = ring.read(c)                             # still there for a bigger buffer
    assert data.size == 200 and meta["seq"] == 0
    with pytest.raises(A.AtkDspError):                        # nbytes above the block
        ring.claim()
        ring.commit(257)


def test_peek_release_detects_an_overrun():
    ring = A.Ring(64, 2, 1)
    c = ring.open()
    ring.write(b"A" * 10)
    view, meta, _ = ring.peek(c)
    assert bytes(view) == b"A" * 10 and meta["seq"] == 0
    assert ring.release(c) is True
    ring.write(b"B" * 10)
    view, meta, _ = ring.peek(c)
    assert meta["seq"] == 1
    ring.write(b"C" * 10)                                    # seq 2, slot 0
    ring.write(b"D" * 10)                                    # seq 3 rewrites seq 1's slot
    assert ring.release(c) is False                          # discard what was computed
    st = ring.stats()
    assert st["lost"][c] == 1 and st["torn"] == 1
    data, meta, lost = ring.read(c)
    assert meta["seq"] == 2 and bytes(data) == b"C" * 10


def test_copy_range_cuts_whole_held_blocks():
    ring = A.Ring(64, 4, 1)
    for k in range(6):
        ring.write(bytes([65 + k]) * (10 + k), first_sample=100 * k)
    data, metas = ring.copy_range(0, 6)
    assert [m["seq"] for m in metas] == [2, 3, 4, 5]           # 0 and 1 are gone
    assert bytes(data) == b"".join(bytes([65 + k]) * (10 + k) for k in range(2, 6))
    data, metas = ring.copy_range(3, 5)
    assert [m["first_sample"] for m in metas] == [300, 400]
    _, metas = ring.copy_range(0, 6, cap=24)                  # stops at the byte budget
    assert [m["seq"] for m in metas] == [2]
    _, metas = ring.copy_range(0, 6, cap=25)                  # 12 + 13 bytes fit exactly
    assert [m["seq"] for m in metas] == [2, 3]
    _, metas = ring.copy_range(0, 6, mcap=3)                  # and at the meta budget
    assert [m["seq"] for m in metas] == [2, 3, 4]


def test_threaded_stress_never_returns_a_torn_block():
    """A producer commits blocks whose bytes are a function of their seq as
    fast as it can into a small ring; three readers copy concurrently. Every
    returned block must be intact, and after draining, read + lost accounts
    for every block each reader was open for."""
    nblocks, total, bb = 8, 20000, 512
    ring = A.Ring(bb, nblocks, 4)
    lib = A.load()
    ids = [ring.open(from_oldest=True) for _ in range(3)]
    errors = []
    counts = {i: [0, 0] for i in ids}
    done = threading.Event()

    def pattern(seq):
        return np.full(bb // 8, seq, np.uint64).view(np.uint8)

    def producer():
        for s in range(total):
            src = pattern(s)                                 # held across the copy
            p = lib.atkdsp_ring_claim(ring.handle)
            C.memmove(p, src.ctypes.data, bb)
            m = _meta(bb)
            lib.atkdsp_ring_commit(ring.handle, C.byref(m))
        done.set()

    def reader(cid):
        buf = np.zeros(bb, np.uint8)
        m, lost = A._RingMeta(), C.c_uint64()
        while True:
            n = lib.atkdsp_ring_read(ring.handle, cid, buf.ctypes.data, bb, C.byref(m), C.byref(lost))
            counts[cid][1] += lost.value
            if n > 0:
                counts[cid][0] += 1
                if not np.array_equal(buf.view(np.uint64), np.full(bb // 8, m.seq, np.uint64)):
                    errors.append((cid, m.seq))
            elif done.is_set():
                n = lib.atkdsp_ring_read(ring.handle, cid, buf.ctypes.data, bb, C.byref(m), C.byref(lost))
                counts[cid][1] += lost.value
                if n <= 0:
                    break
                counts[cid][0] += 1

    ts = [threading.Thread(target=reader, args=(i,)) for i in ids] + [threading.Thread(target=producer)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=120)
    assert not errors, errors[:5]
    for cid in ids:
        assert sum(counts[cid]) == total, (cid, counts[cid])
    st = ring.stats()
    assert st["committed"] == total
    assert "torn reads" in ring.describe()
