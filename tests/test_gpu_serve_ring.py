"""Served streams fed from a block ring (atkgpu_server_submit_ring): a ring
block becomes one chunk, copied once from the slot into staging. The output
must be bit-identical to submitting the same blocks directly; a lapped
consumer's lost blocks are reported and skipped; a block of the wrong format
is refused without being consumed; and with a producer writing into a small
ring at full speed beside a consumer that stalls now and then (so it is
lapped), every chunk taken must be an intact
block (the served output equals direct submission of exactly the blocks
taken, regenerated from their sequence numbers) and taken + lost must account
for every block."""
import threading
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FS = 40e6
CU8, CI8, CI16, CI16Q11, CF32 = 0, 1, 2, 3, 4
NAMES = {CU8: "cu8", CI8: "ci8", CI16: "ci16", CI16Q11: "ci16q11", CF32: "cf32"}


def _block(seq, fmt=CI16, nel=1):
    """the bytes of block seq: a deterministic function of seq, 1..8192 samples per element"""
    rng = np.random.default_rng(1000 + seq)
    n = int(rng.integers(1, 8193))
    if fmt in (CU8, CI8):
        return rng.integers(0, 256, size=2 * n * nel, dtype=np.uint8)
    if fmt in (CI16, CI16Q11):
        return rng.integers(-2048, 2048, size=2 * n * nel, dtype=np.int16).view(np.uint8)
    return rng.standard_normal(2 * n * nel).astype(np.float32).view(np.uint8)


def _server(nel=1, fmt="ci16", slots=256):
    from atkdsp_amd.serve import Server
    srv = Server(hbm_budget=2 << 30, max_streams=4, max_batch=16, staging_bytes=32 << 20)
    sid = srv.add_stream(fmt=fmt, fs=FS, nchan=256, oversample=2, taps_per_poly=12, chunk_max=8192,
                         out_slots=slots, nel=nel, coherent=nel > 1)
    return srv, sid


def _direct(blocks, nel=1, fmt="ci16"):
    srv, sid = _server(nel, fmt)
    ts = [srv.submit(sid, b) for b in blocks]
    srv.drain()
    return [srv.read(sid, t) for t in ts]


@pytest.mark.parametrize("code,nel", [(CU8, 1), (CI8, 1), (CI16, 1), (CI16Q11, 1), (CF32, 1), (CF32, 3), (CI16, 2)])
def test_ring_blocks_serve_like_direct_submission(gpu, code, nel):
    fmt = NAMES[code]
    blocks = [_block(s, code, nel) for s in range(12)]
    ring = gpu.Ring(8192 * 8 * nel, 16, 2)
    cons = ring.open()
    srv, sid = _server(nel, fmt)
    ts = []
    for s, b in enumerate(blocks):
        ring.write(b, first_sample=10_000 * s, format=code)
        t, meta, lost = srv.submit_ring(sid, ring, cons)
        assert t == s and lost == 0 and meta["seq"] == s and meta["first_sample"] == 10_000 * s
        ts.append(t)
    assert srv.submit_ring(sid, ring, cons) == (None, None, 0)
    srv.drain()
    ref = _direct(blocks, nel, fmt)
    for t, y in zip(ts, ref):
        assert np.array_equal(srv.read(sid, t), y), (fmt, t)
    st = ring.stats()
    assert st["read"][cons] == 12 and st["lost"][cons] == 0


def test_lapped_consumer_reports_lost_blocks_and_skips_them(gpu):
    ring = gpu.Ring(8192 * 4, 4, 1)
    cons = ring.open()
    blocks = [_block(s) for s in range(10)]
    for s, b in enumerate(blocks):
        ring.write(b, first_sample=s, format=CI16)
    srv, sid = _server()
    got = []
    t, meta, lost = srv.submit_ring(sid, ring, cons)
    assert lost == 6 and meta["seq"] == 6
    got.append(t)
    while True:
        t, meta, lost = srv.submit_ring(sid, ring, cons)
        if t is None:
            break
        assert lost == 0
        got.append(t)
    assert got == [0, 1, 2, 3]
    srv.drain()
    ref = _direct(blocks[6:])
    for t, y in zip(got, ref):
        assert np.array_equal(srv.read(sid, t), y)


def test_refused_blocks_stay_in_the_ring(gpu):
    ring = gpu.Ring(8192 * 8, 8, 2)
    cons = ring.open()
    srv, sid = _server()
    cf = _block(0, CF32)
    ring.write(cf, format=CF32)                                # not the stream's format
    with pytest.raises(gpu.AtkDspError):
        srv.submit_ring(sid, ring, cons)
    data, meta, lost = ring.read(cons)                         # still there, unconsumed
    assert meta["seq"] == 0 and np.array_equal(data, cf)
    big = np.zeros(4 * 9000, np.uint8)                         # 9000 samples > chunk_max
    ring.write(big, format=CI16)
    with pytest.raises(gpu.AtkDspError):
        srv.submit_ring(sid, ring, cons)
    assert ring.read(cons)[1]["seq"] == 1
    with pytest.raises(gpu.AtkDspError):                       # not an open consumer
        srv.submit_ring(sid, ring, 1)
    assert srv.submit_ring(sid, ring, cons) == (None, None, 0)


def test_concurrent_producer_every_taken_block_is_intact(gpu):
    total = 240
    ring = gpu.Ring(8192 * 4, 6, 1)
    cons = ring.open(from_oldest=True)
    srv, sid = _server(slots=256)
    done = threading.Event()
    blocks = [_block(s) for s in range(total)]

    def producer():
        for s in range(total):
            ring.write(blocks[s], first_sample=s, format=CI16)
        done.set()

    taken, lost_total = [], 0
    th = threading.Thread(target=producer)
    th.start()
    while True:
        t, meta, lost = srv.submit_ring(sid, ring, cons)
        lost_total += lost
        if t is None:
            if done.is_set():
                t, meta, lost = srv.submit_ring(sid, ring, cons)
                lost_total += lost
                if t is None:
                    break
            else:
                continue
        assert t == len(taken)
        taken.append(meta["seq"])
        if len(taken) % 8 == 0:
            time.sleep(0.002)                                  # let the producer lap this consumer
    th.join(timeout=60)
    assert taken == sorted(taken) and len(set(taken)) == len(taken)
    assert len(taken) + lost_total == total, (len(taken), lost_total)
    assert lost_total > 0 and taken
    srv.drain()
    ref = _direct([blocks[s] for s in taken])
    for t, y in enumerate(ref):
        assert np.array_equal(srv.read(sid, t), y), (t, taken[t])
    print(f"taken {len(taken)}, lost {lost_total}, torn {ring.stats()['torn']}")
