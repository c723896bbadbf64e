"""The CFAR oracle (oracle/oracle.py: cfar_tables, cfar_test, Cfar; reference
src/cfar.c:270-819) pinned against the reference build's events
(tests/golden/cfar_events.npz): the same eight scenarios as the GPU parity
test, replayed on the CPU with the reference's alpha, CA alpha and E[T];
every event, the counters and the last frame's hits. The reference-cell
table is also checked against its definition on small masked cases, by
brute force over every bin."""
import json

import numpy as np
import pytest

from conftest import golden, same_cfar_events
from oracle import oracle as O

G = golden("cfar_events")
META = json.loads(str(G["meta"]))


@pytest.mark.parametrize("m", META, ids=[m["name"] for m in META])
def test_oracle_events_match_the_reference(m):
    name = m["name"]
    frames = G[f"{name}_frames"].astype(np.float32)
    a, aca, et = G[f"{name}_info"]
    c = O.Cfar(m["nbins"], m["cfg"], a, aca, et)
    ev, evop = G[f"{name}_events"], G[f"{name}_evop"]
    for i, op in enumerate(m["ops"]):
        if op[0] == "process":
            got = c.process(frames[op[1]:op[1] + op[2]], op[1] * 0.01, 0.01)
        elif op[0] == "mask":
            c.set_mask(None if op[1] is None else G[f"{name}_{op[1]}"])
            got = np.zeros(0, O.CFAR_EVENT)
        elif op[0] == "reset":
            c.reset()
            got = np.zeros(0, O.CFAR_EVENT)
        else:
            got = c.flush()
        same_cfar_events(got, ev[evop == i], f"{name} op {i}")
    assert list(c.stats().values()) == [int(v) for v in G[f"{name}_stats"]]
    X, thr, hits = c.last
    la = G[f"{name}_last"]
    assert np.array_equal(hits, la[2] != 0) and np.array_equal(np.isnan(thr), np.isnan(la[1]))


def test_reference_cells_by_their_definition():
    rng = np.random.default_rng(3)
    for nbins, n_ref, n_guard in [(40, 3, 1), (24, 4, 0), (64, 2, 3)]:
        mask = rng.random(nbins) < 0.3
        cells, ok, sided = O.cfar_tables(nbins, n_ref, n_guard, mask)
        for i in range(nbins):
            free = [j for j in range(nbins) if not mask[j]]
            left = [j for j in reversed(free) if j < i - n_guard][:n_ref]
            right = [j for j in free if j > i + n_guard][:n_ref]
            far_r = [j for j in free if j > i + n_guard][n_ref:][:n_ref - len(left)]
            far_l = [j for j in reversed(free) if j < i - n_guard][n_ref:][:n_ref - len(right)]
            want = left + right + far_r + far_l
            if mask[i] or len(want) < 2 * n_ref:
                assert not ok[i]
                continue
            assert ok[i] and list(cells[i]) == want and sided[i] == (len(left) == len(right) == n_ref)


def _replay_all(mod):
    """Every scenario through a (possibly altered) oracle module; True when
    all of them match the reference's events, counters and hits."""
    try:
        for m in META:
            name = m["name"]
            frames = G[f"{name}_frames"].astype(np.float32)
            c = mod.Cfar(m["nbins"], m["cfg"], *G[f"{name}_info"])
            ev, evop = G[f"{name}_events"], G[f"{name}_evop"]
            for i, op in enumerate(m["ops"]):
                if op[0] == "process":
                    got = c.process(frames[op[1]:op[1] + op[2]], op[1] * 0.01, 0.01)
                elif op[0] == "mask":
                    c.set_mask(None if op[1] is None else G[f"{name}_{op[1]}"])
                    continue
                elif op[0] == "reset":
                    c.reset()
                    continue
                else:
                    got = c.flush()
                same_cfar_events(got, ev[evop == i], name)
            assert list(c.stats().values()) == [int(v) for v in G[f"{name}_stats"]]
    except AssertionError:
        return False
    return True


@pytest.mark.parametrize("old,new", [
    ("if best < 0 or ov > bov:", "if best < 0 or ov >= bov:"),                      # ties to the newer track
    ("lo_side = free[free < i - n_guard][::-1]", "lo_side = free[free < i - n_guard]"),        # far cells first
    ('if bin(t["hist"]).count("1") >= c["open_m"]:', 'if bin(t["hist"]).count("1") > c["open_m"]:'),
    ('t1 = t["t_hit_end"] - 0.5 * (k - 1) * t["dt_hit"]', 't1 = t["t_hit_end"] - 0.5 * k * t["dt_hit"]'),
    ("side = np.maximum(L, R) if mode == 2 else np.minimum(L, R)",
     "side = np.minimum(L, R) if mode == 2 else np.maximum(L, R)"),
    ("cuts = np.flatnonzero(np.diff(idx) - 1 > g) + 1", "cuts = np.flatnonzero(np.diff(idx) - 1 >= g) + 1"),
    ("for r in range(k):\n                            s = (F + 1 + r) % k", "for r in range(k):\n                            s = (F + k - r) % k"),
])
def test_planted_mistakes_fail_the_fixtures(old, new):
    """Plausible mistakes planted in a copy of the oracle on the CPU (ties to
    the newer track, the left cells taken farthest first, opening one hit
    late, the coarse t1 off by half a frame, GO and SO swapped, a merge gap
    one bin short, t0 searched newest first) each fail the reference's
    fixtures; the unaltered copy passes. (The order of borrowed cells only
    moves CA's sum by rounding; the brute-force table test pins it.)"""
    import types
    src = open(O.__file__).read()
    assert src.count(old) == 1, old
    for text, want in ((src, True), (src.replace(old, new), False)):
        mod = types.ModuleType("oracle_copy")
        exec(compile(text, O.__file__, "exec"), mod.__dict__)
        assert _replay_all(mod) is want
