"""The display's single-FMA sample conversion against the unpack's form.

The chain's display reads raw samples and converts them with
fma((float)b, k, -c'k) where the unpack computes ((float)b - c') * k (reference
src/convert.c:44-70, contraction off). For a power-of-two k the two round to
the same float whenever c' is zero or c'k is a normal float. These tests check
that claim with exact rational arithmetic, rounding each exact value once to
float32, over every 8-bit code and a spread of 16-bit codes and offsets, and
check that the condition is needed (an underflowed c'k gives a different
float for some samples).
"""
from fractions import Fraction

import numpy as np
import pytest

F32_MIN = Fraction(2) ** -126


def rn32(q: Fraction) -> np.float32:
    """q rounded once to float32, to nearest, ties to even."""
    if q == 0:
        return np.float32(0.0)
    guess = np.float32(float(q))            # within one float32 step of the answer
    best = None
    for c in (np.nextafter(guess, np.float32(-np.inf)), guess, np.nextafter(guess, np.float32(np.inf))):
        if not np.isfinite(c):
            continue
        d = abs(Fraction(float(c)) - q)
        if best is None or d < best[0] or (d == best[0] and (int(c.view(np.uint32)) & 1) == 0):
            best = (d, c)
    return np.float32(best[1])


def unpack_form(b: int, c: np.float32, k: np.float32) -> np.float32:
    """((float)b - c') * k, each operation rounded to float32."""
    t = rn32(Fraction(int(b)) - Fraction(float(c)))
    return rn32(Fraction(float(t)) * Fraction(float(k)))


def fma_form(b: int, d: np.float32, k: np.float32) -> np.float32:
    """fma((float)b, k, -d): the exact b*k - d rounded once."""
    return rn32(Fraction(int(b)) * Fraction(float(k)) - Fraction(float(d)))


def fm_ok(c: np.float32, k: np.float32):
    """The kernel's condition: k a power of two, c' zero or c'k normal; d = c'k."""
    kb = int(np.float32(k).view(np.uint32))
    pow2 = (kb & 0x807FFFFF) == 0 and (kb >> 23) not in (0, 255)
    d = np.float32(c * k)
    return pow2 and (c == 0 or abs(Fraction(float(d))) >= F32_MIN), d


def offsets(rng):
    """c' = off/k for DC offsets across magnitudes, as the host forms it."""
    mags = [0.0, 1e-30, 1e-9, 3e-6, 1e-3, 0.02, 0.4, 0.99]
    out = []
    for m in mags:
        for s in (1.0, -1.0):
            out.append(np.float32(s * m * (1.0 + 0.37 * rng.random())))
    return out


@pytest.mark.parametrize("k,codes", [(np.float32(1.0 / 128.0), range(-128, 128)),
                                     (np.float32(1.0 / 2048.0), list(range(-2048, 2048, 7)) + [-2048, 2047]),
                                     (np.float32(1.0 / 32768.0), list(range(-32768, 32768, 251)) + [-32768, 32767])])
def test_fma_conversion_equals_unpack(k, codes):
    rng = np.random.default_rng(11)
    checked = 0
    for off in offsets(rng):
        c = np.float32(np.float32(0.0) + np.float32(off) / k)     # c' = c0 + off/k with c0 = 0
        ok, d = fm_ok(c, k)
        if not ok:
            continue
        for b in codes:
            u, f = unpack_form(b, c, k), fma_form(b, d, k)
            assert u.view(np.uint32) == f.view(np.uint32), (b, float(c), float(k), float(u), float(f))
            checked += 1
    assert checked > 1000


def test_condition_is_needed():
    """Where c'k rounds away (here it underflows to zero) the single FMA loses
    c': for b = 0 it gives +0 where the unpack's form gives -0. The kernel's
    condition refuses the FMA form for such a c' (nonzero, c'k not normal).
    (A c' = off/k the host forms from a float offset always passes: c'k is
    then the offset itself.)"""
    k = np.float32(2.0 ** -15)
    c = np.float32(2.0 ** -149)              # c'k = 2^-164 rounds to 0
    ok, d = fm_ok(c, k)
    assert not ok and d == 0
    u, f = unpack_form(0, c, k), fma_form(0, d, k)
    assert u == f == 0 and np.signbit(u) and not np.signbit(f)


def test_host_offsets_take_the_fma_form():
    """c' = off/k for any normal float offset gives c'k = off exactly."""
    rng = np.random.default_rng(5)
    for k in (np.float32(2.0 ** -7), np.float32(2.0 ** -11), np.float32(2.0 ** -15)):
        for off in np.float32(rng.standard_normal(200) * 10.0 ** rng.uniform(-30, 0, 200)):
            c = np.float32(off) / k
            ok, d = fm_ok(np.float32(c), k)
            assert ok and d == off


def test_rn32_rounds_once_to_nearest_even():
    # ties between two float32 neighbours go to the even one
    one = Fraction(1)
    ulp = Fraction(2) ** -23
    assert rn32(one + ulp / 2) == np.float32(1.0)
    assert rn32(one + ulp + ulp / 2) == np.nextafter(np.nextafter(np.float32(1.0), np.float32(2)), np.float32(2))
    assert rn32(Fraction(1, 3)) == np.float32(1.0) / np.float32(3.0)
