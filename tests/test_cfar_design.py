"""CFAR threshold design (reference include/atkdsp.h:1733-1860, src/cfar.c:23-268),
host code in this library: atkdsp_cfar_alpha, _n_eff, _window_rho2 and
_defaults, on the CPU.

Pinned three ways:
- closed forms at K = 1 look (the header's own): CA alpha = N (Pfa^(-1/N) - 1);
  OS Pfa = prod_{i<k} (N - i) / (N - i + alpha); GO and SO by Gandhi and
  Kassam (1988); E[T] = 1 for CA and 1/N for the smallest of N (OS, k = 1);
- the reference build's values (tests/golden/cfar.npz, 1000 cases over mode,
  cells, rate, looks, OS rank and Hann correlation): alpha within 1e-8 and
  E[T] within 5e-6, the accuracy of the reference's Simpson grid;
- OS with correlated cells, where the rate it is solved at falls to 1e-20 ..
  1e-26 and the reference's grid loses accuracy (up to 24 % in alpha): an
  independent integration with scipy's incomplete gamma and beta confirms
  that this library's alpha gives the target rate."""
import math

import numpy as np
import pytest

import atkdsp_amd as A
from conftest import golden

HANN = (4.0 / 9.0, 1.0 / 36.0, 0.0, 0.0)


def _solve(f, target):
    lo, hi = 1e-6, 1e30
    for _ in range(500):
        mid = math.sqrt(lo * hi)
        if f(mid) > target:
            lo = mid
        else:
            hi = mid
    return math.sqrt(lo * hi)


@pytest.mark.parametrize("n", [1, 2, 8, 16, 64])
@pytest.mark.parametrize("pfa", [1e-3, 1e-6, 1e-12])
def test_closed_forms_at_one_look(n, pfa):
    N = 2 * n
    a, et = A.cfar_alpha("ca", n, pfa)
    assert abs(a / (N * (pfa ** (-1.0 / N) - 1.0)) - 1.0) < 1e-12 and abs(et - 1.0) < 1e-9
    for k in sorted({1, (3 * N + 3) // 4, N}):
        a, et = A.cfar_alpha("os", n, pfa, 1.0, k)
        exact = _solve(lambda al: math.prod((N - i) / (N - i + al) for i in range(k)), pfa)
        assert abs(a / exact - 1.0) < 1e-11, (k, a, exact)
        if k == 1:
            assert abs(et * N - 1.0) < 1e-10                  # E[min of N unit exponentials] = 1/N

    def so(T):
        return 2.0 * sum(math.comb(n - 1 + i, i) * (2.0 + T / n) ** (-(n + i)) for i in range(n))

    assert abs(A.cfar_alpha("so", n, pfa)[0] / _solve(so, pfa) - 1.0) < 1e-11
    assert abs(A.cfar_alpha("go", n, pfa)[0] / _solve(lambda T: 2.0 * (1.0 + T / n) ** (-n) - so(T), pfa) - 1.0) < 1e-10


def _golden_cases():
    g = golden("cfar")
    return [(int(m), int(n), float(p), float(K), int(k), int(r), int(rc), float(a), float(e))
            for (m, n, p, K, k), r, rc, a, e in zip(g["c_args"], g["c_rho"], g["c_rc"], g["c_alpha"], g["c_et"])]


def test_agrees_with_the_reference_build():
    cases = _golden_cases()
    rng = np.random.default_rng(7)
    pick = [c for i, c in enumerate(cases) if c[0] != 1 or not c[5]]
    pick = [pick[i] for i in sorted(rng.choice(len(pick), 120, replace=False))]
    for mode, n, pfa, K, k, rho, rc, ar, er in pick:
        if rc:
            with pytest.raises(A.AtkDspError):
                A.cfar_alpha(mode, n, pfa, K, k, HANN if rho else None)
            continue
        a, e = A.cfar_alpha(mode, n, pfa, K, k, HANN if rho else None)
        assert abs(a / ar - 1.0) < 1e-8, (mode, n, pfa, K, k, rho, a, ar)
        assert abs(e / er - 1.0) < 5e-6, (mode, n, pfa, K, k, rho, e, er)


def test_os_with_correlated_cells_meets_its_rate():
    from scipy import integrate
    from scipy import special as sp
    worst = []
    for mode, n, pfa, K, k, rho, rc, ar, er in _golden_cases():
        if mode != 1 or not rho or rc or pfa > 1e-12 or n not in (2, 8):
            continue
        N = 2 * n
        kk = k or (3 * N + 3) // 4
        a_ca = A.cfar_alpha("ca", n, pfa, K, 0, HANN)[0]
        target = sp.betainc(N * K, K, 1.0 / (1.0 + a_ca / N))   # the independent-cell CA rate there
        a = A.cfar_alpha("os", n, pfa, K, k, HANN)[0]

        def g(u):
            x = math.exp(u)
            fx = math.exp(K * math.log(K) + (K - 1.0) * u - K * x - math.lgamma(K)) * x
            return fx * sp.betainc(kk, N - kk + 1, sp.gammainc(K, K * x / a))

        us = np.linspace(-80.0, 8.0, 1761)
        m = int(np.argmax([g(u) for u in us]))
        val = integrate.quad(g, us[max(m - 200, 0)], us[min(m + 200, len(us) - 1)], limit=500, epsabs=0,
                             epsrel=1e-12)[0]
        assert abs(val / target - 1.0) < 1e-6, (n, pfa, K, kk, val, target)
        worst.append(abs(a / ar - 1.0))
    assert worst and max(worst) > 1e-3          # cases where the reference's grid is the inaccurate one


def test_n_eff_window_rho2_and_defaults_are_the_references():
    g = golden("cfar")
    for n in range(1, 40):
        assert A.cfar_n_eff(n, HANN) == g["n_eff"][n - 1, 0] and A.cfar_n_eff(n) == g["n_eff"][n - 1, 1] == 2 * n
    for kind in (0, 1, 2, 3):
        for n in (1000, 4096):
            got = A.cfar_window_rho2(A.window(kind, n))
            assert np.allclose(got, g[f"rho2_{kind}_{n}"], rtol=1e-12, atol=1e-15), (kind, n)
    hann = A.cfar_window_rho2(A.window("hann", 4096))
    assert abs(hann[0] - 4.0 / 9.0) < 1e-3 and abs(hann[1] - 1.0 / 36.0) < 1e-3 and max(hann[2:]) < 1e-6
    assert max(A.cfar_window_rho2(np.ones(1024, np.float32))) < 1e-12      # rectangular: none
    lib = A.load()
    c = A._CfarConfig()
    lib.atkdsp_cfar_defaults(c)
    assert bytes(c) == g["defaults"].tobytes()


def test_arguments_are_refused_as_the_reference_does():
    for args in [(4, 16, 1e-6), (0, 0, 1e-6), (0, 257, 1e-6), (0, 16, 0.6), (0, 16, 1e-16), (0, 16, 1e-6, 0.1),
                 (1, 16, 1e-6, 1.0, 33), (0, 16, 1e-6, 1.0, 0, (1.0, 0, 0, 0))]:
        with pytest.raises(A.AtkDspError):
            A.cfar_alpha(*args)
    with pytest.raises(A.AtkDspError):
        A.cfar_window_rho2(np.zeros(8, np.float32))
    assert A.cfar_n_eff(0) == 0.0
