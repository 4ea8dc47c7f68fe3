"""Oracle restatements downstream of eps^-1: the screened interaction
W = v^1/2 eps^-1 v^1/2 (correlation.f:868-877) and the RPA
correlation-energy integrand sum(log e + 1 - e) (correlation.f:686-690),
pinned by closed forms and by independent algorithms."""
import math

import numpy as np

import oracle


def _hpd(g, nb, lo=0.5, hi=4.0):
    q, _ = np.linalg.qr(g.standard_normal((nb, nb)) + 1j * g.standard_normal((nb, nb)))
    lam = g.uniform(lo, hi, nb)
    return (q * lam) @ q.conj().T, lam


def test_rpa_integrand_identity_is_zero():
    assert oracle.rpa_integrand(oracle.pack_upper(np.eye(7, dtype=complex)), 1.0) == 0.0


def test_rpa_integrand_diagonal_closed_form():
    d = np.array([1.5, 0.25, 3.0, 1.0, 7.5])
    want = sum(math.log(x) + 1.0 - x for x in d) * 0.125
    got = oracle.rpa_integrand(oracle.pack_upper(np.diag(d).astype(complex)), 0.125)
    assert abs(got - want) < 1e-13 * abs(want)


def test_rpa_integrand_2x2_eigenvalues_in_closed_form():
    a, c, b = 2.0, 0.75, 0.4 - 0.3j
    m = (a + c) / 2
    r = math.sqrt(((a - c) / 2) ** 2 + abs(b) ** 2)
    want = sum(math.log(e) + 1 - e for e in (m - r, m + r))
    got = oracle.rpa_integrand(oracle.pack_upper(np.array([[a, b], [np.conj(b), c]])), 1.0)
    assert abs(got - want) < 1e-14


def test_rpa_integrand_matches_log_determinant_and_trace():
    """sum log e = log det (LU, an independent algorithm); sum e = trace."""
    g = np.random.default_rng(3)
    e, lam = _hpd(g, 40)
    sign, logdet = np.linalg.slogdet(e)
    assert sign.real > 0
    want = logdet + 40 - np.trace(e).real
    got = oracle.rpa_integrand(oracle.pack_upper(e), 1.0)
    assert abs(got - want) < 1e-11 * abs(want)
    assert abs(got - np.sum(np.log(lam) + 1 - lam)) < 1e-11 * abs(want)


def test_screened_w_diagonal_closed_form():
    s = np.array([0.5, 2.0, 1.5])
    einv = np.array([[1.0, 0.2j, 0.1], [-0.2j, 2.0, 0.3], [0.1, 0.3, 0.5]])
    w = oracle.unpack_herm(oracle.screened_w(oracle.pack_upper(einv), s), 3)
    for i in range(3):
        for j in range(3):
            assert abs(w[i, j] - s[i] * einv[i, j] * s[j]) < 1e-15


def test_screened_w_dense_unit_eps_gives_coulomb():
    """eps^-1 = 1 (no screening) leaves W = v = (v^1/2)^2."""
    g = np.random.default_rng(4)
    v, _ = _hpd(g, 12)
    lam, q = np.linalg.eigh(v)
    vh = (q * np.sqrt(lam)) @ q.conj().T
    w = oracle.unpack_herm(oracle.screened_w(oracle.pack_upper(np.eye(12, dtype=complex)), vh), 12)
    assert np.max(np.abs(w - v)) < 1e-12 * np.max(np.abs(v))
