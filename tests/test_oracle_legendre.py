"""Pins for the oracle's normalised Wigner-d rows (no GPU)."""

import math

import numpy as np
import pytest
from scipy.special import roots_legendre, sph_harm_y

from oracle import legendre

TH = np.array([0.05, 0.7, 1.3, np.pi / 2, 2.0, 3.05])


def _rows(L, m, s, th):
    return legendre.lambda_rows(L, m, s, np.cos(th), np.sin(th / 2), np.cos(th / 2))


def test_spin0_matches_scipy_sph_harm():
    L = 40
    for m in range(0, 35):
        lam = _rows(L, m, 0, TH)
        for ell in range(m, L):
            ref = sph_harm_y(ell, m, TH, 0.0).real
            assert np.max(np.abs(lam[ell] - ref)) < 5e-14 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("s", [1, 2, 3])
def test_spin_s_matches_wigner_explicit_sum(s):
    for m in range(0, 7):
        lam = _rows(11, m, s, TH)
        for ell in range(max(m, s), 11):
            for k, t in enumerate(TH):
                ref = math.sqrt((2 * ell + 1) / (4 * math.pi)) * legendre.wigner_d_explicit(ell, m, -s, t)
                assert abs(lam[ell, k] - ref) < 1e-13


def test_known_answers():
    t = np.array([0.4, 1.1, 2.5])
    y10 = _rows(2, 0, 0, t)[1]
    assert np.allclose(y10, np.sqrt(3 / (4 * np.pi)) * np.cos(t), atol=1e-15)
    y11 = _rows(2, 1, 0, t)[1]
    assert np.allclose(y11, -np.sqrt(3 / (8 * np.pi)) * np.sin(t), atol=1e-15)
    # 2Y_22 = (-1)^s lambda: sqrt(5/(64 pi)) (1 - cos t)^2
    y22 = _rows(3, 2, 2, t)[2]
    assert np.allclose(y22, np.sqrt(5 / (64 * np.pi)) * (1 - np.cos(t)) ** 2, atol=1e-15)


def test_parity_identities():
    L = 30
    th = TH[:3]
    for m in range(0, 12):
        a = _rows(L, m, 0, th)
        b = _rows(L, m, 0, np.pi - th)
        ell = np.arange(L)[:, None]
        assert np.max(np.abs(b - (-1.0) ** (ell + m) * a)) < 1e-13
    # mirror identity used by the spin-s kernels: d_{-m,-s}(t) = (-1)^{l+s} d_{m,-s}(pi - t)
    for s in (1, 2):
        for m in range(1, 6):
            for ell in range(max(m, s), 9):
                for t in th:
                    lhs = legendre.wigner_d_explicit(ell, -m, -s, t)
                    rhs = (-1) ** (ell + s) * legendre.wigner_d_explicit(ell, m, -s, np.pi - t)
                    assert abs(lhs - rhs) < 1e-13


def test_gauss_legendre_orthonormality():
    # int lambda_lm lambda_l'm dphi dcos = delta  ->  2 pi sum_k w_k lambda lambda' = delta
    L = 60
    x, w = roots_legendre(L + 2)
    for m, s in ((0, 0), (5, 0), (0, 2), (7, 2)):
        lam = legendre.lambda_rows(L, m, s, x, np.sqrt((1 - x) / 2), np.sqrt((1 + x) / 2))
        G = 2 * np.pi * (lam * w) @ lam.T
        M = max(m, s)
        assert np.max(np.abs(G[M:, M:] - np.eye(L - M))) < 5e-13


def test_large_l_and_underflowed_seed_match_mpmath():
    # l near 3000 and a seed far below the FP64 range (m = 1500 at theta = 0.3)
    cases = [(2999, 1500, 2, 0.3), (2999, 5, 0, 1.2), (1200, 1100, 2, 2.9)]
    for ell, m, s, t in cases:
        lam = _rows(ell + 1, m, s, np.array([t]))[ell, 0]
        ref = legendre.lambda_mp(ell, m, s, t)
        assert abs(lam - ref) <= 1e-11 * abs(ref) + 1e-300


def test_scaled_coefficients_reproduce_rows():
    # mu_l = (alpha x + gamma) mu_{l-1} - mu_{l-2}, lambda = P mu: the HIP kernels' form
    L = 50
    t = np.array([0.3, 1.9])
    for m, s in ((0, 0), (3, 0), (0, 2), (4, 2)):
        alpha, gamma, P = legendre.scaled_coefficients(L, m, s)
        lam = _rows(L, m, s, t)
        M = max(m, s)
        mu_prev, mu = np.zeros(2), lam[M].copy()
        for ell in range(M + 1, L):
            mu_prev, mu = mu, (alpha[ell] * np.cos(t) + gamma[ell]) * mu - mu_prev
            assert np.allclose(P[ell] * mu, lam[ell], rtol=1e-12, atol=1e-300)


def test_planted_sign_error_is_caught(monkeypatch):
    """A plausible mistake (the wrong sign of the m s / (l (l+1)) term) fails the explicit-sum pin."""
    good = legendre.recurrence_coefficients

    def bad(L, m, s):
        A, C, B = good(L, m, s)
        return A, -C, B

    monkeypatch.setattr(legendre, "recurrence_coefficients", bad)
    lam = _rows(8, 3, 2, np.array([0.9]))
    ref = math.sqrt(15 / (4 * math.pi)) * legendre.wigner_d_explicit(7, 3, -2, 0.9)
    assert abs(lam[7, 0] - ref) > 1e-6
