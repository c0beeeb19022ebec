"""The spin-2 transforms written over the spin-0 rows lambda_lm (the identity the batched
latitudinal kernels use so that T, the mask and Q/U share one recurrence), checked in
numpy against the oracle's spin-2 transforms built from d^l_{m,-2} directly.

With W = (lambda+_lm + (-1)^m lambda-_lm)/2 and X = (lambda+_lm - (-1)^m lambda-_lm)/2 (the
spin-2 rows of orders +m and -m), the HEALPix functions of Gorski et al. (2005) give

    W_l = aW_l lambda_l / s^2 + bW_l lambda_l + cW_l lambda_{l-2} / s^2
    X_l = dX_l lambda_{l+1} / s^2 + eX_l lambda_{l-1} / s^2,       s^2 = sin^2(theta),

after the recurrence x lambda_{l-1} = eps_l lambda_l + eps_{l-1} lambda_{l-2} removes the
cos(theta) factors.  So the analysis needs only S0_j = sum_r lambda_j G_r and
S1_j = sum_r lambda_j G_r / s_r^2 of the Q and U ring spectra, and the synthesis only
sums of lambda_j times per-degree coefficients, split into a 1/s^2 part and a plain part;
the plain part folds into the 1/s^2 part (one coefficient per degree, up to degree lmax + 2).
"""

import numpy as np

from oracle import sht
from oracle.geometry import AlmInfo, Rings
from oracle.legendre import lambda_rows


def _coefs(L, m):
    """(aW, bW, cW, dX, eX) for l = 0..L-1 (zero for l < max(m, 2))."""
    out = np.zeros((5, L))
    for l in range(max(m, 2), L):
        c = 2.0 / np.sqrt((l + 2.0) * (l + 1.0) * l * (l - 1.0))
        k = np.sqrt((2 * l + 1.0) * (l * l - m * m) / (2 * l - 1.0))

        def eps(j):
            return np.sqrt(max(j * j - m * m, 0) / (4.0 * j * j - 1.0)) if j >= m and j > 0 else 0.0

        out[0, l] = c * (m * m - l + (l * l - m * m) / (2 * l - 1.0))
        out[1, l] = -c * l * (l - 1) / 2.0
        out[2, l] = c * k * eps(l - 1)
        out[3, l] = -c * m * (l - 1) * eps(l + 1)
        out[4, l] = c * m * (k - (l - 1) * eps(l))
    return out


def _rows(nside, lmax):
    R = Rings(nside)
    return R, 1.0 - R.z * R.z


def test_spin2_rows_from_spin0_rows():
    """W and X of the oracle's d^l_{m,-2} rows equal the spin-0 combinations, every order of
    Nside 16 (lmax 47), every ring including the polar ones."""
    nside, lmax = 16, 47
    L = lmax + 1
    R, s2 = _rows(nside, lmax)
    mir = R.nring - 1 - np.arange(R.nring)
    ell = np.arange(L)[:, None]
    for m in range(L):
        lp = lambda_rows(L, m, 2, R.z, R.shalf, R.chalf)
        lm = ((-1.0) ** ell) * lp[:, mir]
        W = (lp + (-1) ** m * lm) / 2
        X = (lp - (-1) ** m * lm) / 2
        lam = lambda_rows(L + 1, m, 0, R.z, R.shalf, R.chalf)
        aW, bW, cW, dX, eX = _coefs(L, m)
        for l in range(max(m, 2), L):
            w = (aW[l] / s2 + bW[l]) * lam[l] + cW[l] / s2 * lam[l - 2]
            x = (dX[l] * lam[l + 1] + eX[l] * lam[l - 1]) / s2
            scale = max(np.max(np.abs(W[l])), np.max(np.abs(X[l])), 1e-300)
            assert np.max(np.abs(w - W[l])) < 1e-12 * scale, (m, l)
            assert np.max(np.abs(x - X[l])) < 1e-12 * scale, (m, l)


def test_spin2_analysis_and_synthesis_in_the_spin0_basis():
    """E/B from the S0/S1 sums of the Q and U spectra equal the oracle's spin-2 analysis, and
    Q/U spectra from the (1/s^2, plain) coefficient sums equal its synthesis (Nside 8)."""
    nside, lmax = 8, 23
    L = lmax + 1
    R, s2 = _rows(nside, lmax)
    ai = AlmInfo(lmax)
    rng = np.random.default_rng(3)
    qu = rng.normal(size=(2, R.npix))
    orders = np.arange(-(L - 1), L)
    Gp = sht.phased_spectrum(qu[0] + 1j * qu[1], R, orders)  # F+ of Q + iU
    ref = sht.plus_to_eb(sht.analysis_plus(Gp, R, ai, 2), ai, 2)
    E = np.zeros(ai.nelem, dtype=np.complex128)
    B = np.zeros(ai.nelem, dtype=np.complex128)
    for m in range(L):
        fp, fm = Gp[:, L - 1 + m], Gp[:, L - 1 - m]
        GQ = (fp + np.conj(fm)) / 2
        GU = (fp - np.conj(fm)) / 2j
        lam = lambda_rows(L + 1, m, 0, R.z, R.shalf, R.chalf)
        S0Q, S0U = lam @ GQ, lam @ GU
        S1Q, S1U = lam @ (GQ / s2), lam @ (GU / s2)
        aW, bW, cW, dX, eX = _coefs(L, m)
        for l in range(max(m, 2), L):
            sw_q = aW[l] * S1Q[l] + bW[l] * S0Q[l] + cW[l] * S1Q[l - 2]
            sw_u = aW[l] * S1U[l] + bW[l] * S0U[l] + cW[l] * S1U[l - 2]
            sx_q = dX[l] * S1Q[l + 1] + eX[l] * S1Q[l - 1]
            sx_u = dX[l] * S1U[l + 1] + eX[l] * S1U[l - 1]
            E[ai.mstart[m] + l] = -(sw_q + 1j * sx_u)
            B[ai.mstart[m] + l] = 1j * sx_q - sw_u
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(E - ref[0])) < 1e-12 * scale
    assert np.max(np.abs(B - ref[1])) < 1e-12 * scale

    # synthesis: random E, B -> F+ at +-m via the two per-degree coefficient sets
    eb = rng.normal(size=(2, ai.nelem)) + 1j * rng.normal(size=(2, ai.nelem))
    eb[:, ai.ell < 2] = 0
    eb[:, :L] = eb[:, :L].real
    Fref = sht.synthesis_plus(sht.eb_to_plus(eb, ai), R, ai, 2)
    for m in range(L):
        Em = np.zeros(L + 3, dtype=np.complex128)
        Bm = np.zeros(L + 3, dtype=np.complex128)
        Em[m:L] = eb[0, ai.mstart[m] + np.arange(m, L)]
        Bm[m:L] = eb[1, ai.mstart[m] + np.arange(m, L)]
        aW, bW, cW, dX, eX = (np.concatenate([v, np.zeros(3)]) for v in _coefs(L, m))
        j = np.arange(m, L + 1)

        def at(v, jj):
            return np.where(jj >= 0, v[np.clip(jj, 0, L + 2)], 0)

        aQ = -(aW[j] * Em[j] + at(cW, j + 2) * at(Em, j + 2)
               + 1j * (at(dX, j - 1) * at(Bm, j - 1) + at(eX, j + 1) * at(Bm, j + 1)))
        bQ = -bW[j] * Em[j]
        aU = -(aW[j] * Bm[j] + at(cW, j + 2) * at(Bm, j + 2)
               - 1j * (at(dX, j - 1) * at(Em, j - 1) + at(eX, j + 1) * at(Em, j + 1)))
        bU = -bW[j] * Bm[j]
        lam = lambda_rows(L + 3, m, 0, R.z, R.shalf, R.chalf)[m:]  # degrees m .. L + 2
        Q = (aQ @ lam[:len(j)]) / s2 + bQ @ lam[:len(j)]
        U = (aU @ lam[:len(j)]) / s2 + bU @ lam[:len(j)]
        sc = max(np.max(np.abs(Fref)), 1e-300)
        assert np.max(np.abs(Q + 1j * U - Fref[:, L - 1 + m])) < 1e-12 * sc, m
        if m:
            assert np.max(np.abs(np.conj(Q) + 1j * np.conj(U) - Fref[:, L - 1 - m])) < 1e-12 * sc, m
        # the plain part folded into the 1/s^2 part: s^2 lambda_j = (1 - eps_{j+1}^2 - eps_j^2) lambda_j
        # - eps_{j+1} eps_{j+2} lambda_{j+2} - eps_j eps_{j-1} lambda_{j-2}, one coefficient per degree
        ee = np.array([np.sqrt(max(k * k - m * m, 0) / (4.0 * k * k - 1.0)) if k >= m and k > 0 else 0.0
                       for k in range(L + 6)])
        jj = np.arange(m, L + 2)  # the folded plain part reaches degree lmax + 2 = L + 1
        bq_full = np.zeros(L + 6, dtype=np.complex128)
        bu_full = np.zeros(L + 6, dtype=np.complex128)
        bq_full[j] = bQ
        bu_full[j] = bU

        def fold(bf):
            lower = np.where(jj >= 2, bf[np.maximum(jj - 2, 0)], 0)
            return bf[jj] * (1 - ee[jj + 1] ** 2 - ee[jj] ** 2) - lower * ee[jj - 1] * ee[jj] \
                - bf[jj + 2] * ee[jj + 1] * ee[jj + 2]

        aq_full = np.zeros(L + 6, dtype=np.complex128)
        au_full = np.zeros(L + 6, dtype=np.complex128)
        aq_full[j] = aQ
        au_full[j] = aU
        gq = aq_full[jj] + fold(bq_full)
        gu = au_full[jj] + fold(bu_full)
        Qf = (gq @ lam[:len(jj)]) / s2
        Uf = (gu @ lam[:len(jj)]) / s2
        assert np.max(np.abs(Qf + 1j * Uf - Fref[:, L - 1 + m])) < 1e-12 * sc, m
