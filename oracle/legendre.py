"""Normalised Wigner-d rows for the spin-s spherical-harmonic transform (oracle, FP64).

    lambda_lm^(s)(theta) = sqrt((2l + 1) / (4 pi)) d^l_{m,-s}(theta),   l >= M = max(m, s)

so that the spin-s harmonic is ``sY_lm = (-1)^s lambda_lm^(s)(theta) e^{i m phi}``
(``gmaster/utils.py:683-699`` builds both transform directions from
``d^l_{m,-s}`` this way; for s = 0 it is the healpy ``Y_lm``).

Closed form and seeds follow ``docs/notes/latitudinal_march_maths.md`` sections 1-3
(``:28-128``): ``d^l_{m,-s} = (-1)^{m+s} 2^{eps LG2N} sin^{m+s}(theta/2)
cos^{|m-s|}(theta/2) P_n^{(a,b)}(cos theta)``, which at ``l = M`` is the seed

    lambda_M = (-1)^{m+s} sqrt((2M + 1)/(4 pi)) sqrt(C(2M, M + min(m, s)))
               sin^{m+s}(theta/2) cos^{|m-s|}(theta/2).

The recurrence in l is the standard three-term Wigner-d recurrence (Kostelec &
Rockmore 2008, eq. 4.3 with m' = -s), written for the normalised rows:

    lambda_{l+1} = (A_{l+1} x + C_{l+1}) lambda_l - B_{l+1} lambda_{l-1}
    K   = (l+1)(2l+1) / sqrt(((l+1)^2 - m^2)((l+1)^2 - s^2))
    A_{l+1} = sqrt((2l+3)/(2l+1)) K,      C_{l+1} = A_{l+1} m s / (l (l+1))   (0 if l = 0)
    B_{l+1} = sqrt((2l+3)/(2l-1)) K sqrt((l^2 - m^2)(l^2 - s^2)) / (l (2l+1))

(B vanishes at l = M, so lambda_{M+1} follows from lambda_M alone).  Values are
carried with an extended exponent (mantissa in [2^-256, 2^256] and an integer power of
two) so seeds far below the FP64 range are exact -- section 3 of the same note.

Pinned by: ``scipy.special.sph_harm_y`` for s = 0 (closed form of Y_lm), Wigner's
explicit sum (``wigner_d_explicit``, exact rational arithmetic in mpmath) for s = 1, 2,
3, the parity identities ``d^l_{m,0}(pi - theta) = (-1)^{l+m} d^l_{m,0}(theta)`` and
``d^l_{m,-s}(pi - theta) = (-1)^{l+m} d^l_{m,s}(theta)`` and Gauss-Legendre
orthonormality; see ``tests/test_oracle_legendre.py``.
"""

import math

import mpmath
import numpy as np

_FOUR_PI = 4.0 * math.pi


def seed_log2(m, s):
    """``(sign, log2 |lambda_M| - (m+s) log2 sin(t/2) - |m-s| log2 cos(t/2))`` to ~1e-18.

    The binomial factor is evaluated with mpmath at 40 digits, so the returned float is
    exact to FP64 rounding (needed because it is added to O(1e4) exponents).
    """
    M = max(m, s)
    k = min(m, s)
    with mpmath.workdps(40):
        val = 0.5 * mpmath.log((2 * M + 1) / (4 * mpmath.pi), 2) + 0.5 * mpmath.log(
            mpmath.binomial(2 * M, M + k), 2
        )
        return (-1) ** (m + s), float(val)


def recurrence_coefficients(L, m, s):
    """Arrays ``A, C, B`` of length L: entry ``l`` holds the coefficients producing
    ``lambda_l`` from ``lambda_{l-1}, lambda_{l-2}`` (valid for ``M < l < L``)."""
    A = np.zeros(L)
    C = np.zeros(L)
    B = np.zeros(L)
    M = max(m, s)
    for lp in range(M + 1, L):
        l = lp - 1
        K = (l + 1) * (2 * l + 1) / math.sqrt(((l + 1) ** 2 - m * m) * ((l + 1) ** 2 - s * s))
        A[lp] = math.sqrt((2 * l + 3) / (2 * l + 1)) * K
        C[lp] = A[lp] * m * s / (l * (l + 1)) if l > 0 else 0.0
        if l > M or (l == M and l > 0):
            B[lp] = (
                math.sqrt((2 * l + 3) / (2 * l - 1))
                * K
                * math.sqrt((l * l - m * m) * (l * l - s * s))
                / (l * (2 * l + 1))
            )
    return A, C, B


def scaled_coefficients(L, m, s):
    """The mu-rescaled form used by the HIP kernels.

    With ``P_M = P_{M+1} = 1`` and ``P_l = B_l P_{l-2}``, the rows ``mu_l = lambda_l / P_l``
    obey ``mu_l = (alpha_l x + gamma_l) mu_{l-1} - mu_{l-2}`` with
    ``alpha_l = A_l P_{l-1}/P_l`` and ``gamma_l = C_l P_{l-1}/P_l``.  Returns
    ``(alpha, gamma, P)`` of length L.
    """
    A, C, B = recurrence_coefficients(L, m, s)
    M = max(m, s)
    P = np.ones(L)
    alpha = np.zeros(L)
    gamma = np.zeros(L)
    for l in range(M + 2, L):
        P[l] = B[l] * P[l - 2]
    for l in range(M + 1, L):
        alpha[l] = A[l] * P[l - 1] / P[l]
        gamma[l] = C[l] * P[l - 1] / P[l]
    return alpha, gamma, P


def lambda_rows(L, m, s, z, shalf, chalf):
    """``lambda_lm^(s)(theta)`` for l = 0..L-1 at the colatitudes given by
    ``z = cos theta``, ``shalf = sin(theta/2)``, ``chalf = cos(theta/2)``.

    Returns an ``(L, nr)`` FP64 array (rows below ``max(m, s)`` are zero).
    """
    z = np.asarray(z, dtype=np.float64)
    shalf = np.asarray(shalf, dtype=np.float64)
    chalf = np.asarray(chalf, dtype=np.float64)
    nr = z.shape[0]
    out = np.zeros((L, nr))
    M = max(m, s)
    if M >= L:
        return out
    sign, c = seed_log2(m, s)
    alpha = m + s
    beta = abs(m - s)
    with np.errstate(divide="ignore"):
        f = np.full(nr, c)
        if alpha:
            f = f + alpha * np.log2(shalf)
        if beta:
            f = f + beta * np.log2(chalf)
    # Extended-exponent lanes: value = v * 2^e.
    finite = np.isfinite(f)
    e = np.where(finite, np.floor(f), 0).astype(np.int64)
    v = np.where(finite, sign * np.exp2(f - np.floor(np.where(finite, f, 0))), 0.0)
    A, C, B = recurrence_coefficients(L, m, s)
    v_prev = np.zeros(nr)
    out[M] = np.ldexp(v, e)
    for l in range(M + 1, L):
        v_new = (A[l] * z + C[l]) * v - B[l] * v_prev
        v_prev, v = v, v_new
        big = np.maximum(np.abs(v), np.abs(v_prev))
        up = big > 2.0**256
        dn = (big < 2.0**-256) & (big > 0)
        if up.any() or dn.any():
            shift = np.where(up, -256, np.where(dn, 256, 0))
            v = np.ldexp(v, shift)
            v_prev = np.ldexp(v_prev, shift)
            e = e - shift
        out[l] = np.ldexp(v, e)
    return out


def lambda_mp(l, m, s, theta, dps=50):
    """One value ``lambda_lm^(s)(theta)`` by the same recurrence in mpmath at ``dps`` digits
    (spot checks at large l, where Wigner's explicit sum cancels catastrophically)."""
    with mpmath.workdps(dps):
        t = mpmath.mpf(theta)
        x = mpmath.cos(t)
        M = max(m, s)
        if l < M:
            return 0.0
        lam = (
            (-1) ** (m + s)
            * mpmath.sqrt((2 * M + 1) / (4 * mpmath.pi))
            * mpmath.sqrt(mpmath.binomial(2 * M, M + min(m, s)))
            * mpmath.sin(t / 2) ** (m + s)
            * mpmath.cos(t / 2) ** abs(m - s)
        )
        prev = mpmath.mpf(0)
        for k in range(M, l):  # produce lambda_{k+1}
            K = (k + 1) * (2 * k + 1) / mpmath.sqrt(((k + 1) ** 2 - m * m) * ((k + 1) ** 2 - s * s))
            A = mpmath.sqrt(mpmath.mpf(2 * k + 3) / (2 * k + 1)) * K
            C = A * m * s / (k * (k + 1)) if k > 0 else 0
            if k > 0 and (k * k - m * m) * (k * k - s * s) > 0:
                B = (
                    mpmath.sqrt(mpmath.mpf(2 * k + 3) / (2 * k - 1))
                    * K
                    * mpmath.sqrt((k * k - m * m) * (k * k - s * s))
                    / (k * (2 * k + 1))
                )
            else:
                B = 0
            lam, prev = (A * x + C) * lam - B * prev, lam
        return float(lam)


def wigner_d_explicit(l, mp, m, theta, dps=40):
    """Wigner's explicit sum for ``d^l_{mp, m}(theta)`` (Sakurai 3.8.33), exact factorials.

    Independent of the recurrence; used to pin ``lambda_rows`` at small l.
    """
    with mpmath.workdps(dps):
        t = mpmath.mpf(theta)
        c = mpmath.cos(t / 2)
        sn = mpmath.sin(t / 2)
        total = mpmath.mpf(0)
        pref = mpmath.sqrt(
            mpmath.factorial(l + m) * mpmath.factorial(l - m) * mpmath.factorial(l + mp) * mpmath.factorial(l - mp)
        )
        for k in range(0, 2 * l + 1):
            a = l + m - k
            b = l - k - mp
            d = k - m + mp
            if a < 0 or b < 0 or d < 0:
                continue
            term = (
                (-1) ** (k - m + mp)
                * pref
                / (mpmath.factorial(a) * mpmath.factorial(k) * mpmath.factorial(b) * mpmath.factorial(d))
                * c ** (2 * l - 2 * k + m - mp)
                * sn ** (2 * k - m + mp)
            )
            total += term
        return float(total)


def mlim(lmax, s, sin_theta, z):
    """ducc0's polar-skip bound (docs/notes/latitudinal_march_maths.md:245-250):
    orders above ``s|z| + sqrt((lmax sin + ofs)^2 - s^2 sin^2)``, ``ofs = max(100, 0.01 lmax)``,
    are negligible on the ring.  The oracle itself never skips; the HIP path does."""
    ofs = max(100.0, 0.01 * lmax)
    b = lmax * sin_theta + ofs
    return s * np.abs(z) + np.sqrt(np.maximum(b * b - s * s * sin_theta * sin_theta, 0.0))
