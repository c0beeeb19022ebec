"""Pure-E/B spin-2 coupling blocks (oracle, FP64 / long double, numpy) -- test infrastructure only.

Restates ``gmaster/workspaces.py:928-1052`` (with ``_wigner_d_table``, ``:609-679``): the
mode-coupling of purified spin-2 fields by Gauss-Legendre quadrature of order 2 lmax + 1
over Wigner d-functions.  Purification brings in the mask's spin-1 and spin-2 derivatives,
i.e. the mask correlation functions

    c_mn(x) = sum_L (2L+1) W_L / (4 pi) f_mn(L) d^L_mn(x),   s_mn(x) = the same with (-1)^L,

with f = 1, sqrt(L(L+1)) per spin-1 and sqrt((L+2)(L+1)L(L-1)) per spin-2 derivative, and
the row factors 2/sqrt((l+2)(l-1)) (spin-1 counter-term) and 1/sqrt((l+2)(l+1)l(l-1))
(spin-2 counter-term).  ``integrate(left, right)[l1, l2] = sum_k w_k left[l1](x_k)
right[l2](x_k) (2 l2 + 1) / 2``; the spin-2 blocks split by the parity of l1 + l2 into
even (M+) and odd (M-) parts.

Eight blocks are returned: standard TE, standard M+, standard M-, pure TE, M+ / M- with one
purified field, M+ / M- with two.

Convention: ``wigner_d_table(x, m, n, lmax)[l]`` is d^l_{mn}(beta) at cos(beta) = x from the
Jacobi-polynomial recurrence, with the reference's sign (-1)^((n - m + |m - n|) / 2); it equals
Wigner's explicit sum (``oracle.legendre.wigner_d_explicit``) with the two orders exchanged.

Pinned in ``tests/test_oracle_pure.py`` by: the d-tables against Wigner's explicit sum; the
three standard blocks against exact Racah 3j sums; the full-sky window (pure = standard);
the vanishing of every M- block with a purified field in the slot it fills (exact for a
smooth window on the sphere: the purified B of a pure-E sky is zero, and likewise E of a
pure-B sky); and the expected pseudo-spectra of purified fields (``oracle.purify``), summed
over every unit-power mode of a band-limited sky seen through a band-limited mask.
"""

import math

import numpy as np

LD = np.longdouble


def gauss_legendre(order):
    """Ascending Gauss-Legendre nodes and weights on [-1, 1], weights normalised to sum 2:
    numpy's rule polished by Newton steps on P_order in long double (numpy's own weights
    carry moment errors growing with the order, ~1e-14 at order 129)."""
    n = int(order)
    x = np.polynomial.legendre.leggauss(n)[0].astype(LD)

    def legendre_pair(z):
        p0, p1 = np.ones_like(z), z.copy()
        for k in range(2, n + 1):
            p0, p1 = p1, ((2 * k - 1) * z * p1 - (k - 1) * p0) / k
        return (p1, p0) if n > 1 else (z, np.ones_like(z))

    for _ in range(3):
        pn, pm = legendre_pair(x)
        x = x - pn / (n * (x * pn - pm) / (x * x - 1))
    pn, pm = legendre_pair(x)
    dp = n * (x * pn - pm) / (x * x - 1)
    w = 2 / ((1 - x * x) * dp * dp)
    return x.astype(np.float64), (w * (2 / w.sum())).astype(np.float64)


def wigner_d_table(x, m, n, lmax):
    """``(lmax + 1, len(x))`` table of d^l_{mn}(beta), cos(beta) = x; rows below max(|m|, |n|)
    are zero (``gmaster/workspaces.py:609-679``)."""
    x = np.asarray(x, dtype=np.float64)
    out = np.zeros((lmax + 1, len(x)))
    lmin = max(abs(m), abs(n))
    if lmin > lmax:
        return out
    mu, nu = abs(m - n), abs(m + n)
    t = mu + nu
    sh = np.sqrt(np.maximum((1 - x) / 2, 0.0))
    ch = np.sqrt(np.maximum((1 + x) / 2, 0.0))
    pre = (-1.0) ** ((n - m + mu) // 2) * sh**mu * ch**nu
    prev, cur = np.zeros_like(x), np.ones_like(x)
    norm = np.sqrt(math.factorial(t) / (math.factorial(mu) * math.factorial(nu)))
    for l in range(lmin, lmax + 1):
        k = l - lmin
        if k == 1:
            prev, cur = cur, 0.5 * ((mu - nu) + (t + 2) * x)
        elif k >= 2:
            nxt = ((2 * k + t - 1) * ((2 * k + t) * (2 * k + t - 2) * x + mu * mu - nu * nu) * cur
                   - 2 * (k + mu - 1) * (k + nu - 1) * (2 * k + t) * prev) / (2 * k * (k + t) * (2 * k + t - 2))
            prev, cur = cur, nxt
        if k >= 1:
            norm *= np.sqrt(k * (k + t) / ((k + mu) * (k + nu)))
        out[l] = pre * norm * cur
    return out


def _factors(lmax):
    ell = np.arange(2 * lmax + 1, dtype=np.float64)
    first = np.sqrt(ell * (ell + 1))
    second = np.sqrt(np.maximum((ell + 2) * (ell + 1) * ell * (ell - 1), 0.0))
    row = np.arange(lmax + 1, dtype=np.float64)
    row_first = np.where(row > 1, 2 / np.sqrt(np.maximum((row + 2) * (row - 1), 1)), 0.0)
    row_second = np.where(row > 1, 1 / np.sqrt(np.maximum((row + 2) * (row + 1) * row * (row - 1), 1)), 0.0)
    return ell, first, second, row_first, row_second


def pure_blocks(window, lmax):
    """The eight ``(lmax + 1, lmax + 1)`` blocks, in the order listed in the module docstring."""
    L = lmax + 1
    W = np.zeros(2 * lmax + 1)
    win = np.asarray(window, dtype=np.float64)[: 2 * lmax + 1]
    W[: len(win)] = win
    x, w = gauss_legendre(2 * lmax + 1)
    ell, mf, ms, rf, rs = _factors(lmax)
    coef = (2 * ell + 1) * W / (4 * np.pi)
    sign = np.where(ell % 2, -1.0, 1.0)

    def corr(m, n, f=1.0):
        T = wigner_d_table(x, m, n, 2 * lmax).astype(LD)
        cw = (coef * f).astype(LD)
        return (cw @ T).astype(np.float64), ((cw * sign) @ T).astype(np.float64)

    c00, s00 = corr(0, 0)
    c10, s10 = corr(1, 0, mf)
    c20, s20 = corr(2, 0, ms)
    c11, s11 = corr(1, 1, mf**2)
    c12, s12 = corr(1, 2, mf * ms)
    c22, s22 = corr(2, 2, ms**2)
    c01, _ = corr(0, 1, mf)
    c02, _ = corr(0, 2, ms)
    d = {mn: wigner_d_table(x, mn[0], mn[1], lmax) for mn in
         ((0, 0), (0, 1), (0, 2), (1, 0), (1, 1), (1, 2), (2, 2), (0, -2), (-2, -2))}
    col = ((2 * np.arange(L) + 1) / 2.0).astype(LD)

    def integrate(left, right):
        return ((left.astype(LD) * w.astype(LD)) @ right.T.astype(LD) * col[None]).astype(np.float64)

    rT, rS = d[(0, -2)], d[(-2, -2)]
    a, b = rf[:, None], rs[:, None]
    temp_left = d[(0, 2)] * c00 + d[(0, 1)] * c01 * a + d[(0, 0)] * c02 * b
    one = d[(2, 2)] * c00 + d[(1, 2)] * c10 * a + d[(0, 2)] * c20 * b
    one_s = d[(2, 2)] * s00 + d[(1, 2)] * s10 * a + d[(0, 2)] * s20 * b
    two = (d[(2, 2)] * c00 + 2 * d[(1, 2)] * c10 * a + 2 * d[(0, 2)] * c20 * b + d[(1, 1)] * c11 * a * a
           + 2 * d[(1, 0)] * c12 * a * b + d[(0, 0)] * c22 * b * b)
    two_s = (d[(2, 2)] * s00 + 2 * d[(1, 2)] * s10 * a + 2 * d[(0, 2)] * s20 * b + d[(1, 1)] * s11 * a * a
             + 2 * d[(1, 0)] * s12 * a * b + d[(0, 0)] * s22 * b * b)
    row = np.arange(L)
    pair = np.where((row[:, None] + row[None]) % 2, -1.0, 1.0)

    def parity(left, signed_left):
        total = integrate(left, rS)
        sgn = integrate(signed_left, rS) * pair
        return (total + sgn) / 2, (total - sgn) / 2

    std_te = integrate(d[(0, 2)] * c00, rT)
    std_even, std_odd = parity(d[(2, 2)] * c00, d[(2, 2)] * s00)
    pure_te = integrate(temp_left, rT)
    one_even, one_odd = parity(one, one_s)
    two_even, two_odd = parity(two, two_s)
    return std_te, std_even, std_odd, pure_te, one_even, one_odd, two_even, two_odd


def layout_pure(ncls, spin1, spin2, pe1, pb1, pe2, pb2):
    """``[(block index into pure_blocks, (row channel, column channel), sign)]`` for spin 0/2
    pairs with purification (``gmaster/workspaces.py:1845-1903``): the block of a slot is
    chosen by how many of the two fields are purified in that slot's channels."""
    te_levels = (0, 3)
    even_levels = (1, 4, 6)
    odd_levels = (2, 5, 7)
    out = []
    if ncls == 2:
        sign = float((-1) ** (spin1 + spin2))
        for c, level in enumerate((pe1 + pe2, pb1 + pb2)):
            out.append((te_levels[level], (c, c), sign))
        return out
    offset = 0
    if ncls == 7:
        offset = 3
        out.append((None, (0, 0), 1.0))
        for c, level in enumerate((int(pe2), int(pb2))):
            out.append((te_levels[level], (1 + c, 1 + c), float((-1) ** spin2)))
        pe1, pb1 = pe2, pb2
        spin_sign = 1.0
    else:
        spin_sign = float((-1) ** (spin1 + spin2))
    levels = (pe1 + pe2, pe1 + pb2, pb1 + pe2, pb1 + pb2)
    for c, level in enumerate(levels):
        out.append((even_levels[level], (offset + c, offset + c), spin_sign))
    for c, level in enumerate(levels):
        out.append((odd_levels[level], (offset + c, offset + 3 - c), -spin_sign if c in (1, 2) else spin_sign))
    return out


def assemble(blocks, layout, ncls, m00=None):
    """Dense MCM ``(L ncls, L ncls)`` in ``l * ncls + c`` ordering from a pure layout."""
    L = blocks[0].shape[0]
    out = np.zeros((L, ncls, L, ncls))
    for idx, (c1, c2), sign in layout:
        out[:, c1, :, c2] += sign * (m00 if idx is None else blocks[idx])
    return out.reshape(L * ncls, L * ncls)
