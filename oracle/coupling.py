"""Mode-coupling matrices for spin pairs 0-0, 0-2 and 2-2 (oracle, FP64).

The reference's general coupling matrix (``gmaster/workspaces.py:1069-1077``) is

    M[l1, l2] = (2 l2 + 1)/(4 pi) sum_l3 (2 l3 + 1) W_l3 P(J) (l1 l2 l3; n1 -s1 s1-n1)(l1 l2 l3; n2 -s2 s2-n2)

with ``W`` the mask pseudo-C_ell padded with zeros above lmax_mask
(``gmaster/workspaces.py:1778-1781``).  The blocks used by MASTER are
``M00`` (0,0,0,0), ``M02`` = sum over both parities of (0,2,0,2) (the TE block,
``gmaster/workspaces.py:912-918``) and ``M22+`` / ``M22-`` = even / odd parity of
(2,2,2,2) (``:919-922``).

The oracle evaluates them by Gauss-Legendre quadrature of Wigner-d products,
``int d^l1_{a a'} d^l2_{b b'} d^l3_{c c'} dx = 2 (3j)(3j)``, with the mask correlation
function ``xi(x) = sum (2 l3 + 1) W_l3 P_l3(x)``: the same identity the reference's
quadrature route uses (``gmaster/workspaces.py:682-732``), and an algorithm independent
of the l3 recursion in the HIP coupling kernel.  Pinned against exact Racah 3j sums
(``oracle.wigner3j.threej_exact``) at small lmax and against the full-sky invariant
(W = 4 pi delta_{l0} gives M = I, M22- = 0) in ``tests/test_oracle_coupling.py``.
"""

import math

import numpy as np
from scipy.special import roots_legendre

from .legendre import lambda_rows


def _d_rows(L, m, s, x):
    """``d^l_{m,-s}(x)`` rows (unnormalised) for l < L at nodes x."""
    lam = lambda_rows(L, m, s, x, np.sqrt((1 - x) / 2), np.sqrt((1 + x) / 2))
    ell = np.arange(L)
    return lam / np.sqrt((2 * ell + 1) / (4 * math.pi))[:, None]


def coupling_blocks(window, lmax):
    """``(M00, M02, M22p, M22m)`` each ``(lmax+1, lmax+1)`` from ``window`` (the mask
    pseudo-C_ell; entries beyond 2 lmax cannot couple and are ignored)."""
    L = lmax + 1
    window = np.asarray(window, dtype=np.float64)[: 2 * lmax + 1]
    Lw = len(window)
    n = (2 * lmax + Lw) // 2 + 2
    x, w = roots_legendre(n)
    P = _d_rows(L, 0, 0, x)  # d_00 = P_l
    Pw = _d_rows(Lw, 0, 0, x)
    d20 = _d_rows(L, 2, 0, x)  # d_{2,0} = d_{0,2}
    d0m2 = _d_rows(L, 0, 2, x)  # d_{0,-2}
    d2m2 = _d_rows(L, 2, 2, x)  # d_{2,-2}  (= d_{-2,2})
    # d_{2,2}(x) = d_{-2,-2}(x) = (-1)^l d_{2,-2}(-x)
    d22 = _d_rows(L, 2, 2, -x) * ((-1.0) ** np.arange(L))[:, None]
    # the node sums cancel by ~|xi| / |M| (up to ~1e3 for a white window), so they are
    # accumulated in long double
    ld = np.longdouble
    wx = w.astype(ld) * (((2 * np.arange(Lw) + 1) * window).astype(ld) @ Pw.astype(ld))
    col = ((2 * np.arange(L) + 1)[None, :] / (8 * np.pi)).astype(ld)

    def contract(a, b):
        return ((a.astype(ld) * wx) @ b.T.astype(ld) * col).astype(np.float64)

    M00 = contract(P, P)
    M02 = contract(d20, d0m2)
    Mall = contract(d22, d22)
    Msig = contract(d2m2, d2m2)
    M22p = 0.5 * (Mall + Msig)
    M22m = 0.5 * (Mall - Msig)
    return M00, M02, M22p, M22m


def coupling_blocks_exact(window, lmax):
    """The same four blocks from exact Racah 3j symbols (tiny lmax only)."""
    from .wigner3j import threej_exact

    L = lmax + 1
    W = np.asarray(window, dtype=np.float64)[: 2 * lmax + 1]
    out = np.zeros((4, L, L))
    for l1 in range(L):
        for l2 in range(L):
            acc = np.zeros(4)
            for l3 in range(abs(l1 - l2), min(l1 + l2, len(W) - 1) + 1):
                w0 = threej_exact(l1, l2, l3, 0, 0, 0)
                w2 = threej_exact(l1, l2, l3, 2, -2, 0)
                f = (2 * l3 + 1) * W[l3]
                acc[0] += f * w0 * w0
                acc[1] += f * w0 * w2
                if (l1 + l2 + l3) % 2 == 0:
                    acc[2] += f * w2 * w2
                else:
                    acc[3] += f * w2 * w2
            out[:, l1, l2] = acc * (2 * l2 + 1) / (4 * math.pi)
    return tuple(out)


def assemble_mcm(blocks, spin1, spin2):
    """Dense MCM in ``l * ncls + c`` ordering (``gmaster/workspaces.py:1851-1903``, ``:313-342``)."""
    M00, M02, M22p, M22m = blocks
    L = M00.shape[0]
    if spin1 == 0 and spin2 == 0:
        return M00.copy(), 1
    if (spin1 == 0) != (spin2 == 0):
        ncls = 2
        out = np.zeros((L, ncls, L, ncls))
        sign = (-1) ** (spin1 + spin2)
        for c in range(2):
            out[:, c, :, c] += sign * M02
        return out.reshape(L * ncls, L * ncls), ncls
    ncls = 4
    sign = (-1) ** (spin1 + spin2)
    out = np.zeros((L, ncls, L, ncls))
    for c in range(4):
        out[:, c, :, c] += sign * M22p
    for c in range(4):
        out[:, c, :, 3 - c] += (-sign if c in (1, 2) else sign) * M22m
    return out.reshape(L * ncls, L * ncls), ncls


def assemble_mcm_teb(blocks, spin2=2):
    """Dense joint T/E/B MCM (ncls 7: TT, TE, TB, EE, EB, BE, BB) in ``l * 7 + c`` ordering
    (``gmaster/workspaces.py:1871-1903`` with ``is_teb``): 00 at (0, 0), the 0-s block at
    (1, 1) and (2, 2) with sign (-1)^spin2, M+ on the spin x spin diagonal and M- on its
    anti-diagonal (signs +, -, -, +) at offset 3."""
    M00, M02, M22p, M22m = blocks
    L = M00.shape[0]
    out = np.zeros((L, 7, L, 7))
    out[:, 0, :, 0] = M00
    for c in (1, 2):
        out[:, c, :, c] = (-1) ** spin2 * M02
    for c in range(4):
        out[:, 3 + c, :, 3 + c] += M22p
        out[:, 3 + c, :, 6 - c] += (-1.0 if c in (1, 2) else 1.0) * M22m
    return out.reshape(L * 7, L * 7)


def general_coupling_exact(pcl, s1, s2, n1, n2, parity="all"):
    """``get_general_coupling_matrix`` (``gmaster/workspaces.py:1069-1122``) from exact Racah
    3j symbols (tiny sizes):

        M[l, l'] = sum_L (2l' + 1)(2L + 1)/(4 pi) C_L P(l + l' + L)
                   (l l' L; n1 -s1 s1-n1)(l l' L; n2 -s2 s2-n2),

    rows and columns below max(s1, s2, n1, n2) zeroed; ``parity`` "all", "even", "odd" or
    "both" (even and odd stacked)."""
    from .wigner3j import threej_exact

    C = np.asarray(pcl, dtype=np.float64)
    L = len(C)
    out = np.zeros((2, L, L))
    for l1 in range(L):
        for l2 in range(L):
            for l3 in range(abs(l1 - l2), min(l1 + l2, L - 1) + 1):
                a = threej_exact(l1, l2, l3, n1, -s1, s1 - n1)
                if a == 0:
                    continue
                b = threej_exact(l1, l2, l3, n2, -s2, s2 - n2)
                out[(l1 + l2 + l3) % 2, l1, l2] += (2 * l3 + 1) * C[l3] * a * b
    out *= (2 * np.arange(L) + 1)[None, None, :] / (4 * math.pi)
    lstart = max(s1, s2, n1, n2)
    out[:, :lstart, :] = 0
    out[:, :, :lstart] = 0
    if parity == "both":
        return out
    if parity == "even":
        return out[0]
    if parity == "odd":
        return out[1]
    return out.sum(axis=0)
