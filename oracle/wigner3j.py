"""Wigner 3j symbols (oracle).

``threej_exact`` evaluates Racah's formula with Python integers and returns the value
rounded once to FP64 -- an exact reference, independent of every recursion.  It pins
``closed_form_000`` (the ``g(p) = Gamma(p+1/2)/(sqrt(pi) Gamma(p+1))`` closed form the
reference uses for the scalar coupling, ``gmaster/workspaces.py:440-517``,
``tests/test_workspaces.py:122-145``) and the recursion-based HIP coupling kernel.
"""

import math
from fractions import Fraction

import mpmath
import numpy as np


def _f(n):
    return math.factorial(n)


def threej_exact(j1, j2, j3, m1, m2, m3):
    """(j1 j2 j3; m1 m2 m3) by Racah's formula, exact rational arithmetic."""
    if m1 + m2 + m3 != 0:
        return 0.0
    if j3 < abs(j1 - j2) or j3 > j1 + j2:
        return 0.0
    if abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.0
    tri = Fraction(_f(j1 + j2 - j3) * _f(j1 - j2 + j3) * _f(-j1 + j2 + j3), _f(j1 + j2 + j3 + 1))
    pre = tri * _f(j1 + m1) * _f(j1 - m1) * _f(j2 + m2) * _f(j2 - m2) * _f(j3 + m3) * _f(j3 - m3)
    kmin = max(0, j2 - j3 - m1, j1 - j3 + m2)
    kmax = min(j1 + j2 - j3, j1 - m1, j2 + m2)
    total = Fraction(0)
    for k in range(kmin, kmax + 1):
        den = (
            _f(k)
            * _f(j1 + j2 - j3 - k)
            * _f(j1 - m1 - k)
            * _f(j2 + m2 - k)
            * _f(j3 - j2 + m1 + k)
            * _f(j3 - j1 - m2 + k)
        )
        total += Fraction((-1) ** k, den)
    sign = (-1) ** (j1 - j2 - m3)
    # value = sign * sqrt(pre) * total
    with mpmath.workdps(40):
        val = sign * mpmath.sqrt(mpmath.mpf(pre.numerator) / pre.denominator) * (
            mpmath.mpf(total.numerator) / total.denominator
        )
        return float(val)


def g_table(n):
    """``g(p) = Gamma(p + 1/2) / (sqrt(pi) Gamma(p + 1))`` for p = 0..n-1 (``g(p) = g(p-1)(2p-1)/(2p)``)."""
    out = np.empty(n, dtype=np.float64)
    acc = mpmath.mpf(1)
    with mpmath.workdps(30):
        for p in range(n):
            if p:
                acc = acc * (2 * p - 1) / (2 * p)
            out[p] = float(acc)
    return out


def closed_form_000_sq(l1, l2, l3, g):
    """``(l1 l2 l3; 0 0 0)^2 = g(J/2-l1) g(J/2-l2) g(J/2-l3) / ((J+1) g(J/2))`` for even J, else 0."""
    J = l1 + l2 + l3
    if J % 2 or l3 < abs(l1 - l2) or l3 > l1 + l2:
        return 0.0
    h = J // 2
    return g[h - l1] * g[h - l2] * g[h - l3] / ((J + 1) * g[h])
