"""Pins of the pure-E/B coupling oracle (``oracle.pure``), CPU only."""

import math

import numpy as np
import pytest

from oracle import pure, sht
from oracle import purify as opur
from oracle.coupling import coupling_blocks_exact
from oracle.geometry import AlmInfo
from oracle.legendre import wigner_d_explicit
from oracle.master import coupled_cell

SPINS = ((0, 0), (0, 1), (0, 2), (1, 0), (1, 1), (1, 2), (2, 2), (0, -2), (-2, -2), (2, 0), (1, -2))


def _window(lmax, seed=1):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.1, 1, 2 * lmax + 1) * np.exp(-np.arange(2 * lmax + 1) / 4.0)


@pytest.mark.parametrize("m,n", SPINS)
def test_d_table_matches_wigner_explicit_sum(m, n):
    x = np.array([-0.93, -0.31, 0.0, 0.24, 0.77, 0.999])
    table = pure.wigner_d_table(x, m, n, 7)
    lmin = max(abs(m), abs(n))
    for l in range(8):
        want = [wigner_d_explicit(l, n, m, math.acos(v)) if l >= lmin else 0.0 for v in x]
        np.testing.assert_allclose(table[l], want, rtol=0, atol=2e-15)


def test_gauss_legendre_integrates_polynomials_exactly():
    order = 13
    x, w = pure.gauss_legendre(order)
    assert np.all(np.diff(x) > 0)
    for k in range(2 * order):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.sum(w * x**k) - exact) < 1e-14


def test_standard_blocks_equal_racah_3j_sums():
    lmax = 10
    W = _window(lmax)
    B = pure.pure_blocks(W, lmax)
    _, m02, m22p, m22m = coupling_blocks_exact(W, lmax)
    for got, want in ((B[0], m02), (B[1], m22p), (B[2], m22m)):
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-13 * np.abs(want).max())


def test_full_sky_window_makes_pure_blocks_standard():
    lmax = 12
    W = np.zeros(2 * lmax + 1)
    W[0] = 4 * np.pi  # mask = 1
    B = pure.pure_blocks(W, lmax)
    eye = np.diag((np.arange(lmax + 1) >= 2).astype(float))
    for i in (0, 1, 3, 4, 6):
        np.testing.assert_allclose(B[i], eye, atol=1e-13)
    for i in (2, 5, 7):
        np.testing.assert_allclose(B[i], 0.0, atol=1e-13)


@pytest.mark.parametrize("seed", [1, 2])
def test_purified_odd_blocks_vanish(seed):
    """A purified B (E) of a pure-E (B) sky is zero for any smooth window, so every M- block
    filling a slot with a purified field vanishes; the standard M- does not."""
    lmax = 14
    B = pure.pure_blocks(_window(lmax, seed), lmax)
    scale = np.abs(B[6]).max()
    assert np.abs(B[2]).max() > 1e-3 * np.abs(B[1]).max()
    assert np.abs(B[5]).max() < 1e-13 * scale
    assert np.abs(B[7]).max() < 1e-13 * scale


def _basis(ai, L):
    """Unit-power real modes at multipole L: a_L0 = 1 and a_Lm = 1/sqrt2, i/sqrt2."""
    out = []
    v = np.zeros(ai.nelem, complex)
    v[ai.idx(L, 0)] = 1
    out.append(v)
    for m in range(1, L + 1):
        for ph in (1, 1j):
            v = np.zeros(ai.nelem, complex)
            v[ai.idx(L, m)] = ph / np.sqrt(2)
            out.append(v)
    return out


def test_pure_blocks_give_expected_pseudo_spectra_of_purified_fields():
    """Expected pseudo-spectra summed over every unit-power mode at multipoles L of a
    band-limited sky (field lmax 6) seen through a band-limited mask (lmax 4), all
    transforms at lmax 12 on Nside 8 (exact for these band limits), match M[:, L] from the
    pure blocks for every purification of the two fields, and TE / TB with a purified
    spin-2 field."""
    nside, lf, lm, la, n_iter = 8, 6, 4, 12, 3
    ai = AlmInfo(la)
    rng = np.random.default_rng(5)
    am = np.zeros(ai.nelem, complex)
    for l in range(1, lm + 1):
        for m in range(l + 1):
            am[ai.idx(l, m)] = 0.5 / (1 + l) * (rng.normal() + (1j * rng.normal() if m else 0))
    am[ai.idx(0, 0)] = 2.5
    mask = sht.alm2map(am, 0, nside, la)[0]
    W = coupled_cell(am, am, la)[0][: 2 * lf + 1]
    B = pure.pure_blocks(W, lf)
    resp = {}
    for kind in ("E", "B", "EB"):
        for L in (2, 5):
            acc = {}
            for v in _basis(ai, L):
                alm = np.zeros((2, ai.nelem), complex)
                if "E" in kind:
                    alm[0] = v
                if "B" in kind:
                    alm[1] = v
                maps = sht.alm2map(alm, 2, nside, la)
                pu = opur.purify(mask, am, maps, nside, la, (True, True), n_iter)
                st = sht.map2alm(maps * mask[None], 2, nside, la, n_iter)
                tt = sht.map2alm(sht.alm2map(v, 0, nside, la) * mask[None], 0, nside, la, n_iter)[0]
                comp = {"sE": st[0], "sB": st[1], "pE": pu[0], "pB": pu[1], "T": tt}
                for a in comp:
                    for b in ("sE", "sB", "pE", "pB"):
                        acc[(a, b)] = acc.get((a, b), 0) + coupled_cell(comp[a], comp[b], la)[0][: lf + 1]
            resp[(kind, L)] = acc
    counter = np.abs(resp[("B", 5)][("pB", "pB")] - resp[("B", 5)][("sB", "sB")]).max()
    assert counter > 1e-2 * np.abs(resp[("B", 5)][("sB", "sB")]).max()  # the counter-terms matter here
    tol = 1e-5 * np.abs(B[1]).max()
    for pe1, pb1, pe2, pb2 in ((0, 0, 0, 0), (0, 1, 1, 1), (1, 0, 0, 1), (1, 1, 1, 1), (0, 1, 0, 1), (1, 0, 1, 0)):
        M = pure.assemble(B, pure.layout_pure(4, 2, 2, pe1, pb1, pe2, pb2), 4).reshape(lf + 1, 4, lf + 1, 4)
        n1 = ("pE" if pe1 else "sE", "pB" if pb1 else "sB")
        n2 = ("pE" if pe2 else "sE", "pB" if pb2 else "sB")
        for L in (2, 5):
            for kind, cols in (("E", (0,)), ("B", (3,)), ("EB", (0, 1, 2, 3))):
                for c in range(4):
                    got = resp[(kind, L)][(n1[c // 2], n2[c % 2])][2:]
                    want = sum(M[2:, c, L, k] for k in cols)
                    np.testing.assert_allclose(got, want, rtol=0, atol=tol)
    # T x (E, B) of a spin-2 field, T correlated with E (or B) mode by mode
    for pe2, pb2 in ((0, 0), (1, 1), (1, 0), (0, 1)):
        M = pure.assemble(B, pure.layout_pure(2, 0, 2, 0, 0, pe2, pb2), 2).reshape(lf + 1, 2, lf + 1, 2)
        for L in (2, 5):
            for c, (kind, name) in enumerate((("E", "pE" if pe2 else "sE"), ("B", "pB" if pb2 else "sB"))):
                np.testing.assert_allclose(resp[(kind, L)][("T", name)][2:], M[2:, c, L, c], rtol=0, atol=tol)


def test_layout_levels_follow_purity():
    lay = pure.layout_pure(4, 2, 2, 0, 1, 1, 1)
    assert [b for b, _, _ in lay] == [4, 4, 6, 6, 5, 5, 7, 7]
    lay = pure.layout_pure(7, 0, 2, 0, 0, 1, 0)
    assert [b for b, _, _ in lay][:3] == [None, 3, 0]
    assert [b for b, _, _ in lay][3:] == [6, 4, 4, 1, 7, 5, 5, 2]
