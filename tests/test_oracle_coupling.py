"""Pins for the oracle's 3j symbols, coupling matrices and MASTER algebra (no GPU)."""

import math
import os

import numpy as np
import pytest

from oracle import coupling, master, wigner3j

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_racah_matches_sympy():
    from sympy.physics.wigner import wigner_3j

    for args in [(3, 4, 5, 2, -2, 0), (7, 3, 6, 0, 0, 0), (10, 12, 5, 2, -2, 0), (2, 2, 2, 2, -2, 0),
                 (6, 6, 1, 2, -2, 0), (9, 4, 7, 1, -3, 2)]:
        assert abs(wigner3j.threej_exact(*args) - float(wigner_3j(*args))) < 1e-15


def test_closed_form_000_matches_racah():
    g = wigner3j.g_table(80)
    for l1, l2, l3 in [(7, 3, 6), (10, 10, 0), (20, 13, 9), (30, 31, 1)]:
        assert abs(wigner3j.closed_form_000_sq(l1, l2, l3, g) - wigner3j.threej_exact(l1, l2, l3, 0, 0, 0) ** 2) < 1e-15


def _shrinking_offset_tt(window, lmax):
    """Per-offset closed form of the scalar MCM, restated from the reference's own test
    (tests/test_workspaces.py:120-145)."""
    n = lmax + 1
    p = np.arange(1, 2 * lmax + 1, dtype=np.float64)
    g = np.exp(np.concatenate([[0.0], np.cumsum(np.log((p - 0.5) / p))]))
    mp = window * (2 * np.arange(2 * lmax + 1) + 1) / (4 * np.pi)
    row = np.arange(n)[:, None]
    col = np.arange(n)[None, :]
    lo, up = np.minimum(row, col), np.maximum(row, col)
    M = np.zeros((n, n))
    for o in range(n):
        pt = up + o
        term = mp[np.minimum(up - lo + 2 * o, 2 * lmax)] * g[up - lo + o] * g[o] * g[np.maximum(lo - o, 0)] / (g[pt] * (2 * pt + 1))
        M += np.where(o <= lo, term, 0)
    return M * (2 * col + 1)


def test_reference_offset_closed_form_is_exact():
    lmax = 9
    w = np.random.default_rng(3).uniform(size=2 * lmax + 1)
    ex = np.zeros((lmax + 1, lmax + 1))
    for l1 in range(lmax + 1):
        for l2 in range(lmax + 1):
            ex[l1, l2] = sum((2 * l3 + 1) * w[l3] * wigner3j.threej_exact(l1, l2, l3, 0, 0, 0) ** 2
                             for l3 in range(abs(l1 - l2), l1 + l2 + 1)) * (2 * l2 + 1) / (4 * math.pi)
    assert np.max(np.abs(_shrinking_offset_tt(w, lmax) - ex)) < 1e-14


@pytest.mark.parametrize("lmax", [15, 16, 31])
def test_quadrature_m00_matches_reference_offset_closed_form(lmax):
    """The quadrature oracle's precision is ~1e-12 of max |M| for a white window: Wigner-d
    rounding is amplified by |xi| / |M| ~ 1e3 (the node sums themselves are long double)."""
    w = np.random.default_rng(20 + lmax).uniform(size=2 * lmax + 1)
    M00 = coupling.coupling_blocks(w, lmax)[0]
    assert np.max(np.abs(M00 - _shrinking_offset_tt(w, lmax))) < 2e-12 * max(1, np.max(np.abs(M00)))


def test_quadrature_blocks_match_exact_fixture():
    fx = np.load(os.path.join(GOLDEN, "coupling_exact_lmax12.npz"))
    q = coupling.coupling_blocks(fx["window"], 12)
    for got, key in zip(q, ("m00", "m02", "m22p", "m22m")):
        assert np.max(np.abs(got - fx[key])) < 1e-12 * np.max(np.abs(fx[key]))


def test_full_sky_identity():
    lmax = 20
    w = np.zeros(lmax + 1)
    w[0] = 4 * np.pi
    M00, M02, M22p, M22m = coupling.coupling_blocks(w, lmax)
    eye2 = np.eye(lmax + 1)
    eye2[:2, :2] = 0
    assert np.max(np.abs(M00 - np.eye(lmax + 1))) < 5e-13
    assert np.max(np.abs(M02 - eye2)) < 5e-13 and np.max(np.abs(M22p - eye2)) < 5e-13
    assert np.max(np.abs(M22m)) < 5e-13


def test_symmetry_of_master_coefficients():
    lmax = 25
    w = np.random.default_rng(2).uniform(size=lmax + 1)
    col = 2 * np.arange(lmax + 1) + 1
    for M in coupling.coupling_blocks(w, lmax):
        S = M / col[None, :]
        assert np.allclose(S, S.T, atol=1e-14)


def test_linear_bins_drop_incomplete_band():
    ells, bpws, w = master.linear_bins(12287, 16)
    assert bpws.max() + 1 == 767 and ells[0] == 2 and ells[-1] == 2 + 767 * 16 - 1
    assert np.allclose(np.bincount(bpws, weights=w), 1.0)


def test_decoupling_recovers_input_under_coupling():
    lmax, nlb = 47, 8
    w = 1.0 / (1.0 + np.arange(lmax + 1)) ** 2
    mcm, mb, ncls, (ells, bpws, wt) = master.workspace(w, lmax, 0, 0, nlb)
    # a band-constant spectrum coupled through M decouples exactly to its band values
    nb = bpws.max() + 1
    vals = np.linspace(1, 2, nb)
    cl = np.zeros(lmax + 1)
    cl[ells] = vals[bpws]
    got = master.decouple(mb, (mcm @ cl)[None], lmax, ells, bpws, wt, ncls)
    assert np.allclose(got[0], vals, rtol=1e-12)


def test_pipeline_fixture_is_reproducible():
    fx = np.load(os.path.join(GOLDEN, "pipeline_nside16.npz"))
    res = master.pipeline_tp(fx["mask"], fx["tqu"], 16, 47, 8, 3)
    for key in ("bp_00", "bp_02", "bp_22", "pcl_mask"):
        assert np.max(np.abs(res[key] - fx[key])) <= 1e-12 * np.max(np.abs(fx[key]))


def test_teb_assembly_is_the_direct_sum_of_the_separate_layouts():
    """The joint layout holds the 00, 0-2 and 2-2 matrices of the separate workspaces as
    diagonal blocks (channels 0, 1-2, 3-6) and nothing else."""
    from oracle.coupling import assemble_mcm, assemble_mcm_teb

    rng = np.random.default_rng(7)
    L = 9
    blocks = [rng.normal(size=(L, L)) for _ in range(4)]
    teb = assemble_mcm_teb(blocks).reshape(L, 7, L, 7)
    m00, _ = assemble_mcm(blocks, 0, 0)
    m02, _ = assemble_mcm(blocks, 0, 2)
    m22, _ = assemble_mcm(blocks, 2, 2)
    np.testing.assert_array_equal(teb[:, 0, :, 0], m00)
    np.testing.assert_array_equal(teb[:, 1:3, :, 1:3].reshape(2 * L, 2 * L), m02)
    np.testing.assert_array_equal(teb[:, 3:, :, 3:].reshape(4 * L, 4 * L), m22)
    off = teb.copy()
    off[:, 0, :, 0] = 0
    off[:, 1:3, :, 1:3] = 0
    off[:, 3:, :, 3:] = 0
    assert not off.any()


def test_general_coupling_oracle_reduces_to_the_spin_blocks():
    """The exact general-coupling restatement equals the 00 / 02 / 22+- blocks for those spins
    (every parity option), and zeroes rows and columns below max(s1, s2, n1, n2)."""
    from oracle.coupling import general_coupling_exact

    rng = np.random.default_rng(12)
    C = rng.uniform(size=8)
    M00, M02, M22p, M22m = coupling.coupling_blocks_exact(np.concatenate([C, np.zeros(7)]), 7)
    np.testing.assert_allclose(general_coupling_exact(C, 0, 0, 0, 0), M00, atol=1e-15)
    np.testing.assert_allclose(general_coupling_exact(C, 0, 2, 0, 2), M02, atol=1e-15)
    both = general_coupling_exact(C, 2, 2, 2, 2, "both")
    np.testing.assert_allclose(both[0], M22p, atol=1e-15)
    np.testing.assert_allclose(both[1], M22m, atol=1e-15)
    np.testing.assert_allclose(general_coupling_exact(C, 2, 2, 2, 2), M22p + M22m, atol=1e-15)
    g = general_coupling_exact(C, 3, 2, -1, 1)
    assert np.all(g[:3] == 0) and np.all(g[:, :3] == 0) and np.abs(g).max() > 1e-3
