"""The Toeplitz approximation oracle (``oracle/toeplitz.py``) pinned by what the method fixes.

* A matrix whose correlation ``S / sqrt(D D)`` is exactly Toeplitz is reproduced by every
  fill (pure Toeplitz, exact columns, band, corner extrapolation), so the output equals the
  input; dropping or mis-indexing any fill breaks this.
* The fill reads only the exact entries listed by ``exact_pairs``
  (``gmaster/workspaces.py:520-539``; the reference's own check is
  ``tests/test_workspaces.py:108-119``).
* An independent per-element statement (the last fill that covers an element wins) agrees
  with the literal fill sequence on random symmetric matrices.
"""

import numpy as np
import pytest

from oracle.coupling import coupling_blocks
from oracle.toeplitz import apply_to_mcm_block, apply_toeplitz, exact_pairs, sanity

PARAMS = [(6, 3, 2), (16, 4, 3), (20, 20, 0), (30, 1, 5), (38, 10, 1), (1, 1, 0)]


def _toeplitz_correlated(n, seed):
    rng = np.random.default_rng(seed)
    c = np.concatenate([[1.0], rng.uniform(-0.5, 0.5, n - 1) / (1 + np.arange(1, n))])
    d = rng.uniform(0.5, 2.0, n)
    i = np.arange(n)
    return c[np.abs(i[:, None] - i[None])] * np.sqrt(d[:, None] * d[None])


def _random_symmetric(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(n, n))
    a = a + a.T
    a[np.diag_indices(n)] = np.abs(a[np.diag_indices(n)]) + 1.0
    return a


def _per_element(S, lt, le, dl):
    n = len(S)
    D = np.abs(np.diag(S))
    out = np.empty_like(S)

    def pure(i, j):
        k = (i - j + lt) % n
        den = np.sqrt(D[k] * D[lt])
        return (S[k, lt] / den if den > 0 else 0.0) * np.sqrt(D[i] * D[j])

    for a in range(n):
        for b in range(n):
            i, j = max(a, b), min(a, b)
            v = pure(i, j)
            if j <= le:
                v = S[i, j]
            if i - j <= dl:
                v = S[j, i]
            if i >= lt and j >= lt:
                v = pure(i, j)
            if i == j:
                v = S[i, i]
            if 1 <= i - j <= dl and j <= lt:
                v = S[j, i]
            if le <= j <= lt and i >= n - 1 - lt + j:
                k = le + i - j
                v = S[k, le] / np.sqrt(D[k] * D[le]) * np.sqrt(D[i] * D[j]) if D[k] > 0 else 0.0
            out[a, b] = v
    return out


@pytest.mark.parametrize("lt,le,dl", PARAMS)
def test_toeplitz_correlation_is_reproduced(lt, le, dl):
    S = _toeplitz_correlated(40, lt + 7 * le + dl)
    np.testing.assert_allclose(apply_toeplitz(S, lt, le, dl), S, rtol=0, atol=1e-14)


@pytest.mark.parametrize("lt,le,dl", PARAMS)
def test_output_is_symmetric_and_keeps_exact_entries(lt, le, dl):
    S = _random_symmetric(40, 3)
    out = apply_toeplitz(S, lt, le, dl)
    np.testing.assert_array_equal(out, out.T)
    np.testing.assert_array_equal(np.diag(out), np.diag(S))
    np.testing.assert_array_equal(out[lt:, :le], S[lt:, :le])
    # column l_exact is rebuilt by the corner extrapolation from itself: equal up to rounding
    np.testing.assert_allclose(out[lt:, le], S[lt:, le], rtol=1e-15)
    for off in range(1, dl + 1):
        k = min(40 - off, lt + 1)
        j = np.arange(k)
        # the corner extrapolation (columns le..lt, rows >= n - 1 - lt + column) is applied last
        keep = (j < le) | (off < 40 - 1 - lt)
        np.testing.assert_array_equal(np.diag(out, off)[:k][keep], np.diag(S, off)[:k][keep])


@pytest.mark.parametrize("lt,le,dl", PARAMS)
def test_fill_reads_only_the_exact_pairs(lt, le, dl):
    S = _random_symmetric(40, 11)
    r, c = exact_pairs(40, lt, le, dl)
    sparse = np.zeros_like(S)
    sparse[r, c] = S[r, c]
    sparse[c, r] = S[r, c]
    np.testing.assert_array_equal(apply_toeplitz(sparse, lt, le, dl), apply_toeplitz(S, lt, le, dl))


@pytest.mark.parametrize("lt,le,dl", PARAMS)
def test_per_element_rule_matches_fill_sequence(lt, le, dl):
    S = _random_symmetric(40, 5)
    np.testing.assert_array_equal(_per_element(S, lt, le, dl), apply_toeplitz(S, lt, le, dl))


def test_coupling_block_column_factor():
    lmax = 30
    w = np.random.default_rng(19).uniform(size=2 * lmax + 1)
    for M in coupling_blocks(w, lmax)[:2]:
        col = 2 * np.arange(lmax + 1) + 1
        got = apply_to_mcm_block(M, 16, 4, 3)
        np.testing.assert_allclose(got / col, apply_toeplitz(M / col, 16, 4, 3), rtol=1e-15, atol=0)
        # exact columns and diagonal of M itself survive
        np.testing.assert_allclose(np.diag(got), np.diag(M), rtol=1e-15)
        np.testing.assert_allclose(got[:, :5], M[:, :5], rtol=1e-15)


def test_disabled_is_identity():
    S = _random_symmetric(10, 1)
    np.testing.assert_array_equal(apply_toeplitz(S, -1, -1, -1), S)


@pytest.mark.parametrize(
    "args,msg",
    [
        ((6, 0, 2, 12), "must be positive"),
        ((6, 3, -1, 12), "must be positive"),
        ((6, 7, 2, 12), "must be `<= l_toeplitz"),
        ((12, 3, 2, 12), "smaller than `l_max`"),
        ((6, 3, 12, 12), "smaller than `l_max`"),
    ],
)
def test_sanity_messages(args, msg):
    with pytest.raises(ValueError, match=msg):
        sanity(*args)
    sanity(-1, -1, -1, 12)


@pytest.mark.parametrize("lt,le,dl", PARAMS)
def test_entry_kind_marks_exactly_the_kept_entries(lt, le, dl):
    from oracle.toeplitz import entry_kind

    S = _random_symmetric(40, 21)
    out = apply_toeplitz(S, lt, le, dl)
    kind = entry_kind(40, lt, le, dl)
    assert np.array_equal(kind, kind.T)
    # kind 0 entries are the input's values (column l_exact rebuilt by the corner fill aside)
    np.testing.assert_array_equal(out[kind == 0], S[kind == 0])
    # the per-element rule agrees on which fill is last
    for a in range(40):
        for b in range(40):
            i, j = max(a, b), min(a, b)
            k = 1
            if j <= le or i - j <= dl:
                k = 0
            if i >= lt and j >= lt:
                k = 1
            if i == j or (1 <= i - j <= dl and j <= lt):
                k = 0
            if le <= j <= lt and i >= 40 - 1 - lt + j:
                k = 2
            assert kind[a, b] == k


def test_correction_with_zero_residual_is_identity_and_is_bilinear():
    from oracle.toeplitz import correct, entry_kind

    S = _random_symmetric(40, 4) + 60 * np.eye(40)
    col = 2 * np.arange(40) + 1
    M = S * col[None]
    lg, dg = np.array([3, 9, 30]), np.array([1, 4, 16])
    np.testing.assert_array_equal(correct(M, 20, 2, 1, lg, dg, np.zeros(9)), M)
    # a constant e adds e sqrt(D_i D_j) (2 col + 1) on every covered non-exact element
    out = correct(M, 20, 2, 1, lg, dg, np.full(9, 0.5))
    kind = entry_kind(40, 20, 2, 1)
    D = np.abs(np.diag(S))
    i, j = np.meshgrid(np.arange(40), np.arange(40), indexing="ij")
    lo, d = np.minimum(i, j), np.abs(i - j)
    cov = (kind != 0) & (lo >= 3) & (lo <= 30) & (d >= 1) & (d <= 16)
    want = M + np.where(cov, 0.5 * np.sqrt(D[i] * D[j]) * col[None], 0.0)
    np.testing.assert_allclose(out, want, rtol=1e-15)
    # bilinear in logs: on a grid node the node value is used
    e = np.arange(9, dtype=float)
    out = correct(M, 20, 2, 1, lg, dg, e)
    a, b = 9 + 4, 9  # j = 9 (node 1), d = 4 (node 1) -> e[1, 1] = 4
    assert kind[a, b] != 0
    assert out[a, b] - M[a, b] == pytest.approx(4.0 * np.sqrt(D[a] * D[b]) * col[b], rel=1e-14)
