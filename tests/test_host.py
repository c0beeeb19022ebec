"""Host-side logic of gmaster_amd that runs without a GPU."""

import math

import numpy as np
import pytest

from gmaster_amd import bins, sht
from oracle import legendre, master


@pytest.mark.parametrize("s", [0, 1, 2, 3])
def test_seed_constants_match_oracle(s):
    mant, exp = sht._seed_constants(2000, s)
    for m in list(range(0, 12)) + [100, 999, 1999]:
        sign, c = legendre.seed_log2(m, s)
        assert np.sign(mant[m]) == sign
        assert abs(math.log2(abs(mant[m])) + exp[m] - c) < 1e-13 * max(1.0, abs(c))


@pytest.mark.parametrize("nside,spin", [(64, 0), (64, 2), (1024, 0), (1024, 2)])
def test_polar_skip_prefix_is_conservative(nside, spin):
    lmax = 3 * nside - 1
    na = sht._nactive(nside, lmax, spin, True)
    z, sin = sht._ring_slots(nside, spin)
    ofs = max(100.0, 0.01 * lmax)
    b = lmax * sin + ofs
    mlim = spin * np.abs(z) + np.sqrt(np.maximum(b * b - spin * spin * sin * sin, 0.0))
    for m in range(0, lmax + 1, max(1, lmax // 50)):
        needed = np.flatnonzero(mlim >= m)
        assert na[m] >= (needed.max() + 1 if len(needed) else 0)
    assert np.all(np.diff(na) <= 0)
    assert np.all(sht._nactive(nside, lmax, spin, False) == len(z))


def test_ring_slot_order_visits_every_ring_once():
    for spin, n in ((0, 16), (2, 16)):
        z, _ = sht._ring_slots(n, spin)
        assert np.all(np.diff(np.abs(z)) >= -1e-15)


def test_nmtbin_matches_oracle_linear_bands():
    b = bins.NmtBin.from_nside_linear(64, 16)
    ells, bpws, w = master.linear_bins(191, 16)
    assert b.get_n_bands() == bpws.max() + 1
    for k in range(b.get_n_bands()):
        assert np.array_equal(b.get_ell_list(k), ells[bpws == k])
        assert np.allclose(b.get_weight_list(k), w[bpws == k])
    assert np.allclose(b.get_effective_ells(), np.bincount(bpws, weights=ells * w))


def test_nmtbin_errors_and_dell():
    with pytest.raises(ValueError):
        bins.NmtBin(bpws=[], ells=[])
    b = bins.NmtBin.from_lmax_linear(40, 4, is_Dell=True)
    f = b.get_fell_list(0)
    ell = b.get_ell_list(0)
    assert np.allclose(f, ell * (ell + 1) / (2 * np.pi))
    e = bins.NmtBin.from_edges([2, 10, 20], [10, 20, 41])
    assert e.get_n_bands() == 3 and e.lmax == 40


@pytest.mark.parametrize("lt,le,dl", [(6, 3, 2), (16, 4, 3), (30, 30, 0), (20, 5, 30), (39, 1, 0), (25, 10, 20)])
def test_toeplitz_pair_list_covers_the_exact_entries(lt, le, dl):
    from gmaster_amd.workspaces import _toeplitz_pairs
    from oracle.toeplitz import exact_pairs

    L = 41
    lo, hi = _toeplitz_pairs(L, lt, le, dl, "cpu")
    r, c = exact_pairs(L, lt, le, dl)
    want = np.unique(np.minimum(r, c) * L + np.maximum(r, c))
    got = lo.numpy().astype(np.int64) * L + hi.numpy()
    assert np.all(lo.numpy() <= hi.numpy())
    assert len(np.unique(got)) == len(got)  # each pair once
    np.testing.assert_array_equal(np.sort(got), want)


@pytest.mark.parametrize("L", [41, 200])
@pytest.mark.parametrize("lt,le,dl", [(6, 3, 2), (16, 4, 3), (30, 30, 0), (20, 5, 30), (39, 1, 0), (25, 10, 20),
                                      (150, 64, 70), (150, 20, 64), (120, 10, 128)])
def test_toeplitz_row_items_cover_the_exact_entries(L, lt, le, dl):
    """The wave items of gm_coupling_rows (row l1, columns l2_0 .. l2_0 + nlive - 1, at most
    64 per item, inside the triangle) and the single pairs beside them cover each exact entry
    of the fill exactly once (band widths that leave a short remainder split it off)."""
    from gmaster_amd.workspaces import _toeplitz_items
    from oracle.toeplitz import exact_pairs

    if lt >= L - 1:
        pytest.skip("l_toeplitz must be below lmax")
    items, plo, phi = _toeplitz_items(L, lt, le, dl, "cpu")
    it = items.numpy().astype(np.int64)
    l1, l20, n = it[:, 0], it[:, 1], it[:, 2]
    assert np.all((n >= 1) & (n <= 64)) and np.all(l1 <= l20) and np.all(l20 + n <= L)
    plo, phi = plo.numpy().astype(np.int64), phi.numpy().astype(np.int64)
    assert np.all(plo <= phi) and np.all(phi < L)
    got = np.concatenate([a * L + b + np.arange(k) for a, b, k in it] + [plo * L + phi])
    r, c = exact_pairs(L, lt, le, dl)
    want = np.unique(np.minimum(r, c) * L + np.maximum(r, c))
    assert len(np.unique(got)) == len(got)
    np.testing.assert_array_equal(np.sort(got), want)


def test_toeplitz_sanity_messages():
    from gmaster_amd.workspaces import toeplitz_sanity

    class Pure:
        pure_e, pure_b = False, True

    with pytest.raises(ValueError, match="purification"):
        toeplitz_sanity(6, 3, 2, 12, (Pure(),))
    with pytest.raises(ValueError, match="must be positive"):
        toeplitz_sanity(6, 0, 2, 12)
    with pytest.raises(ValueError, match="smaller than"):
        toeplitz_sanity(6, 3, 12, 12)
    toeplitz_sanity(-1, -1, -1, 12)


def test_residual_targets_handle_lmax_below_the_grid():
    """Grid multipoles above lmax get weight 0 and are never used as indices (CPU tensors
    raise on an out-of-range index, as the device does)."""
    import torch

    from gmaster_amd import learned

    g = learned.ResidualGrid()
    lmax = 767
    rng = np.random.default_rng(0)
    ex = {k: torch.as_tensor(rng.uniform(1, 2, (lmax + 1, lmax + 1))) for k in learned.BLOCKS}
    ap = {k: v * 0.5 for k, v in ex.items()}
    t, w = learned.residual_targets(ex, ap, lmax, g)
    assert t.shape == (g.n_out,) and w.shape == (g.n_out,)
    J = np.repeat(g.l_grid, len(g.d_grid))
    I = J + np.tile(g.d_grid, len(g.l_grid))
    np.testing.assert_array_equal(w.numpy().reshape(4, -1)[0], (I <= lmax).astype(np.float32))
    assert bool(torch.all(t[w == 0] == 0))
    # S - S_T = S / 2 here, so e = 0.5 S_ij / sqrt(D_i D_j) at valid points
    col = 2 * np.arange(lmax + 1) + 1
    S = ex["00"].numpy() / col[None, :]
    D = np.abs(np.diag(S))
    ok = I <= lmax
    want = 0.5 * S[J[ok], I[ok]] / np.sqrt(D[J[ok]] * D[I[ok]])
    np.testing.assert_allclose(t.numpy()[: len(J)][ok], want, rtol=1e-6)


def test_coupling_precision_setting():
    import gmaster_amd as nmt
    assert nmt.coupling_precision() == "auto"
    nmt.set_coupling_precision("fp32")
    assert nmt.coupling_precision() == "fp32"
    nmt.set_coupling_precision("auto")
    with pytest.raises(KeyError):
        nmt.set_coupling_precision("fp16")


@pytest.mark.parametrize("nside", [16, 64, 256, 1024, 2048, 4096, 8192])
def test_residual_polar_split_is_whole_workgroups_or_the_fallback(nside):
    """The FP16 analysis's FP64 polar split: a whole number of analysis workgroups between
    the sin^2 = 1/24 and 1/96 splits when one exists there, else the 1/48 split; never
    closer to the pole than 1/96 nor further than 1/24 (checked on the ring geometry)."""
    from gmaster_amd import batched

    z, sin = sht._ring_slots(nside, 0)
    s2 = sin * sin
    j = batched.residual_split(nside)
    lo = int(np.count_nonzero(s2 >= 1.0 / 24.0))
    hi = int(np.count_nonzero(s2 >= 1.0 / 96.0))
    assert lo <= j <= hi
    blocks = [k * batched.ANA_BLOCK for k in range(1, 2 * nside // batched.ANA_BLOCK + 1)]
    inside = [b for b in blocks if lo <= b <= hi]
    if inside:
        assert j == max(inside) and j % batched.ANA_BLOCK == 0
    else:
        assert j == int(np.count_nonzero(s2 >= 1.0 / 48.0))
    # ring pairs are ordered from the equator: every pair before the split is at least as far
    # from the pole as every pair after it
    assert np.all(np.diff(s2) <= 0)
    if 0 < j < len(s2):
        assert s2[j - 1] >= s2[j]
