"""NaMaster's Toeplitz approximation of a coupling matrix (oracle).

Restates ``gmaster/workspaces.py:1141-1194`` (``_apply_toeplitz``, after Louis et al.
2020) step by step, on the master coefficients ``S[l, l'] = M[l, l'] / (2 l' + 1)``
(the reference divides the column factor out before and restores it after,
``gmaster/workspaces.py:1834-1840``), and ``_toeplitz_sanity`` (``:1125-1138``).

Pinned by: exactness when the correlation ``S / sqrt(diag diag)`` is Toeplitz (then every
fill reproduces the matrix), symmetry of the output, and the untouched columns / band.
"""

import numpy as np


def sanity(l_toeplitz, l_exact, dl_band, lmax):
    if l_toeplitz <= 0:
        return
    if l_exact <= 0 or dl_band < 0:
        raise ValueError("`l_exact` and `dl_band` must be positive numbers")
    if l_exact > l_toeplitz:
        raise ValueError("`l_exact` must be `<= l_toeplitz")
    if l_toeplitz >= lmax or l_exact >= lmax or dl_band >= lmax:
        raise ValueError("`l_toeplitz`, `l_exact` and `dl_band` must be smaller than `l_max`")


def apply_toeplitz(S, l_toeplitz, l_exact, dl_band):
    """The fill sequence on a symmetric ``(n, n)`` matrix ``S``; returns a new matrix."""
    if l_toeplitz <= 0:
        return S.copy()
    S = np.asarray(S, dtype=np.float64)
    n = len(S)
    lt, le = l_toeplitz, l_exact
    diag = np.abs(np.diag(S))
    den = np.sqrt(diag * diag[lt])
    corr = np.where(den > 0, S[:, lt] / np.where(den > 0, den, 1), 0.0)
    idx = (np.abs(np.arange(n)[:, None] - np.arange(n)[None]) + lt) % n
    root = np.sqrt(diag[:, None] * diag[None])
    pure = corr[idx] * root
    res = pure.copy()
    res[:, : le + 1] = S[:, : le + 1]
    for off in range(dl_band + 1):
        i = np.arange(n - off)
        res[i + off, i] = np.diag(S, off)
    res[lt:, lt:] = pure[lt:, lt:]
    i = np.arange(n)
    res[i, i] = np.diag(S)
    for off in range(1, dl_band + 1):
        cnt = min(n - off, lt + 1)
        i = np.arange(cnt)
        res[i + off, i] = np.diag(S, off)[:cnt]
    # the reference guards this division by diag > 0 only (x / 0 stays inf or nan)
    with np.errstate(divide="ignore", invalid="ignore"):
        ccol = np.where(diag > 0, S[:, le] / np.sqrt(diag * diag[le]), 0.0)
    src = n - 1 - lt + le
    for off in range(lt - le + 1):
        col = off + le
        tgt = n - 1 - lt + col
        res[tgt:, col] = ccol[src: n - off] * np.sqrt(diag[col] * diag[tgt:])
    lower = np.tril(res)
    return lower + np.tril(res, -1).T


def exact_pairs(size, l_toeplitz, l_exact, dl_band):
    """``(rows, cols)`` of the exact entries the fill reads (``gmaster/workspaces.py:520-539``):
    the diagonal, column ``l_toeplitz``, columns ``<= l_exact`` and ``dl_band`` upper diagonals."""
    pairs = {(l, l) for l in range(size)}
    pairs.update((l, l_toeplitz) for l in range(size))
    pairs.update((l, c) for c in range(l_exact + 1) for l in range(size))
    pairs.update((l, l + off) for off in range(dl_band + 1) for l in range(size - off))
    enc = np.sort(np.fromiter((r * size + c for r, c in pairs), dtype=np.int64))
    return enc // size, enc % size


def apply_to_mcm_block(M, l_toeplitz, l_exact, dl_band):
    """Toeplitz-approximate a coupling block ``M`` (column factor (2 l' + 1) included)."""
    col = (2 * np.arange(len(M)) + 1)[None, :]
    return apply_toeplitz(M / col, l_toeplitz, l_exact, dl_band) * col


def entry_kind(n, l_toeplitz, l_exact, dl_band):
    """``(n, n)`` int array: 0 where the fill keeps the exact value, 1 where it writes the pure
    Toeplitz extension, 2 where it writes the corner extrapolation from column l_exact
    (read off the fill sequence of ``apply_toeplitz``)."""
    lt, le, dl = l_toeplitz, l_exact, dl_band
    kind = np.ones((n, n), dtype=np.int64)
    kind[:, : le + 1] = 0
    for off in range(dl + 1):
        i = np.arange(n - off)
        kind[i + off, i] = 0
    kind[lt:, lt:] = 1
    i = np.arange(n)
    kind[i, i] = 0
    for off in range(1, dl + 1):
        cnt = min(n - off, lt + 1)
        i = np.arange(cnt)
        kind[i + off, i] = 0
    for off in range(lt - le + 1):
        col = off + le
        kind[n - 1 - lt + col:, col] = 2
    lower = np.tril(kind)
    return lower + np.tril(kind, -1).T


def correct(M, l_toeplitz, l_exact, dl_band, l_grid, d_grid, e):
    """Learned correction of a Toeplitz block ``M`` (column factor included): every
    non-exact element gets ``e(j, d) sqrt(D_j D_{j+d}) (2 col + 1)``, e bilinear in
    (log j, log d) on the grids, zero outside them."""
    M = np.array(M, dtype=np.float64)
    n = len(M)
    kind = entry_kind(n, l_toeplitz, l_exact, dl_band)
    lg = np.asarray(l_grid, np.float64)
    dg = np.asarray(d_grid, np.float64)
    e = np.asarray(e, np.float64).reshape(len(lg), len(dg))
    D = np.abs(np.diag(M)) / (2 * np.arange(n) + 1)
    out = M.copy()
    for a in range(n):
        for b in range(n):
            i, j = max(a, b), min(a, b)
            d = i - j
            if kind[a, b] == 0 or j < lg[0] or j > lg[-1] or d < dg[0] or d > dg[-1]:
                continue
            p = min(np.searchsorted(lg, j, side="right") - 1, len(lg) - 2) if len(lg) > 1 else 0
            q = min(np.searchsorted(dg, d, side="right") - 1, len(dg) - 2) if len(dg) > 1 else 0
            p1, q1 = min(p + 1, len(lg) - 1), min(q + 1, len(dg) - 1)
            tl = (np.log(j) - np.log(lg[p])) / (np.log(lg[p1]) - np.log(lg[p])) if p1 != p else 0.0
            td = (np.log(d) - np.log(dg[q])) / (np.log(dg[q1]) - np.log(dg[q])) if q1 != q else 0.0
            v = (1 - tl) * ((1 - td) * e[p, q] + td * e[p, q1]) + tl * ((1 - td) * e[p1, q] + td * e[p1, q1])
            out[a, b] += v * np.sqrt(D[i] * D[j]) * (2 * b + 1)
    return out
