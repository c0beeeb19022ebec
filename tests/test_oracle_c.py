"""The C/OpenMP oracle (CPU baseline) agrees with the numpy oracle (no GPU)."""

import os
import subprocess

import numpy as np
import pytest

from oracle import coupling, sht
from oracle.geometry import Rings

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def cb():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle", "c")], check=True)
    from oracle.c import cpu_baseline

    return cpu_baseline


def test_c_legendre_analysis_matches_numpy_oracle(cb):
    nside, lmax, s = 8, 23, 2
    L = lmax + 1
    rng = np.random.default_rng(0)
    R = Rings(nside)
    qu = rng.normal(size=(2, R.npix))
    orders = np.arange(-(L - 1), L)
    G = sht.phased_spectrum(qu[0] + 1j * qu[1], R, orders)
    ap = sht.analysis_plus(G, R, sht.AlmInfo(lmax), s)  # (L, 2L-1)
    out = cb.lat_ana_s(nside, lmax, s, G, np.arange(L))
    ell = np.arange(L)
    for m in range(L):
        direct = (-1) ** s * (out[m, :, 0] + 1j * out[m, :, 1])
        assert np.allclose(direct[m:], ap[m:, L - 1 + m], atol=1e-12)
        if m:
            mirror = (-1) ** s * (-1.0) ** (ell + s) * (out[m, :, 2] + 1j * out[m, :, 3])
            assert np.allclose(mirror[m:], ap[m:, L - 1 - m], atol=1e-12)


def test_c_coupling_matches_quadrature(cb):
    lmax = 40
    w = np.random.default_rng(4).uniform(size=lmax + 1)
    q = coupling.coupling_blocks(w, lmax)
    raw = cb.coupling_rows(lmax, w / (4 * np.pi), np.arange(lmax + 1))
    col = 2 * np.arange(lmax + 1) + 1
    for k in range(4):
        M = np.zeros((lmax + 1, lmax + 1))
        for l1 in range(lmax + 1):
            M[l1, l1:] = raw[l1, l1:, k] * col[l1:]
            M[l1:, l1] = raw[l1, l1:, k] * col[l1]
        # the quadrature oracle carries ~1e-12 round-off relative to max |M| (the l3 recursion
        # matches exact Racah sums; see tests/test_gpu_master.py)
        assert np.max(np.abs(M - q[k])) < 1e-11 * np.max(np.abs(q[k]))


def test_n3j_count():
    from oracle.c.cpu_baseline import n3j

    assert n3j(3) == sum(min(a + b, 3) - (b - a) + 1 for a in range(4) for b in range(a, 4))
    assert abs(n3j(12287) - 4.64e11) / 4.64e11 < 0.01


def test_column_entry_matches_full_spectrum_entry():
    """``lat_ana_s_cols`` (per-order columns, for spectra too large for the host) gives
    exactly what ``lat_ana_s`` gives from the full spectrum."""
    import subprocess
    import os
    subprocess.run(["make", "-s", "-C", os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                                      "oracle", "c")], check=True)
    from oracle.c import cpu_baseline as cb
    nside, lmax, spin = 8, 20, 2
    nring, L = 4 * nside - 1, lmax + 1
    rng = np.random.default_rng(3)
    spec = rng.normal(size=(nring, 2 * L - 1)) + 1j * rng.normal(size=(nring, 2 * L - 1))
    orders = np.array([0, 1, 5, 13, 20])
    full = cb.lat_ana_s(nside, lmax, spin, spec, orders)
    cols = cb.lat_ana_s_cols(nside, lmax, spin, spec[:, L - 1 + orders].T, spec[:, L - 1 - orders].T, orders)
    assert np.array_equal(full, cols)


@pytest.mark.parametrize("extended", [False, True])
@pytest.mark.parametrize("spin", [0, 1, 2])
def test_c_legendre_synthesis_matches_numpy_oracle(cb, spin, extended):
    """``gmo_lat_syn_s`` (direct and mirror channels; FP64 and extended precision) against
    ``synthesis_plus`` / ``synthesis_s0`` of the numpy oracle, every order of a small
    geometry."""
    nside, lmax = 8, 23
    L = lmax + 1
    R = Rings(nside)
    ai = sht.AlmInfo(lmax)
    rng = np.random.default_rng(10 + spin)
    ap = rng.normal(size=(L, 2 * L - 1)) + 1j * rng.normal(size=(L, 2 * L - 1))
    ell = np.arange(L)
    ap[ell < spin] = 0
    for m in range(L):
        ap[:m, L - 1 + m] = 0
        ap[:m, L - 1 - m] = 0
    orders = np.arange(L)
    sg = (-1.0) ** spin
    cd = sg * ap[:, L - 1 + orders].T
    cm = sg * ((-1.0) ** (ell + spin))[None, :] * ap[:, L - 1 - orders].T
    out = cb.lat_syn_s(nside, lmax, spin, cd, cm, orders, extended=extended)
    if spin == 0:
        alm = np.zeros(ai.nelem, dtype=np.complex128)
        for m in range(L):
            alm[ai.mstart[m] + np.arange(m, L)] = ap[m:, L - 1 + m]
        ref = sht.synthesis_s0(alm, R, ai)
        got = out[:, :, 0] + 1j * out[:, :, 1]
        assert np.max(np.abs(got.T - ref)) < 1e-12 * np.max(np.abs(ref))
        return
    F = sht.synthesis_plus(ap, R, ai, spin)
    scale = np.max(np.abs(F))
    for m in range(L):
        assert np.max(np.abs(out[m, :, 0] + 1j * out[m, :, 1] - F[:, L - 1 + m])) < 1e-12 * scale
        if m:
            assert np.max(np.abs(out[m, :, 2] + 1j * out[m, :, 3] - F[:, L - 1 - m])) < 1e-12 * scale


def test_c_extended_synthesis_bounds_fp64_at_the_pole(cb):
    """At Nside 4096, m = 0, the FP64 and extended-precision syntheses agree to ~1e-10 of the
    largest value, with the difference largest at the polar rings (the FP64 recurrence's
    conditioning as x = cos(theta) -> 1) and ~1e-12 in the belt."""
    nside = 4096
    lmax = 3 * nside - 1
    L = lmax + 1
    rng = np.random.default_rng(4)
    orders = np.array([0])
    cd = rng.normal(size=(1, L)) + 1j * rng.normal(size=(1, L))
    cd[:, 0] = cd[:, 0].real
    cm = np.zeros_like(cd)
    a = cb.lat_syn_s(nside, lmax, 0, cd, cm, orders)
    b = cb.lat_syn_s(nside, lmax, 0, cd, cm, orders, extended=True)
    d = np.abs((a[0, :, 0] - b[0, :, 0]) + 1j * (a[0, :, 1] - b[0, :, 1]))
    scale = np.max(np.abs(b[0, :, 0] + 1j * b[0, :, 1]))
    assert d.max() / scale < 1e-9
    belt = slice(nside, 3 * nside)
    assert d[belt].max() / scale < 1e-11
    assert d.max() > 10 * d[belt].max()


def test_c_synthesis_is_the_analysis_transpose(cb):
    """<S c, G> = <c, A G> for the direct channel at a larger geometry (Nside 64, lmax 191),
    spin 2: the two C entries walk the same lambda rows in opposite directions."""
    nside, lmax, spin = 64, 191, 2
    L, nring = lmax + 1, 4 * nside - 1
    rng = np.random.default_rng(8)
    orders = np.array([0, 1, 2, 50, 120, 191])
    cd = rng.normal(size=(len(orders), L)) + 1j * rng.normal(size=(len(orders), L))
    cm = rng.normal(size=(len(orders), L)) + 1j * rng.normal(size=(len(orders), L))
    for k, m in enumerate(orders):
        cd[k, :max(m, spin)] = 0
        cm[k, :max(m, spin)] = 0
    syn = cb.lat_syn_s(nside, lmax, spin, cd, cm, orders)
    spec = rng.normal(size=(nring, 2 * L - 1)) + 1j * rng.normal(size=(nring, 2 * L - 1))
    ana = cb.lat_ana_s(nside, lmax, spin, spec, orders)
    for k, m in enumerate(orders):
        # real bilinear forms: sum_r S[r] g[r] = sum_l c[l] (A g)[l] with lambda real
        gd = spec[:, L - 1 + m]
        lhs = np.sum((syn[k, :, 0] + 1j * syn[k, :, 1]) * gd)
        rhs = np.sum(cd[k] * (ana[k, :, 0] + 1j * ana[k, :, 1]))
        assert abs(lhs - rhs) < 1e-10 * max(abs(lhs), 1.0)
        if m:  # the mirror channel of the analysis reads the reflected ring
            lhs = np.sum((syn[k, :, 2] + 1j * syn[k, :, 3]) * spec[:, L - 1 - m])
            rhs = np.sum(cm[k] * (ana[k, :, 2] + 1j * ana[k, :, 3]))
            assert abs(lhs - rhs) < 1e-10 * max(abs(lhs), 1.0), (m, lhs, rhs)
