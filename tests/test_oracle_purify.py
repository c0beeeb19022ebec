"""The purification oracle against what purification is for (CPU)."""

import numpy as np

from oracle import purify as op
from oracle import sht
from oracle.geometry import AlmInfo, Rings

NSIDE, LMAX, NITER = 16, 32, 3


def _cap_mask(theta0_deg=60.0, apo_deg=25.0):
    """1 near the north pole, C2-apodized to 0 between theta0 - apo and theta0 (smooth,
    vanishing with its first derivative at the edge)."""
    R = Rings(NSIDE)
    theta = np.repeat(R.theta, R.nphi)
    t0, a = np.radians(theta0_deg), np.radians(apo_deg)
    x = np.clip((t0 - theta) / a, 0.0, 1.0)
    return 0.5 * (1 - np.cos(np.pi * x))


def _sky(which, seed=0):
    ainfo = AlmInfo(LMAX)
    rng = np.random.default_rng(seed)
    cl = np.where(ainfo.ell >= 2, 1.0 / (ainfo.ell + 5.0) ** 2, 0.0) * (ainfo.ell <= LMAX // 2)
    a = (rng.normal(size=ainfo.nelem) + 1j * rng.normal(size=ainfo.nelem) * (ainfo.m > 0)) * np.sqrt(cl / 2)
    alm = np.zeros((2, ainfo.nelem), dtype=np.complex128)
    alm[0 if which == "E" else 1] = a
    return sht.alm2map(alm, 2, NSIDE, LMAX)


def test_full_sky_purification_adds_nothing():
    # the full-sky mask's spectrum is a_00 = sqrt(4 pi) alone: no derivative terms
    mask = np.ones(12 * NSIDE * NSIDE)
    amask = np.zeros(AlmInfo(LMAX).nelem, dtype=np.complex128)
    amask[0] = np.sqrt(4 * np.pi)
    qu = _sky("E", 1) + _sky("B", 2)
    plain = sht.map2alm(qu, 2, NSIDE, LMAX, NITER)
    pure = op.purify(mask, amask, qu, NSIDE, LMAX, (True, True), NITER)
    assert np.max(np.abs(pure - plain)) < 1e-9 * np.max(np.abs(plain))


def test_purified_b_suppresses_e_to_b_leakage():
    mask = _cap_mask()
    amask = sht.map2alm(mask[None], 0, NSIDE, LMAX, NITER)[0]
    qu = _sky("E", 3)  # a pure-E sky: every B mode of its masked map is leakage
    plain = sht.map2alm(qu * mask[None], 2, NSIDE, LMAX, NITER)
    pure = op.purify(mask, amask, qu, NSIDE, LMAX, (False, True), NITER)
    ainfo = AlmInfo(LMAX)
    w = np.where(ainfo.m == 0, 1.0, 2.0)
    sel = (ainfo.ell >= 4) & (ainfo.ell <= LMAX // 2)
    leak_plain = np.sum(w[sel] * np.abs(plain[1][sel]) ** 2)
    leak_pure = np.sum(w[sel] * np.abs(pure[1][sel]) ** 2)
    print(f"B leakage: plain {leak_plain:.3e}, purified {leak_pure:.3e}")
    assert leak_pure < 0.05 * leak_plain
    # the E modes are untouched when only B is purified
    assert np.array_equal(pure[0], plain[0])


def test_spin_products_are_conjugate_products():
    rng = np.random.default_rng(4)
    w, p = rng.normal(size=(2, 2, 40))
    out = op.spin_mul(w, p)
    z = np.conj(w[0] + 1j * w[1]) * (p[0] + 1j * p[1])
    assert np.allclose(out[0], z.real) and np.allclose(out[1], z.imag)
