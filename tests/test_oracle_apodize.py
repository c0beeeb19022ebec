"""Pins of the apodization oracle (``oracle.apodize``) by its definition, CPU only."""

import numpy as np
import pytest

from oracle.apodize import mask_apodization
from oracle.geometry import pix2vec_ring


def _holes_mask(nside, seed):
    rng = np.random.default_rng(seed)
    npix = 12 * nside * nside
    vec = pix2vec_ring(nside, np.arange(npix))
    mask = np.ones(npix)
    for _ in range(4):
        c = rng.normal(size=3)
        c /= np.linalg.norm(c)
        mask[vec @ c > np.cos(np.radians(rng.uniform(8, 25)))] = 0.0
    mask[vec[:, 2] > 0.95] = 0.0  # a polar cap
    return mask * rng.uniform(0.5, 1.0, npix) * (mask > 0)


@pytest.mark.parametrize("apotype", ["C1", "C2"])
@pytest.mark.parametrize("nside,apo", [(8, 20.0), (16, 10.0)])
def test_apodization_matches_brute_force_definition(nside, apo, apotype):
    mask = _holes_mask(nside, nside)
    got = mask_apodization(mask, apo, apotype)
    vec = pix2vec_ring(nside, np.arange(mask.size))
    masked = vec[mask <= 0]
    want = mask.copy()
    thr = 1 - np.cos(np.radians(apo))
    n_aff = 0
    for p in np.flatnonzero(mask > 0):
        d = np.min(0.5 * np.sum((masked - vec[p]) ** 2, axis=1))
        if d < thr:
            n_aff += 1
            x = np.sqrt(d / thr)
            f = x - np.sin(2 * np.pi * x) / (2 * np.pi) if apotype == "C1" else 0.5 * (1 - np.cos(np.pi * x))
            want[p] *= f
    assert n_aff > 50
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-13)


def test_apodization_factors_at_the_ends():
    """f(0) = 0 and f(1) = 1 for both types: a pixel next to a hole is nearly zeroed, a pixel
    beyond the scale is untouched; masked pixels stay masked."""
    nside = 16
    mask = _holes_mask(nside, 3)
    out = mask_apodization(mask, 10.0, "C2")
    assert np.all(out[mask <= 0] == mask[mask <= 0])
    assert np.all(out <= mask + 1e-15)
    assert np.all(out >= 0)


def test_ang2pix_and_largest_pixel_radius():
    """ang2pix_ring inverts pix2ang_ring on every centre; over random points the distance to
    the containing pixel's centre never exceeds max_pixrad and comes within 2% of it."""
    from oracle.geometry import ang2pix_ring, max_pixrad, pix2ang_ring

    for nside in (1, 2, 4, 8):
        npix = 12 * nside * nside
        th, ph = pix2ang_ring(nside, np.arange(npix))
        assert np.array_equal(ang2pix_ring(nside, th, ph), np.arange(npix))
        rng = np.random.default_rng(nside)
        z = rng.uniform(-1, 1, 400_000)
        phi = rng.uniform(0, 2 * np.pi, 400_000)
        pix = ang2pix_ring(nside, np.arccos(z), phi)
        assert np.bincount(pix, minlength=npix).min() > 0
        v = np.stack([np.sqrt(1 - z * z) * np.cos(phi), np.sqrt(1 - z * z) * np.sin(phi), z], 1)
        ang = np.arccos(np.clip(np.sum(v * pix2vec_ring(nside, pix), 1), -1, 1))
        r = max_pixrad(nside)
        assert ang.max() <= r * (1 + 1e-12) and ang.max() > 0.98 * r


def test_smooth_apodization_zeroed_set_and_smoothing():
    from oracle.apodize import smooth_apodization
    from oracle.geometry import max_pixrad

    nside, apo = 16, 4.0
    mask = _holes_mask(nside, 7)
    mask = (mask > 0).astype(float)
    out = smooth_apodization(mask, apo)
    vec = pix2vec_ring(nside, np.arange(mask.size))
    masked = vec[mask <= 0]
    rbig = 2.5 * np.radians(apo) + max_pixrad(nside)
    d = np.array([np.min(0.5 * np.sum((masked - v) ** 2, axis=1)) for v in vec])
    zeroed = d < 1 - np.cos(rbig)
    assert zeroed.sum() > mask.size // 10
    # masked pixels stay zero; pixels far beyond the zeroed region stay near 1
    assert np.all(out[mask <= 0] == 0)
    far = d > 1 - np.cos(rbig + 5 * np.radians(apo))
    if far.any():
        assert np.all(np.abs(out[far] - 1) < 0.05)
    assert np.all(out < 1.2) and np.all(out > -0.2)
