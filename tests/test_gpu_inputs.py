"""GPU parity of the input side: C1/C2 mask apodization (``gm_mask_apodize``) against the
oracle's nearest-masked-pixel definition, and the Gaussian alms of synfast
(``gm_gaussian_alms``, ``gm_gauss_draws``) against the oracle's Philox4x32-10 restatement."""

import numpy as np
import pytest
import torch

from oracle import apodize as oap
from oracle import sht as osht
from oracle import synfast as osf
from oracle.geometry import pix2vec_ring

pytestmark = pytest.mark.gpu


def _holes_mask(nside, seed, nholes=12, cap=True, weights=True):
    rng = np.random.default_rng(seed)
    npix = 12 * nside * nside
    vec = pix2vec_ring(nside, np.arange(npix))
    mask = np.ones(npix)
    for _ in range(nholes):
        c = rng.normal(size=3)
        c /= np.linalg.norm(c)
        mask[vec @ c > np.cos(np.radians(rng.uniform(2, 15)))] = 0.0
    if cap:
        mask[vec[:, 2] > 0.97] = 0.0
    if weights:
        mask *= rng.uniform(0.5, 1.0, npix)
    return mask


@pytest.mark.parametrize("apotype", ["C1", "C2"])
@pytest.mark.parametrize("nside,apo", [(16, 10.0), (64, 5.0), (256, 2.0), (256, 0.3)])
def test_apodization_matches_oracle(nside, apo, apotype):
    import gmaster_amd as nmt
    mask = _holes_mask(nside, nside + int(10 * apo))
    got = np.asarray(nmt.mask_apodization(mask, apo, apotype))
    want = oap.mask_apodization(mask, apo, apotype)
    changed = np.sum(want != mask)
    err = np.max(np.abs(got - want))
    print(f"nside {nside} apo {apo} {apotype}: {changed} pixels apodized, max err {err:.1e}", flush=True)
    assert changed > 100
    assert err < 1e-10


def test_apodization_edge_cases():
    import gmaster_amd as nmt
    nside = 32
    npix = 12 * nside * nside
    ones = np.ones(npix)
    assert np.array_equal(np.asarray(nmt.mask_apodization(ones, 3.0, "C1")), ones)  # nothing masked
    single = ones.copy()
    single[5000] = 0.0  # one masked pixel in an equatorial ring
    want = oap.mask_apodization(single, 4.0, "C2")
    assert np.max(np.abs(np.asarray(nmt.mask_apodization(single, 4.0, "C2")) - want)) < 1e-10
    pole = ones.copy()
    pole[:4] = 0.0  # the first ring (4 pixels at the north pole)
    want = oap.mask_apodization(pole, 6.0, "C1")
    assert np.max(np.abs(np.asarray(nmt.mask_apodization(pole, 6.0, "C1")) - want)) < 1e-10
    assert np.array_equal(np.asarray(nmt.mask_apodization(single, 0.0)), single)
    with pytest.raises(ValueError):
        nmt.mask_apodization(single, 1e-4, "C1")  # below the pixel size
    with pytest.raises(ValueError):
        nmt.mask_apodization(single, 1.0, "C3")


def test_apodization_at_nside_1024_on_sampled_pixels():
    """Config-2 geometry (|b| > 20 deg plus 500 holes of 5 arcmin, C2 over 1 deg): the device
    result on 200k sampled pixels (a third of them within the scale of the mask) against the
    nearest masked pixel from a k-d tree over every masked pixel."""
    import gmaster_amd as nmt
    from scipy.spatial import cKDTree
    nside = 1024
    npix = 12 * nside * nside
    vec = pix2vec_ring(nside, np.arange(npix))
    rng = np.random.default_rng(2)
    mask = (np.abs(vec[:, 2]) > np.sin(np.radians(20.0))).astype(np.float64)
    c = rng.normal(size=(500, 3))
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    tree = cKDTree(vec)
    for centre in c:
        mask[tree.query_ball_point(centre, 2 * np.sin(np.radians(5 / 60) / 2))] = 0.0
    got = np.asarray(nmt.mask_apodization(mask, 1.0, "C2"))
    masked = np.flatnonzero(mask <= 0)
    dist, _ = cKDTree(vec[masked]).query(vec, workers=-1)
    dm = 0.5 * dist**2
    thr = 1 - np.cos(np.radians(1.0))
    near = np.flatnonzero((mask > 0) & (dm < thr))
    far = np.flatnonzero((mask > 0) & (dm >= thr))
    sample = np.concatenate([rng.choice(near, 70_000, replace=False), rng.choice(far, 130_000, replace=False)])
    x = np.sqrt(dm[sample] / thr)
    want = np.where(dm[sample] < thr, 0.5 * (1 - np.cos(np.pi * np.minimum(x, 1))), 1.0) * mask[sample]
    err = np.max(np.abs(got[sample] - want))
    print(f"Nside 1024 config-2 mask: {len(near)} pixels apodized, sampled max err {err:.1e}", flush=True)
    assert err < 1e-10
    assert np.array_equal(got[masked], mask[masked])


def test_gauss_draws_match_oracle():
    from gmaster_amd import _lib
    n, ndraw, seed = 100_000, 3, 0x1234_5678_9ABC
    out = torch.empty(2 * n * ndraw, dtype=torch.float64, device="cuda")
    _lib.call("gm_gauss_draws", seed, n, ndraw, _lib.ptr(out), _lib.stream_ptr())
    got = out.cpu().numpy().reshape(n, ndraw, 2)
    for d in range(ndraw):
        z0, z1 = osf.gauss_pairs(seed, np.arange(n), d)
        assert np.max(np.abs(got[:, d, 0] - z0)) < 1e-13
        assert np.max(np.abs(got[:, d, 1] - z1)) < 1e-13


def test_synfast_matches_oracle_alms_and_maps():
    import gmaster_amd as nmt
    nside, lmax, seed = 32, 95, 99
    L = lmax + 1
    ell = np.arange(L)
    tt = 1.0 / (ell + 10.0) ** 2
    ee, bb = 0.1 * tt, 0.01 * tt
    te = 0.3 * np.sqrt(tt * ee)
    z = np.zeros(L)
    cls = np.array([tt, te, z, ee, z, bb])  # T, E, B upper triangle
    beam = np.array([np.exp(-ell * (ell + 1) * 1e-4), np.exp(-ell * (ell + 1) * 2e-4)])
    maps = np.asarray(nmt.synfast_spherical(nside, cls, [0, 2], beam=beam, seed=seed))
    cov = np.zeros((L, 3, 3))
    k = 0
    for i in range(3):
        for j in range(i, 3):
            cov[:, i, j] = cov[:, j, i] = cls[k]
            k += 1
    w, v = np.linalg.eigh(cov)
    root = v * np.sqrt(np.maximum(w, 0))[:, None, :]
    alm = osf.gaussian_alms(root, lmax, seed)
    lidx = np.concatenate([np.arange(mm, L) for mm in range(L)])
    alm = alm * np.stack([beam[0], beam[1], beam[1]])[:, lidx]
    want = np.concatenate([osht.alm2map(alm[0], 0, nside, lmax), osht.alm2map(alm[1:], 2, nside, lmax)])
    err = np.max(np.abs(maps - want)) / np.max(np.abs(want))
    print(f"synfast maps vs oracle: rel {err:.1e}", flush=True)
    assert maps.shape == (3, 12 * nside * nside)
    assert err < 1e-11
    with pytest.raises(ValueError):
        nmt.synfast_spherical(nside, cls[:5], [0, 2])
    bad = cls.copy()
    bad[1] = 10 * np.sqrt(tt * ee)  # |TE| > sqrt(TT EE)
    with pytest.raises(ValueError):
        nmt.synfast_spherical(nside, bad, [0, 2])


@pytest.mark.parametrize("nside,apo", [(16, 4.0), (32, 3.0), (64, 1.5)])
def test_smooth_apodization_matches_oracle(nside, apo):
    import gmaster_amd as nmt
    mask = (_holes_mask(nside, nside, nholes=6, weights=False) > 0).astype(np.float64)
    got = np.asarray(nmt.mask_apodization(mask, apo, "Smooth"))
    want = oap.smooth_apodization(mask, apo)
    err = np.max(np.abs(got - want))
    print(f"Smooth nside {nside} apo {apo}: max err {err:.1e}", flush=True)
    assert err < 1e-10
    assert np.array_equal(np.asarray(nmt.mask_apodization(np.ones(12 * nside * nside), apo, "Smooth")),
                          np.ones(12 * nside * nside))
