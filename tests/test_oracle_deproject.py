"""The deprojection oracle against closed forms and invariants (CPU).

Full sky, a template that is one real spherical harmonic Y_{l0,0}: deprojection removes
that single mode, so the coupled signal bias is -C_{l0} / (M (2 l0 + 1)) at l0 and zero
elsewhere (M = w sum t^2, 1 up to the pixel quadrature), and the uncorrelated-noise bias
is -v / (M (2 l0 + 1)) for a uniform variance v per steradian.  The bias does not depend
on the basis of the template span.  The cleaned map is orthogonal to every template.
"""

import numpy as np
import pytest

from oracle import deproject as dp
from oracle import sht
from oracle.geometry import AlmInfo

NSIDE, LMAX, NITER = 16, 12, 3
NPIX = 12 * NSIDE * NSIDE
W = 4 * np.pi / NPIX


def _mode_map(spin, l0, comp=0):
    ainfo = AlmInfo(LMAX)
    nmaps = 1 if spin == 0 else 2
    a = np.zeros((nmaps, ainfo.nelem), dtype=np.complex128)
    a[comp, ainfo.idx(l0, 0)] = 1.0
    return sht.alm2map(a, spin, NSIDE, LMAX)


def test_pinvh_matches_numpy_pinv():
    rng = np.random.default_rng(0)
    q, _ = np.linalg.qr(rng.normal(size=(5, 5)))
    ev = np.array([3.0, 1.0, 0.5, 1e-3, 1e-14])
    a = (q * ev) @ q.T
    np.testing.assert_allclose(dp.pinvh(a, 1e-10), np.linalg.pinv(a, rcond=1e-10, hermitian=True), atol=1e-10)
    b = (q * np.array([3.0, 1.0, 0.5, 1e-3, 0.2])) @ q.T
    np.testing.assert_allclose(dp.pinvh(b, 0), np.linalg.inv(b), rtol=1e-10, atol=1e-10)
    # the cut drops the 1e-14 direction entirely
    assert abs(q[:, 4] @ dp.pinvh(a, 1e-10) @ q[:, 4]) < 1e-8


def test_cleaned_maps_are_orthogonal_and_exact_fits_vanish():
    rng = np.random.default_rng(1)
    mask = rng.uniform(0.2, 1.0, NPIX)
    temps = rng.normal(size=(3, 2, NPIX))
    maps = rng.normal(size=(2, NPIX)) + 0.7 * temps[0] - 1.3 * temps[2]
    clean, tm, alphas, iM = dp.deproject(mask, maps, temps)
    dots = np.einsum("iap,ap->i", tm, clean) * W
    assert np.max(np.abs(dots)) < 1e-12 * np.sqrt(np.sum(clean ** 2) * np.sum(tm ** 2)) * W
    exact = 2.5 * temps[1] - 0.4 * temps[0]
    clean2, _, alphas2, _ = dp.deproject(mask, exact, temps)
    np.testing.assert_allclose(alphas2, [-0.4, 2.5, 0.0], atol=1e-12)
    assert np.max(np.abs(clean2)) < 1e-12


@pytest.mark.parametrize("l0", [3, 7])
def test_signal_bias_of_one_removed_mode(l0):
    mask = np.ones(NPIX)
    rng = np.random.default_rng(l0)
    t = _mode_map(0, l0)
    maps = rng.normal(size=(1, NPIX))
    f = dp.Field(mask, maps, t[None], 0, NSIDE, LMAX, NITER)
    cl = rng.uniform(0.5, 2.0, size=(1, LMAX + 1))
    M = W * np.sum(t * t)
    bias = dp.deprojection_bias(f, f, cl)
    expect = np.zeros((1, LMAX + 1))
    expect[0, l0] = -cl[0, l0] / (M * (2 * l0 + 1))
    np.testing.assert_allclose(bias, expect, atol=1e-9 * cl.max())
    # a cross spectrum against the same map without templates loses the mode once
    g = dp.Field(mask, maps, None, 0, NSIDE, LMAX, NITER)
    np.testing.assert_allclose(dp.deprojection_bias(f, g, cl), expect, atol=1e-9 * cl.max())
    np.testing.assert_allclose(dp.deprojection_bias(g, f, cl), expect, atol=1e-9 * cl.max())


def test_noise_bias_of_one_removed_mode():
    l0, v = 5, 0.37
    mask = np.ones(NPIX)
    t = _mode_map(0, l0)
    f = dp.Field(mask, np.random.default_rng(2).normal(size=(1, NPIX)), t[None], 0, NSIDE, LMAX, NITER)
    M = W * np.sum(t * t)
    bias = dp.uncorr_noise_deprojection_bias(f, np.full(NPIX, v))
    expect = np.zeros((1, LMAX + 1))
    expect[0, l0] = -v / (M * (2 * l0 + 1))
    np.testing.assert_allclose(bias, expect, atol=1e-10)


def test_bias_depends_only_on_the_template_span():
    rng = np.random.default_rng(3)
    mask = rng.uniform(0.3, 1.0, NPIX)
    ta, tb = _mode_map(0, 4), _mode_map(0, 9)
    maps = rng.normal(size=(1, NPIX))
    cl = rng.uniform(0.5, 2.0, size=(1, LMAX + 1))
    f1 = dp.Field(mask, maps, np.stack([ta, tb]), 0, NSIDE, LMAX, NITER)
    f2 = dp.Field(mask, maps, np.stack([ta + tb, 2.0 * ta - tb]), 0, NSIDE, LMAX, NITER)
    np.testing.assert_allclose(f1.maps, f2.maps, atol=1e-12)
    np.testing.assert_allclose(dp.deprojection_bias(f1, f1, cl), dp.deprojection_bias(f2, f2, cl), atol=1e-12)
    v = rng.uniform(0.1, 1.0, NPIX)
    np.testing.assert_allclose(dp.uncorr_noise_deprojection_bias(f1, v), dp.uncorr_noise_deprojection_bias(f2, v),
                               atol=1e-12)


def test_spin2_bias_of_one_removed_e_mode():
    l0 = 6
    mask = np.ones(NPIX)
    rng = np.random.default_rng(4)
    t = _mode_map(2, l0, comp=0)  # a pure-E template
    f = dp.Field(mask, rng.normal(size=(2, NPIX)), t[None], 2, NSIDE, LMAX, NITER)
    cl = np.zeros((4, LMAX + 1))
    cl[0] = rng.uniform(0.5, 2.0, LMAX + 1)  # EE
    cl[3] = rng.uniform(0.5, 2.0, LMAX + 1)  # BB
    M = W * np.sum(t * t)
    bias = dp.deprojection_bias(f, f, cl)
    expect = np.zeros((4, LMAX + 1))
    expect[0, l0] = -cl[0, l0] / (M * (2 * l0 + 1))
    np.testing.assert_allclose(bias, expect, atol=1e-9)


def test_filter_alms_by_definition():
    rng = np.random.default_rng(5)
    ainfo = AlmInfo(LMAX)
    alms = rng.normal(size=(2, ainfo.nelem)) + 1j * rng.normal(size=(2, ainfo.nelem))
    spec = rng.normal(size=(2, 3, LMAX + 1))
    out = dp.filter_alms(alms, spec, LMAX)
    for o in range(3):
        for m in range(LMAX + 1):
            for l in range(m, LMAX + 1):
                i = ainfo.idx(l, m)
                assert abs(out[o, i] - (spec[0, o, l] * alms[0, i] + spec[1, o, l] * alms[1, i])) < 1e-13
