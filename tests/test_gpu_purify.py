"""GPU parity of E/B purification against the oracle (FP64 form).

Purified spin-2 fields (E, B, E+B; with templates; pre-masked input), their deprojection
bias, and the spin-1 transforms purification runs.  The oracle is pinned in
``tests/test_oracle_purify.py`` (the purified B modes of a pure-E sky carry ~1e-7 of the
ordinary E->B leakage); the same suppression is checked here on the device.
"""

import numpy as np
import pytest
import torch

from oracle import deproject as odp
from oracle import purify as op
from oracle import sht as osht
from oracle.geometry import AlmInfo, Rings

pytestmark = pytest.mark.gpu

NSIDE, LMAX, NITER = 16, 32, 3
NPIX = 12 * NSIDE * NSIDE


@pytest.fixture(autouse=True)
def _fp64_form():
    import gmaster_amd as nmt
    from gmaster_amd.field import clear_mask_cache
    nmt.set_form("fp64")
    clear_mask_cache()
    yield
    clear_mask_cache()


def _rel(got, ref):
    got, ref = np.asarray(got), np.asarray(ref)
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300))


def _cap_mask():
    R = Rings(NSIDE)
    theta = np.repeat(R.theta, R.nphi)
    x = np.clip((np.radians(60.0) - theta) / np.radians(25.0), 0.0, 1.0)
    return 0.5 * (1 - np.cos(np.pi * x))


def _e_sky(seed):
    ainfo = AlmInfo(LMAX)
    rng = np.random.default_rng(seed)
    cl = np.where(ainfo.ell >= 2, 1.0 / (ainfo.ell + 5.0) ** 2, 0.0) * (ainfo.ell <= LMAX // 2)
    a = (rng.normal(size=ainfo.nelem) + 1j * rng.normal(size=ainfo.nelem) * (ainfo.m > 0)) * np.sqrt(cl / 2)
    alm = np.zeros((2, ainfo.nelem), dtype=np.complex128)
    alm[0] = a
    return osht.alm2map(alm, 2, NSIDE, LMAX)


def test_spin1_transforms_match_oracle():
    from gmaster_amd import sht
    rng = np.random.default_rng(0)
    qu = rng.normal(size=(2, NPIX))
    got = sht.map2alm(torch.as_tensor(qu, device="cuda"), 1, LMAX, n_iter=NITER).cpu().numpy()
    assert _rel(got, osht.map2alm(qu, 1, NSIDE, LMAX, NITER)) < 1e-10
    alm = osht.map2alm(qu, 1, NSIDE, LMAX, 0)
    back = sht.alm2map(torch.as_tensor(alm, device="cuda"), 1, NSIDE, LMAX).cpu().numpy()
    assert _rel(back, osht.alm2map(alm, 1, NSIDE, LMAX)) < 1e-11


@pytest.mark.parametrize("task", [(False, True), (True, False), (True, True)])
def test_pure_fields_match_oracle(task):
    import gmaster_amd as nmt
    rng = np.random.default_rng(11)
    mask = _cap_mask()
    template = rng.normal(size=(1, 2, NPIX))
    maps = 0.3 * template[0] + _e_sky(5) + 0.05 * rng.normal(size=(2, NPIX))
    f = nmt.NmtField(mask, maps, templates=template, purify_e=task[0], purify_b=task[1], lmax=LMAX,
                     lmax_mask=LMAX, n_iter=NITER)
    o = odp.Field(mask, maps, template, 2, NSIDE, LMAX, NITER, task=task)
    print(f"task {task}: alphas {f.alphas} oracle {o.alphas}", flush=True)
    assert _rel(f.alphas, o.alphas) < 1e-12
    assert _rel(f.get_alms(), o.alm) < 1e-10
    assert _rel(f.get_maps(), o.maps) < 1e-10
    assert _rel(f.alm_temp.cpu(), o.alm_temp) < 1e-10
    # pre-masked input: the unmasked maps are recovered where the mask is positive, which
    # on a mask without zeros gives the same field
    pos = 0.2 + 0.8 * mask
    g = nmt.NmtField(pos, maps * pos[None, :], purify_e=task[0], purify_b=task[1], masked_on_input=True,
                     lmax=LMAX, lmax_mask=LMAX, n_iter=NITER)
    h = nmt.NmtField(pos, maps, purify_e=task[0], purify_b=task[1], lmax=LMAX, lmax_mask=LMAX, n_iter=NITER)
    assert _rel(g.get_alms(), h.get_alms()) < 1e-12


def test_purified_b_leakage_on_gpu():
    import gmaster_amd as nmt
    mask = _cap_mask()
    qu = _e_sky(3)
    plain = nmt.NmtField(mask, qu, lmax=LMAX, lmax_mask=LMAX, n_iter=NITER).get_alms().cpu().numpy()
    pure = nmt.NmtField(mask, qu, purify_b=True, lmax=LMAX, lmax_mask=LMAX, n_iter=NITER).get_alms().cpu().numpy()
    ainfo = AlmInfo(LMAX)
    w = np.where(ainfo.m == 0, 1.0, 2.0)
    sel = (ainfo.ell >= 4) & (ainfo.ell <= LMAX // 2)
    lp = np.sum(w[sel] * np.abs(plain[1][sel]) ** 2)
    lq = np.sum(w[sel] * np.abs(pure[1][sel]) ** 2)
    print(f"B leakage plain {lp:.3e} purified {lq:.3e}", flush=True)
    assert lq < 1e-4 * lp
    ref = op.purify(mask, osht.map2alm(mask[None], 0, NSIDE, LMAX, NITER)[0], qu, NSIDE, LMAX, (False, True), NITER)
    assert _rel(pure, ref) < 1e-10


def test_pure_deprojection_bias_matches_oracle():
    import gmaster_amd as nmt
    rng = np.random.default_rng(18)
    mask = _cap_mask()
    template = rng.normal(size=(1, 2, NPIX))
    maps = rng.normal(size=(2, NPIX))
    f = nmt.NmtField(mask, maps, templates=template, purify_b=True, lmax=LMAX, lmax_mask=LMAX, n_iter=NITER)
    o = odp.Field(mask, maps, template, 2, NSIDE, LMAX, NITER, task=(False, True))
    guess = rng.uniform(size=(4, LMAX + 1))
    got = np.asarray(nmt.deprojection_bias(f, f, guess))
    ref = odp.deprojection_bias(o, o, guess)
    print(f"pure deprojection bias rel={_rel(got, ref):.2e}", flush=True)
    assert _rel(got, ref) < 1e-9
