"""GPU parity of contaminant deprojection against the oracle (FP64 form).

Fields with templates (spin 0 and spin 2, random masks), the deprojection and
uncorrelated-noise biases, ``compute_full_master``, and the three device primitives
(``gm_map_dots``, ``gm_map_deproject``, ``gm_alm_filter``) by their definitions.  The
oracle itself is pinned by closed forms in ``tests/test_oracle_deproject.py``; the GPU
closed-form case is repeated here so a wrong sign or factor fails on the device too.
"""

import numpy as np
import pytest
import torch

from oracle import deproject as odp
from oracle import master as omaster
from oracle import sht as osht
from oracle.geometry import AlmInfo

pytestmark = pytest.mark.gpu

NSIDE, LMAX, NITER = 16, 24, 3
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


def test_map_dots_deproject_and_filter_primitives():
    from gmaster_amd import deproject as dp
    rng = np.random.default_rng(0)
    a = rng.normal(size=(3, 2, NPIX))
    b = rng.normal(size=(2, 2, NPIX))
    v = rng.uniform(0.1, 1.0, NPIX)
    w = 4 * np.pi / NPIX
    got = dp.map_dots(torch.as_tensor(a, device="cuda"), torch.as_tensor(b, device="cuda"),
                      weight_map=torch.as_tensor(v, device="cuda")).cpu().numpy()
    ref = w * np.einsum("iap,jap,p->ij", a, b, v)
    print(f"map_dots rel={_rel(got, ref):.2e}", flush=True)
    assert _rel(got, ref) < 1e-13
    maps = torch.as_tensor(rng.normal(size=(2, NPIX)), device="cuda")
    m0 = maps.cpu().numpy().copy()
    temps = torch.as_tensor(a, device="cuda")
    alphas, iM = dp.deproject(maps, temps, 1e-10)
    clean, _, oal, oiM = odp.deproject(np.ones(NPIX), m0, a)
    assert _rel(alphas, oal) < 1e-12 and _rel(iM, oiM) < 1e-12
    assert _rel(maps.cpu(), clean) < 1e-12
    ainfo = AlmInfo(LMAX)
    alms = rng.normal(size=(2, ainfo.nelem)) + 1j * rng.normal(size=(2, ainfo.nelem))
    spec = rng.normal(size=(2, 3, LMAX + 1))
    out = dp.filter_alms(torch.as_tensor(alms, device="cuda"), torch.as_tensor(spec, device="cuda"), LMAX)
    assert _rel(out.cpu(), odp.filter_alms(alms, spec, LMAX)) < 1e-15


@pytest.mark.parametrize("spin", [0, 2])
def test_field_with_templates_matches_oracle(spin):
    import gmaster_amd as nmt
    rng = np.random.default_rng(10 + spin)
    nmaps = 1 if spin == 0 else 2
    mask = rng.uniform(0.3, 1.0, NPIX)
    temps = rng.normal(size=(2, nmaps, NPIX))
    maps = 2.5 * temps[0] - 0.7 * temps[1] + rng.normal(size=(nmaps, NPIX))
    f = nmt.NmtField(mask, maps, templates=temps, lmax=LMAX, lmax_mask=LMAX, n_iter=NITER)
    o = odp.Field(mask, maps, temps, spin, NSIDE, LMAX, NITER)
    print(f"spin {spin}: alphas {f.alphas} oracle {o.alphas}", flush=True)
    assert _rel(f.alphas, o.alphas) < 1e-12
    assert _rel(f.get_maps(), o.maps) < 1e-12
    assert _rel(f.get_templates(), o.temp) < 1e-15
    assert _rel(f.get_alms(), o.alm) < 1e-10
    assert _rel(f.alm_temp.cpu(), o.alm_temp) < 1e-10
    # pre-masked device maps: the subtraction works on a copy, the caller's tensor is untouched
    masked = torch.as_tensor(maps * mask[None, :], device="cuda")
    before = masked.clone()
    g = nmt.NmtField(mask, masked, templates=temps * mask[None, None, :], masked_on_input=True, lmax=LMAX,
                     lmax_mask=LMAX, n_iter=NITER)
    assert torch.equal(masked, before)
    assert _rel(g.get_maps(), o.maps) < 1e-12


def test_deprojection_biases_and_full_master_match_oracle():
    import gmaster_amd as nmt
    rng = np.random.default_rng(17)
    mask = rng.uniform(0.3, 1.0, NPIX)
    t1 = rng.normal(size=(1, 1, NPIX))
    t2 = rng.normal(size=(2, 1, NPIX))
    m1 = rng.normal(size=(1, NPIX))
    m2 = rng.normal(size=(1, NPIX))
    f1 = nmt.NmtField(mask, m1, templates=t1, lmax=LMAX, lmax_mask=LMAX, n_iter=NITER)
    f2 = nmt.NmtField(mask, m2, templates=t2, lmax=LMAX, lmax_mask=LMAX, n_iter=NITER)
    o1 = odp.Field(mask, m1, t1, 0, NSIDE, LMAX, NITER)
    o2 = odp.Field(mask, m2, t2, 0, NSIDE, LMAX, NITER)
    guess = rng.uniform(0.5, 2.0, size=(1, LMAX + 1))
    got = np.asarray(nmt.deprojection_bias(f1, f2, guess))
    ref = odp.deprojection_bias(o1, o2, guess)
    print(f"deprojection bias rel={_rel(got, ref):.2e}", flush=True)
    assert _rel(got, ref) < 1e-9
    var = rng.uniform(0.1, 1.0, NPIX)
    got = np.asarray(nmt.uncorr_noise_deprojection_bias(f2, var))
    ref = odp.uncorr_noise_deprojection_bias(o2, var)
    print(f"noise deprojection bias rel={_rel(got, ref):.2e}", flush=True)
    assert _rel(got, ref) < 1e-9
    b = nmt.NmtBin.from_lmax_linear(LMAX, 4)
    bp = np.asarray(nmt.compute_full_master(f1, f2, b, cl_guess=guess))
    # the same from the oracle's pieces: decouple(coupled - bias)
    w = nmt.NmtWorkspace.from_fields(f1, f2, b)
    coupled = omaster.coupled_cell(o1.alm, o2.alm, LMAX)
    ref_bp = np.asarray(w.decouple_cell(coupled - odp.deprojection_bias(o1, o2, guess)))
    print(f"full master rel={_rel(bp, ref_bp):.2e}", flush=True)
    assert _rel(bp, ref_bp) < 1e-9


def test_spin2_bias_of_one_removed_e_mode_on_gpu():
    import gmaster_amd as nmt
    l0 = 6
    ainfo = AlmInfo(LMAX)
    a = np.zeros((2, ainfo.nelem), dtype=np.complex128)
    a[0, ainfo.idx(l0, 0)] = 1.0
    t = osht.alm2map(a, 2, NSIDE, LMAX)
    rng = np.random.default_rng(4)
    f = nmt.NmtField(np.ones(NPIX), rng.normal(size=(2, NPIX)), templates=t[None], lmax=LMAX, lmax_mask=LMAX,
                     n_iter=NITER)
    cl = np.zeros((4, LMAX + 1))
    cl[0] = rng.uniform(0.5, 2.0, LMAX + 1)
    cl[3] = rng.uniform(0.5, 2.0, LMAX + 1)
    M = 4 * np.pi / NPIX * np.sum(t * t)
    expect = np.zeros((4, LMAX + 1))
    expect[0, l0] = -cl[0, l0] / (M * (2 * l0 + 1))
    got = np.asarray(nmt.deprojection_bias(f, f, cl))
    print(f"closed form EE bias {got[0, l0]:.12e} expect {expect[0, l0]:.12e}", flush=True)
    np.testing.assert_allclose(got, expect, atol=1e-9)


def test_fields_without_templates_have_zero_bias():
    import gmaster_amd as nmt
    rng = np.random.default_rng(3)
    mask = rng.uniform(0.3, 1.0, NPIX)
    f = nmt.NmtField(mask, rng.normal(size=(1, NPIX)), lmax=LMAX, lmax_mask=LMAX)
    assert np.all(np.asarray(nmt.deprojection_bias(f, f, np.ones((1, LMAX + 1)))) == 0)
    assert np.all(np.asarray(nmt.uncorr_noise_deprojection_bias(f, np.ones(NPIX))) == 0)
    with pytest.raises(ValueError):
        nmt.NmtField(mask, rng.normal(size=(1, NPIX)), templates=rng.normal(size=(1, 2, NPIX)))
