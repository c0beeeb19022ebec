"""GPU parity of the pure-E/B coupling blocks (FP64) against the oracle.

Gauss-Legendre nodes (the defining moment property at order 2049 and the oracle's rule at 129),
Wigner-d tables and mask correlations, the eight-block quadrature against ``oracle.pure``,
pure workspaces (ncls 2, 4, 7, every purity pattern) against the oracle MCM, the standard
blocks by the same quadrature against the Wigner-3j kernel at lmax 3071 and 12287 (the rule
and the d-tables at scale), the vanishing purified M- blocks there, the expected
pseudo-spectra of fields purified on the device against the device blocks, and a purified-B
MASTER run on a pure-E sky at Nside 256.
"""

import numpy as np
import pytest
import torch

from oracle import pure as opure
from oracle.geometry import AlmInfo

pytestmark = pytest.mark.gpu


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


def _window(lmax, seed=1):
    rng = np.random.default_rng(seed)
    ell = np.arange(2 * lmax + 1)
    return rng.uniform(0.5, 1, 2 * lmax + 1) * 4 * np.pi / (1 + ell / 3.0) ** 3


def test_gauss_legendre_rule():
    from gmaster_amd import purecoupling as pc
    x, w = pc.gauss_legendre(129, "cuda")
    xr, wr = opure.gauss_legendre(129)
    assert np.max(np.abs(x.cpu().numpy() - xr)) < 1e-15
    assert _rel(w.cpu().numpy(), wr) < 1e-13
    # order 2049: ascending, and sum_k w_k P_l(x_k) = 2 delta_l0 for l < 2 order
    order = 2049
    x, w = pc.gauss_legendre(order, "cuda")
    x, w = x.cpu().numpy(), w.cpu().numpy()
    assert np.all(np.diff(x) > 0)
    p0, p1 = np.ones_like(x), x.copy()
    worst = abs(np.sum(w * p1))
    for l in range(2, 2 * order):
        p0, p1 = p1, ((2 * l - 1) * x * p1 - (l - 1) * p0) / l
        worst = max(worst, abs(np.sum(w * p1)))
    print(f"GL order {order}: max |sum w P_l| = {worst:.2e}", flush=True)
    assert abs(np.sum(w) - 2) < 1e-13
    assert worst < 1e-12


@pytest.mark.parametrize("m,n", [(0, 0), (0, 1), (0, 2), (1, 0), (1, 1), (1, 2), (2, 2), (0, -2), (-2, -2)])
def test_wigner_tables_and_correlations(m, n):
    from gmaster_amd import purecoupling as pc
    rng = np.random.default_rng(3)
    xs = np.sort(rng.uniform(-1, 1, 300))
    xs[0], xs[-1] = -0.9999999, 0.9999999
    x = torch.as_tensor(xs, device="cuda")
    lmax = 300
    scale = torch.as_tensor(rng.uniform(0.5, 2, lmax + 1), device="cuda")
    got = pc.wigner_d_table(x, m, n, lmax, scale).cpu().numpy()
    ref = opure.wigner_d_table(xs, m, n, lmax) * scale.cpu().numpy()[:, None]
    assert np.max(np.abs(got - ref)) < 1e-12
    coef = torch.as_tensor(rng.normal(size=lmax + 1), device="cuda")
    c, s = pc.wigner_d_corr(x, m, n, lmax, coef)
    cr = coef.cpu().numpy() @ opure.wigner_d_table(xs, m, n, lmax)
    sr = (coef.cpu().numpy() * (-1.0) ** np.arange(lmax + 1)) @ opure.wigner_d_table(xs, m, n, lmax)
    assert _rel(c.cpu().numpy(), cr) < 1e-12
    assert _rel(s.cpu().numpy(), sr) < 1e-12


@pytest.mark.parametrize("lmax", [24, 95])
def test_pure_blocks_match_oracle(lmax):
    from gmaster_amd.purecoupling import pure_coupling_blocks
    W = _window(lmax)
    got = pure_coupling_blocks(torch.as_tensor(W, device="cuda"), lmax)
    ref = opure.pure_blocks(W, lmax)
    for name, i in (("02pure", 3), ("22p1", 4), ("22p2", 6)):
        err = _rel(got[name].cpu().numpy(), ref[i])
        print(f"lmax {lmax} {name}: rel {err:.2e}", flush=True)
        assert err < 1e-11
    scale = np.abs(ref[6]).max()
    for name in ("22m1", "22m2"):
        assert np.abs(got[name].cpu().numpy()).max() < 1e-11 * scale


def _cap_mask(nside):
    from oracle.geometry import Rings
    R = Rings(nside)
    theta = np.repeat(R.theta, R.nphi)
    x = np.clip((np.radians(60.0) - theta) / np.radians(25.0), 0.0, 1.0)
    return 0.5 * (1 - np.cos(np.pi * x))


@pytest.mark.parametrize("case", ["ee_eb", "b_b", "e_b", "std_b", "t_b", "teb_e", "teb_b"])
def test_pure_workspace_matches_oracle(case):
    import gmaster_amd as nmt
    nside, lmax = 16, 32
    npix = 12 * nside * nside
    rng = np.random.default_rng(7)
    mask = _cap_mask(nside)
    qu = rng.normal(size=(2, npix))
    t = rng.normal(size=(1, npix))
    mk = lambda e, b: nmt.NmtField(mask, qu, purify_e=e, purify_b=b, lmax=lmax, lmax_mask=lmax, n_iter=1)
    ft = nmt.NmtField(mask, t, lmax=lmax, lmax_mask=lmax, n_iter=1)
    f1, f2, teb = {
        "ee_eb": (mk(True, False), mk(True, True), False),
        "b_b": (mk(False, True), mk(False, True), False),
        "e_b": (mk(True, False), mk(False, True), False),
        "std_b": (mk(False, False), mk(False, True), False),
        "t_b": (ft, mk(False, True), False),
        "teb_e": (ft, mk(True, False), True),
        "teb_b": (ft, mk(False, True), True),
    }[case]
    b = nmt.NmtBin.from_lmax_linear(lmax, 4)
    w = nmt.NmtWorkspace.from_fields(f1, f2, b, is_teb=teb)
    pcl = w.pcl_mask.cpu().numpy()
    blocks = opure.pure_blocks(pcl, lmax)
    from oracle.coupling import coupling_blocks_exact
    purity = (f1.pure_e, f1.pure_b, f2.pure_e, f2.pure_b)
    ncls = w.ncls
    lay = opure.layout_pure(ncls, f1.spin, f2.spin, *[int(v) for v in purity])
    m00 = coupling_blocks_exact(pcl, lmax)[0] if ncls == 7 else None
    ref = opure.assemble(blocks, lay, ncls, m00)
    got = w.get_coupling_matrix().cpu().numpy()
    err = _rel(got, ref)
    print(f"{case}: ncls {ncls} purity {purity} rel {err:.2e}", flush=True)
    assert err < 1e-11
    # FITS round trip keeps the purity flags and the blocks
    import os
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        fn = os.path.join(d, "w.fits")
        w.write_to(fn)
        w2 = nmt.NmtWorkspace.from_file(fn)
    assert w2.purity == w.purity
    assert _rel(w2.get_coupling_matrix().cpu().numpy(), got) < 1e-15
    cl = rng.uniform(size=(ncls, lmax + 1))
    assert _rel(w2.decouple_cell(cl).cpu().numpy(), w.decouple_cell(cl).cpu().numpy()) < 1e-12


def test_master_coefficients_with_purification():
    import gmaster_amd as nmt
    lmax = 40
    W = _window(lmax, 4)
    out = nmt.get_master_coefficients(W, lmax, 2, 2, pure_any=True)
    ref = opure.pure_blocks(W, lmax)
    cols = 2 * np.arange(lmax + 1) + 1.0
    assert out["pure_any"] and tuple(out["pp"].shape) == (3, lmax + 1, lmax + 1)
    for lvl, i in enumerate((1, 4, 6)):
        assert _rel(out["pp"][lvl].cpu().numpy(), ref[i] / cols) < 1e-11
    assert _rel(out["mm"][0].cpu().numpy(), ref[2] / cols) < 1e-11
    out = nmt.get_master_coefficients(W, lmax, 0, 2, pure_any=True)
    assert tuple(out["0s"].shape) == (2, lmax + 1, lmax + 1)
    assert _rel(out["0s"][1].cpu().numpy(), ref[3] / cols) < 1e-11


@pytest.mark.parametrize("lmax", [3071, 12287])
def test_quadrature_standard_blocks_match_3j_kernel_at_scale(lmax):
    """The rule and the d-tables at scale (up to the Nside-4096 lmax): the standard TE / M+ /
    M- blocks by the same quadrature pipeline equal the Wigner-3j kernel's, and the purified
    M- blocks vanish."""
    from gmaster_amd import purecoupling as pc
    from gmaster_amd.workspaces import clear_coupling_cache, coupling_blocks
    clear_coupling_cache()
    pc.clear_cache()
    W = torch.as_tensor(_window(lmax, 9), device="cuda")
    L, N = lmax + 1, 2 * lmax + 1
    x, wq = pc.gauss_legendre(N, "cuda")
    ell = torch.arange(N, dtype=torch.float64, device="cuda")
    coef = (2 * ell + 1) * W / (4 * np.pi)
    c00, s00 = pc.wigner_d_corr(x, 0, 0, 2 * lmax, coef)
    half = (2 * torch.arange(L, dtype=torch.float64, device="cuda") + 1) / 2
    d02 = pc.wigner_d_table(x, 0, 2, lmax)
    d22 = pc.wigner_d_table(x, 2, 2, lmax)
    te = pc.table_combine([(1, d02, c00, None)], wq) @ pc.wigner_d_table(x, 0, -2, lmax, half).T
    rS = pc.wigner_d_table(x, -2, -2, lmax, half)
    tot = pc.table_combine([(1, d22, c00, None)], wq) @ rS.T
    sg = pc.table_combine([(1, d22, s00, None)], wq) @ rS.T
    even, odd = pc.parity_split(tot, sg)
    ref = coupling_blocks(W, lmax)
    for name, got in (("02", te), ("22p", even), ("22m", odd)):
        err = (got - ref[name]).abs().max().item() / ref[name].abs().max().item()
        print(f"quadrature vs 3j {name} at lmax {lmax}: rel {err:.2e}", flush=True)
        assert err < 1e-10
    del d02, d22, rS, tot, sg
    del ref, te, even, odd
    clear_coupling_cache()
    blocks = pc.pure_coupling_blocks(W, lmax)
    scale = blocks["22p2"].abs().max().item()
    # the two-field M- cancels terms weighted by up to ((L+2)(L+1)L(L-1))^2 / (l^4) between
    # mask and row multipoles, so its rounding floor grows ~lmax^2 (8.7e-11 at 3071, 2.1e-9
    # at 12287 measured); the one-field M- vanishes to ~1e-20
    bound = {"22m1": 1e-15, "22m2": 4e-10 * (lmax / 3071) ** 2}
    for name in ("22m1", "22m2"):
        r = blocks[name].abs().max().item() / scale
        print(f"{name} / max|22p2| at lmax {lmax}: {r:.2e}", flush=True)
        assert r < bound[name]
    del blocks
    pc.clear_cache()
    torch.cuda.empty_cache()


def test_device_purified_pseudo_spectra_follow_device_blocks():
    """Expected pseudo-spectra of fields purified on the device (every unit-power mode at a
    few multipoles of a band-limited sky through a band-limited mask, Nside 32, transforms at
    lmax 48) against M[:, L] from the device's pure workspace blocks."""
    from gmaster_amd import sht
    from gmaster_amd.purecoupling import pure_coupling_blocks
    from gmaster_amd.purify import purify
    from gmaster_amd.workspaces import _coupled_cell, _layout, coupling_blocks
    nside, lf, lm, la, n_iter = 32, 20, 8, 48, 4
    ai = AlmInfo(la)
    rng = np.random.default_rng(11)
    am = np.zeros(ai.nelem, complex)
    for l in range(1, lm + 1):
        for m in range(l + 1):
            am[ai.idx(l, m)] = 0.6 / (1 + l) * (rng.normal() + (1j * rng.normal() if m else 0))
    am[ai.idx(0, 0)] = 2.5
    amask = torch.as_tensor(am, device="cuda")
    mask = sht.alm2map(amask[None], 0, nside, la, precision="fp64")[0]
    W = _coupled_cell(amask[None], amask[None], la)[0][: 2 * lf + 1]
    blocks = dict(coupling_blocks(W, lf))
    blocks.update(pure_coupling_blocks(W, lf))
    Ls = (2, 7, 16)
    resp = {}
    for kind in ("E", "B"):
        for L in Ls:
            acc = {}
            modes = [(0, 1.0)] + [(m, ph / np.sqrt(2)) for m in range(1, L + 1) for ph in (1, 1j)]
            for m, val in modes:
                alm = torch.zeros((2, ai.nelem), dtype=torch.complex128, device="cuda")
                alm[0 if kind == "E" else 1, ai.idx(L, m)] = val
                maps = sht.alm2map(alm, 2, nside, la, precision="fp64")
                pu = purify(mask, amask, maps, la, (True, True), n_iter, precision="fp64", return_maps=False)
                st = sht.map2alm(maps * mask[None], 2, la, n_iter=n_iter, precision="fp64")
                comp = {"sE": st[0], "sB": st[1], "pE": pu[0], "pB": pu[1]}
                for a in comp:
                    for b in comp:
                        acc[(a, b)] = acc.get((a, b), 0) + _coupled_cell(comp[a][None], comp[b][None], la)[0][: lf + 1]
            resp[(kind, L)] = {k: v.cpu().numpy() for k, v in acc.items()}
    tol = 1e-6 * blocks["22p"].abs().max().item()
    for purity in ((0, 0, 0, 0), (0, 1, 1, 1), (1, 0, 0, 1), (1, 1, 1, 1), (0, 1, 0, 1)):
        L1 = lf + 1
        M = np.zeros((L1, 4, L1, 4))
        for name, (c1, c2), sign in _layout(4, 2, 2, purity):
            M[:, c1, :, c2] += sign * blocks[name].cpu().numpy()
        n1 = ("pE" if purity[0] else "sE", "pB" if purity[1] else "sB")
        n2 = ("pE" if purity[2] else "sE", "pB" if purity[3] else "sB")
        worst = 0.0
        for L in Ls:
            for kind, col in (("E", 0), ("B", 3)):
                for c in range(4):
                    got = resp[(kind, L)][(n1[c // 2], n2[c % 2])][2:]
                    worst = max(worst, np.abs(got - M[2:, c, L, col]).max())
        print(f"purity {purity}: max |response - M| = {worst:.2e} (tol {tol:.1e})", flush=True)
        assert worst < tol


def test_purified_b_master_suppresses_leakage_scatter():
    """End to end at Nside 256: a pure-E sky (no B) through an apodized cut; the decoupled BB
    bandpowers of B-purified fields scatter far less around zero than the standard ones,
    and EE is recovered by both."""
    import gmaster_amd as nmt
    nside = 256
    lmax = 3 * nside - 1
    npix = 12 * nside * nside
    from oracle.geometry import pix2vec_ring
    vec = pix2vec_ring(nside, np.arange(npix))
    binary = (vec[:, 2] < np.cos(np.radians(50.0))).astype(np.float64)  # a cap of 50 deg removed
    mask = np.asarray(nmt.mask_apodization(binary, 10.0, "C2"))
    ell = np.arange(lmax + 1)
    ee = np.where(ell >= 2, 1.0 / (ell + 10.0) ** 2, 0.0)
    z = np.zeros(lmax + 1)
    qu = np.asarray(nmt.synfast_spherical(nside, np.array([ee, z, z]), [2], seed=5))
    b = nmt.NmtBin.from_nside_linear(nside, 32)
    out = {}
    for pure in (False, True):
        f = nmt.NmtField(mask, qu, purify_b=pure, n_iter=3)
        w = nmt.NmtWorkspace.from_fields(f, f, b)
        out[pure] = np.asarray(w.decouple_cell(nmt.compute_coupled_cell(f, f)))
    ee_b = np.asarray(b.bin_cell(ee[None]))[0]
    sel = slice(2, b.get_n_bands() - 2)
    bb_std = np.sqrt(np.mean(out[False][3, sel] ** 2))
    bb_pure = np.sqrt(np.mean(out[True][3, sel] ** 2))
    print(f"BB rms on a pure-E sky: standard {bb_std:.2e}, purified {bb_pure:.2e}; "
          f"EE/input std {np.mean(out[False][0, sel] / ee_b[sel]):.3f} pure {np.mean(out[True][0, sel] / ee_b[sel]):.3f}",
          flush=True)
    assert bb_pure < 0.1 * bb_std  # measured 0.05 (a 20x smaller scatter)
    for pure in (False, True):
        assert abs(np.mean(out[pure][0, sel] / ee_b[sel]) - 1) < 0.1
