"""GPU parity of the MASTER path (coupling kernel, coupled spectra, workspaces, decoupled
bandpowers) against the oracle and its committed fixtures.

The metric's bar for the plain FP64 path is 1e-6 relative on bandpowers; FP64 end to end
gives ~1e-12 here, and the tests hold 1e-9.
"""

import os

import numpy as np
import pytest
import torch

from oracle import master as omaster
from oracle.coupling import coupling_blocks as oracle_blocks

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _rel(got, ref):
    got = np.asarray(got)
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300))


def test_coupling_blocks_match_exact_3j_fixture():
    from gmaster_amd.workspaces import clear_coupling_cache, coupling_blocks

    clear_coupling_cache()
    fx = np.load(os.path.join(GOLDEN, "coupling_exact_lmax12.npz"))
    b = coupling_blocks(torch.as_tensor(fx["window"], device="cuda"), 12)
    for k, name in (("m00", "00"), ("m02", "02"), ("m22p", "22p"), ("m22m", "22m")):
        err = _rel(b[name].cpu(), fx[k])
        print(f"coupling {name} vs exact 3j rel={err:.2e}", flush=True)
        assert err < 1e-12


@pytest.mark.parametrize("lmax,lw", [(95, 95), (95, 190), (200, 200)])
def test_coupling_blocks_match_oracle_quadrature(lmax, lw):
    from gmaster_amd.workspaces import clear_coupling_cache, coupling_blocks

    clear_coupling_cache()
    rng = np.random.default_rng(lmax + lw)
    window = rng.uniform(size=lw + 1) / (1.0 + np.arange(lw + 1)) ** 2
    ref = oracle_blocks(window, lmax)
    b = coupling_blocks(torch.as_tensor(window, device="cuda"), lmax)
    for r, name in zip(ref, ("00", "02", "22p", "22m")):
        err = _rel(b[name].cpu(), r)
        print(f"coupling {name} lmax={lmax} lw={lw} rel={err:.2e}", flush=True)
        # the quadrature oracle itself carries ~1e-11 round-off (relative to max |M|) for
        # windows reaching 2 lmax; exactness is pinned by the Racah fixture test above
        assert err < 5e-11


def test_coupled_cell_matches_oracle():
    from gmaster_amd.workspaces import _coupled_cell

    lmax = 300
    rng = np.random.default_rng(5)
    n = (lmax + 1) * (lmax + 2) // 2
    a = rng.normal(size=(2, n)) + 1j * rng.normal(size=(2, n))
    b = rng.normal(size=(2, n)) + 1j * rng.normal(size=(2, n))
    ref = omaster.coupled_cell(a, b, lmax)
    got = _coupled_cell(torch.as_tensor(a, device="cuda"), torch.as_tensor(b, device="cuda"), lmax).cpu().numpy()
    assert _rel(got, ref) < 1e-13


def _pipeline(fx, nside, nlb):
    import gmaster_amd as nmt

    mask = torch.as_tensor(fx["mask"], device="cuda")
    tqu = torch.as_tensor(fx["tqu"], device="cuda")
    fT = nmt.NmtField(mask, tqu[0:1], n_iter=3)
    fP = nmt.NmtField(mask, tqu[1:3], n_iter=3)
    b = nmt.NmtBin.from_nside_linear(nside, nlb)
    out = {}
    for name, (f1, f2) in {"00": (fT, fT), "02": (fT, fP), "22": (fP, fP)}.items():
        w = nmt.NmtWorkspace.from_fields(f1, f2, b)
        cl = nmt.compute_coupled_cell(f1, f2)
        out["cl_" + name] = np.asarray(cl)
        out["mcm_binned_" + name] = np.asarray(w.mcm_binned.cpu())
        out["bp_" + name] = np.asarray(w.decouple_cell(cl))
    return out


@pytest.mark.parametrize("nside,nlb,fname", [(16, 8, "pipeline_nside16.npz"), (64, 16, "pipeline_nside64.npz")])
def test_tp_pipeline_matches_oracle_fixture(nside, nlb, fname):
    from gmaster_amd.workspaces import clear_coupling_cache

    clear_coupling_cache()
    fx = np.load(os.path.join(GOLDEN, fname))
    got = _pipeline(fx, nside, nlb)
    for key in sorted(got):
        err = _rel(got[key], fx[key])
        print(f"nside={nside} {key} rel={err:.2e}", flush=True)
        assert err < 1e-9, key


def test_full_sky_window_gives_identity_coupling():
    """W = 4 pi delta_l0 (full sky): M00 = I, M02 = M22+ = I for l >= 2, M22- = 0 -- an
    exact invariant independent of the oracle."""
    from gmaster_amd.workspaces import clear_coupling_cache, coupling_blocks

    clear_coupling_cache()
    lmax = 150
    w = np.zeros(lmax + 1)
    w[0] = 4 * np.pi
    b = coupling_blocks(torch.as_tensor(w, device="cuda"), lmax)
    eye = np.eye(lmax + 1)
    eye2 = eye.copy()
    eye2[:2, :2] = 0
    assert np.max(np.abs(b["00"].cpu().numpy() - eye)) < 1e-13
    assert np.max(np.abs(b["02"].cpu().numpy() - eye2)) < 1e-13
    assert np.max(np.abs(b["22p"].cpu().numpy() - eye2)) < 1e-13
    assert np.max(np.abs(b["22m"].cpu().numpy())) < 1e-13


def test_full_sky_field_decouples_close_to_binned_pseudo_cl():
    """Mask = 1 end to end: the mask pseudo-C_ell is 4 pi delta_l0 up to the HEALPix
    quadrature residual after 3 Jacobi steps, so decoupled ~ binned (loose bound)."""
    import gmaster_amd as nmt

    nside = 32
    rng = np.random.default_rng(1)
    npix = 12 * nside * nside
    mask = torch.ones(npix, dtype=torch.float64, device="cuda")
    f = nmt.NmtField(mask, torch.as_tensor(rng.normal(size=(2, npix)), device="cuda"), n_iter=3)
    b = nmt.NmtBin.from_nside_linear(nside, 8)
    w = nmt.NmtWorkspace.from_fields(f, f, b)
    cl = nmt.compute_coupled_cell(f, f)
    bp = np.asarray(w.decouple_cell(cl))
    ref = np.asarray(b.bin_cell(cl))
    assert _rel(bp, ref) < 1e-4


def test_results_convert_with_numpy_asarray():
    import gmaster_amd as nmt

    nside = 8
    m = torch.ones(12 * nside * nside, dtype=torch.float64, device="cuda")
    f = nmt.NmtField(m, m[None] * 2.0)
    a = np.asarray(f.get_alms())
    assert a.shape == (1, (3 * nside) * (3 * nside + 1) // 2) and a.dtype == np.complex128


TOEPLITZ = [(6, 3, 2), (16, 4, 3), (60, 60, 0), (90, 1, 5)]


def test_coupling_pairs_match_full_kernel():
    from gmaster_amd import _lib
    from gmaster_amd.workspaces import _g_table, _toeplitz_pairs, clear_coupling_cache, coupling_blocks

    clear_coupling_cache()
    lmax = 95
    L = lmax + 1
    window = np.random.default_rng(19).uniform(size=2 * lmax + 1)
    w = torch.as_tensor(window, device="cuda")
    full = coupling_blocks(w, lmax)
    wp = (w * (2 * torch.arange(len(w), dtype=torch.float64, device="cuda") + 1) / (4 * np.pi)).contiguous()
    lo, hi = _toeplitz_pairs(L, 16, 4, 3, "cuda")
    out = {k: torch.full((L, L), float("nan"), dtype=torch.float64, device="cuda") for k in ("00", "02", "22p", "22m")}
    _lib.call("gm_coupling_pairs", _lib.ptr(wp), len(w) - 1, _lib.ptr(_g_table(L + 1, "cuda")), lmax, _lib.ptr(lo),
              _lib.ptr(hi), lo.numel(), *(_lib.ptr(out[k]) for k in ("00", "02", "22p", "22m")), _lib.stream_ptr("cuda"))
    torch.cuda.synchronize()
    sel = torch.zeros((L, L), dtype=torch.bool, device="cuda")
    sel[lo.long(), hi.long()] = True
    sel[hi.long(), lo.long()] = True
    for k in out:
        assert torch.equal(out[k][sel], full[k][sel]), k
        assert bool(torch.isnan(out[k][~sel]).all()), k


@pytest.mark.parametrize("lmax,lw", [(300, 600), (300, 300), (700, 700)])
def test_coupling_rows_match_full_kernel(lmax, lw):
    """gm_coupling_rows over the Toeplitz fill's row items (the workspace path) against the
    full kernel: bit for bit where the window reaches 2 lmax (every lane of a wave runs the
    same degrees either way), to 1e-14 of the block where it stops at lmax (lanes cut short
    take a few more zero-weight steps than in the full kernel's waves); untouched elsewhere."""
    from gmaster_amd import _lib
    from gmaster_amd.workspaces import _g_table, _toeplitz_items, clear_coupling_cache, coupling_blocks

    clear_coupling_cache()
    L = lmax + 1
    window = np.random.default_rng(lmax + lw).uniform(size=lw + 1) / (1.0 + np.arange(lw + 1))
    w = torch.as_tensor(window, device="cuda")
    full = coupling_blocks(w, lmax)
    wp = (w * (2 * torch.arange(len(w), dtype=torch.float64, device="cuda") + 1) / (4 * np.pi)).contiguous()
    lt, le, dl = lmax // 2, lmax // 10, lmax // 8
    items, _, _ = _toeplitz_items(L, lt, le, dl, "cuda")
    print(f"coupling rows lmax={lmax} lw={lw}: {items.shape[0]} items", flush=True)
    out = {k: torch.full((L, L), float("nan"), dtype=torch.float64, device="cuda") for k in ("00", "02", "22p", "22m")}
    _lib.call("gm_coupling_rows", _lib.ptr(wp), len(w) - 1, _lib.ptr(_g_table(L + 1, "cuda")), lmax, _lib.ptr(items),
              items.shape[0], *(_lib.ptr(out[k]) for k in ("00", "02", "22p", "22m")), _lib.stream_ptr("cuda"))
    torch.cuda.synchronize()
    it = items.long()
    sel = torch.zeros((L, L), dtype=torch.bool, device="cuda")
    for k in range(64):
        live = it[:, 2] > k
        sel[it[live, 0], it[live, 1] + k] = True
    sel = sel | sel.T
    for k in out:
        if lw >= 2 * lmax:
            assert torch.equal(out[k][sel], full[k][sel]), k
        else:
            err = float((out[k][sel] - full[k][sel]).abs().max() / full[k].abs().max())
            assert err < 1e-14, (k, err)
        assert bool(torch.isnan(out[k][~sel]).all()), k


@pytest.mark.parametrize("lt,le,dl", TOEPLITZ)
def test_toeplitz_kernel_matches_oracle(lt, le, dl):
    from gmaster_amd import _lib
    from oracle.toeplitz import apply_to_mcm_block

    lmax = 95
    window = np.random.default_rng(lt).uniform(size=2 * lmax + 1) / (1.0 + np.arange(2 * lmax + 1))
    for M in oracle_blocks(window, lmax):
        Md = torch.as_tensor(M, device="cuda")
        out = torch.empty_like(Md)
        _lib.call("gm_toeplitz", _lib.ptr(Md), lmax, lt, le, dl, _lib.ptr(out), _lib.stream_ptr("cuda"))
        ref = apply_to_mcm_block(M, lt, le, dl)
        got = out.cpu().numpy()
        # spin-2 blocks vanish below l = 2, so l_exact < 2 divides by a zero diagonal: the
        # reference yields inf / nan there too, and those entries must agree in kind
        fin = np.isfinite(ref)
        np.testing.assert_array_equal(np.isfinite(got), fin)
        np.testing.assert_array_equal(np.isnan(got), np.isnan(ref))
        err = _rel(got[fin], ref[fin])
        print(f"toeplitz lt={lt} le={le} dl={dl} rel={err:.2e} non-finite={int((~fin).sum())}", flush=True)
        assert err < 1e-14


@pytest.mark.parametrize("lmax,lt,le,dl", [(600, 300, 40, 20), (777, 500, 120, 64), (1000, 513, 513, 0)])
def test_toeplitz_kernel_tiles_match_oracle(lmax, lt, le, dl):
    """Blocks wider than one 32 x 256 tile of the fill kernel (tiles straddling the diagonal,
    the band, the exact columns and the corner extrapolation, ragged last tiles), on a
    random matrix with a positive diagonal, against the oracle's fill."""
    from gmaster_amd import _lib
    from oracle.toeplitz import apply_to_mcm_block

    rng = np.random.default_rng(lmax)
    M = rng.uniform(-1.0, 1.0, size=(lmax + 1, lmax + 1)) * 1e-3
    M[np.diag_indices(lmax + 1)] = rng.uniform(0.5, 2.0, size=lmax + 1)
    Md = torch.as_tensor(M, device="cuda")
    out = torch.empty_like(Md)
    print(f"toeplitz tiles lmax={lmax} lt={lt} le={le} dl={dl}", flush=True)
    _lib.call("gm_toeplitz", _lib.ptr(Md), lmax, lt, le, dl, _lib.ptr(out), _lib.stream_ptr("cuda"))
    ref = apply_to_mcm_block(M, lt, le, dl)
    err = _rel(out.cpu().numpy(), ref)
    print(f"  rel={err:.2e}", flush=True)
    assert err < 1e-14


@pytest.mark.parametrize("lt,le,dl", TOEPLITZ[:2])
def test_master_coefficients_with_toeplitz_match_oracle(lt, le, dl):
    """The reference's own check (tests/test_workspaces.py:89-105) at lmax 12, held against
    the oracle's exact blocks under the oracle's fill instead of NaMaster."""
    import gmaster as nmt
    from gmaster_amd.workspaces import clear_coupling_cache
    from oracle.toeplitz import apply_toeplitz

    clear_coupling_cache()
    lmax = 12 if lt < 12 else 31
    mask_cell = np.random.default_rng(18).uniform(size=lmax + 1)
    padded = np.zeros(2 * lmax + 1)
    padded[: lmax + 1] = mask_cell
    cols = 2 * np.arange(lmax + 1) + 1
    ref = dict(zip(("00", "0s", "pp", "mm"), (b / cols for b in oracle_blocks(padded, lmax))))
    for s1, s2, names in ((0, 0, ("00",)), (0, 2, ("0s",)), (2, 2, ("pp", "mm"))):
        got = nmt.get_master_coefficients(mask_cell, lmax, s1, s2, l_toeplitz=lt, l_exact=le, dl_band=dl)
        for name in names:
            g = np.asarray(got[name])
            g = g if name == "00" else g[0]
            want = apply_toeplitz(ref[name], lt, le, dl)
            # the quadrature oracle carries ~1e-12 relative round-off (see the coupling tests)
            np.testing.assert_allclose(g, want, rtol=0, atol=5e-11 * np.abs(want).max())


def test_workspace_toeplitz_blocks_and_bandpowers():
    import gmaster as nmt
    from gmaster_amd.workspaces import clear_coupling_cache
    from oracle.toeplitz import apply_to_mcm_block

    clear_coupling_cache()
    fx = np.load(os.path.join(GOLDEN, "pipeline_nside16.npz"))
    nside = 16
    mask = fx["mask"]
    f0 = nmt.NmtField(mask, [fx["tqu"][0]], n_iter=3)
    f2 = nmt.NmtField(mask, [fx["tqu"][1], fx["tqu"][2]], n_iter=3)
    b = nmt.NmtBin.from_nside_linear(nside, 4)
    exact = nmt.NmtWorkspace.from_fields(f0, f2, b)
    w = nmt.NmtWorkspace.from_fields(f0, f2, b, l_toeplitz=30, l_exact=10, dl_band=8)
    for name in ("00", "02", "22p", "22m"):
        ref = apply_to_mcm_block(exact.blocks[name].cpu().numpy(), 30, 10, 8)
        err = _rel(w.blocks[name].cpu(), ref)
        print(f"workspace toeplitz block {name} rel={err:.2e}", flush=True)
        assert err < 1e-13
    cl = nmt.compute_coupled_cell(f0, f2)
    bp_t = np.asarray(w.decouple_cell(cl))
    bp_e = np.asarray(exact.decouple_cell(cl))
    # the approximation moves the bandpowers (its size depends on the mask; printed only)
    rel = np.abs(bp_t - bp_e).max() / np.abs(bp_e).max()
    print(f"toeplitz vs exact bandpowers rel={rel:.2e}", flush=True)
    assert np.all(np.isfinite(bp_t)) and rel > 0
    with pytest.raises(ValueError, match="smaller than"):
        nmt.NmtWorkspace.from_fields(f0, f2, b, l_toeplitz=47, l_exact=10, dl_band=8)


@pytest.mark.parametrize("spins", [(0, 0), (0, 2), (2, 2)])
def test_workspace_fits_round_trip(tmp_path, spins):
    """write_to / from_file keep the coupling matrix bit for bit and the decoupled
    bandpowers (the reference's own check, tests/test_workspaces.py:684-716, without
    pymaster)."""
    import gmaster as nmt

    fx = np.load(os.path.join(GOLDEN, "pipeline_nside16.npz"))
    mask = torch.as_tensor(fx["mask"], device="cuda")
    tqu = torch.as_tensor(fx["tqu"], device="cuda")
    fields = {0: nmt.NmtField(mask, tqu[0:1], n_iter=3), 2: nmt.NmtField(mask, tqu[1:3], n_iter=3)}
    f1, f2 = fields[spins[0]], fields[spins[1]]
    b = nmt.NmtBin.from_nside_linear(16, 4)
    w = nmt.NmtWorkspace.from_fields(f1, f2, b)
    w.update_beams(np.linspace(1, 0.5, 48), np.linspace(1, 0.7, 48))
    path = str(tmp_path / "w.fits")
    w.write_to(path)
    r = nmt.NmtWorkspace.from_file(path)
    assert torch.equal(r.mcm.cpu(), w.mcm.cpu())
    assert (r.ncls, r.spin1, r.spin2, r.lmax) == (w.ncls, w.spin1, w.spin2, w.lmax)
    cl = nmt.compute_coupled_cell(f1, f2)
    np.testing.assert_allclose(np.asarray(r.decouple_cell(cl)), np.asarray(w.decouple_cell(cl)), rtol=1e-13,
                               atol=0)
    np.testing.assert_allclose(np.asarray(r.get_bandpower_windows()), np.asarray(w.get_bandpower_windows()),
                               rtol=1e-12, atol=1e-15)


def test_teb_workspace_matches_oracle_layout_and_separate_workspaces(tmp_path):
    """is_teb: the dense MCM is the oracle's joint layout of the same blocks, and with one
    mask the 7 decoupled spectra equal the 00, 02 and 22 workspaces' (a block-diagonal
    system), also after a FITS round trip."""
    import gmaster as nmt
    from oracle.coupling import assemble_mcm_teb

    fx = np.load(os.path.join(GOLDEN, "pipeline_nside16.npz"))
    mask = torch.as_tensor(fx["mask"], device="cuda")
    tqu = torch.as_tensor(fx["tqu"], device="cuda")
    f0 = nmt.NmtField(mask, tqu[0:1], n_iter=3)
    f2 = nmt.NmtField(mask, tqu[1:3], n_iter=3)
    b = nmt.NmtBin.from_nside_linear(16, 4)
    with pytest.raises(ValueError, match="is_teb"):
        nmt.NmtWorkspace.from_fields(f2, f0, b, is_teb=True)
    wt = nmt.NmtWorkspace.from_fields(f0, f2, b, is_teb=True)
    assert wt.ncls == 7
    blocks = [wt.blocks[k].cpu().numpy() for k in ("00", "02", "22p", "22m")]
    np.testing.assert_array_equal(np.asarray(wt.get_coupling_matrix().cpu()), assemble_mcm_teb(blocks))
    cl = torch.cat([nmt.compute_coupled_cell(f0, f0), nmt.compute_coupled_cell(f0, f2),
                    nmt.compute_coupled_cell(f2, f2)])
    bp = np.asarray(wt.decouple_cell(cl))
    sep = [np.asarray(nmt.NmtWorkspace.from_fields(a, c, b).decouple_cell(nmt.compute_coupled_cell(a, c)))
           for a, c in ((f0, f0), (f0, f2), (f2, f2))]
    np.testing.assert_allclose(bp, np.concatenate(sep), rtol=1e-10, atol=1e-15 * np.abs(bp).max())
    path = str(tmp_path / "teb.fits")
    wt.write_to(path)
    r = nmt.NmtWorkspace.from_file(path)
    assert r.is_teb and r.ncls == 7
    np.testing.assert_allclose(np.asarray(r.decouple_cell(cl)), bp, rtol=1e-13, atol=0)


@pytest.mark.parametrize("nside,nlb,fname", [(16, 4, "pipeline_nside16.npz"), (64, 16, "pipeline_nside64.npz")])
@pytest.mark.parametrize("spins", [(0, 0), (0, 2), (2, 2)])
def test_bandpower_windows_and_couple_cell_match_oracle(nside, nlb, fname, spins):
    """``get_bandpower_windows`` and ``couple_cell`` (with beams) against
    ``oracle.master.bandpower_windows`` / ``couple_cell`` built from the fixture's mask
    pseudo-C_ell (the oracle's windows are pinned in tests/test_oracle_master.py)."""
    import gmaster as nmt
    from oracle.coupling import assemble_mcm

    fx = np.load(os.path.join(GOLDEN, fname))
    mask = torch.as_tensor(fx["mask"], device="cuda")
    tqu = torch.as_tensor(fx["tqu"], device="cuda")
    fields = {0: nmt.NmtField(mask, tqu[0:1], n_iter=3), 2: nmt.NmtField(mask, tqu[1:3], n_iter=3)}
    lmax = 3 * nside - 1
    b = nmt.NmtBin.from_nside_linear(nside, nlb)
    w = nmt.NmtWorkspace.from_fields(fields[spins[0]], fields[spins[1]], b)
    beam1, beam2 = np.linspace(1, 0.5, lmax + 1), np.linspace(1, 0.8, lmax + 1)
    w.update_beams(beam1, beam2)
    mcm, ncls = assemble_mcm(oracle_blocks(fx["pcl_mask"], lmax), *spins)
    ells, bpws, wts = omaster.linear_bins(lmax, nlb)
    out_op, th_op = omaster.binning_operators(lmax, ells, bpws, wts)
    ref = omaster.bandpower_windows(mcm, ncls, lmax, out_op, th_op, beam1 * beam2)
    got = np.asarray(w.get_bandpower_windows())
    err = _rel(got, ref)
    print(f"nside={nside} spins={spins} windows rel={err:.2e}", flush=True)
    assert got.shape == ref.shape and err < 1e-9
    rng = np.random.default_rng(nside + spins[1])
    cl = rng.normal(size=(ncls, lmax + 3))
    cc = np.asarray(w.couple_cell(cl))
    err_c = _rel(cc, omaster.couple_cell(mcm, cl, ncls, lmax, beam1 * beam2))
    print(f"  couple_cell rel={err_c:.2e}", flush=True)
    assert err_c < 1e-9
    # the device windows reproduce the device decoupling of the coupled expectation
    bp = np.asarray(w.decouple_cell(w.couple_cell(cl)))
    assert _rel(np.einsum("cbdl,dl->cb", got, cl[:, :lmax + 1]), bp) < 1e-10
