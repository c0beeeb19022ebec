"""GPU checks of the mixed-precision form: FP32 refinement passes against the FP64 path and
the oracle fixtures, and the fast coupling form."""

import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("nside,spin", [(16, 0), (64, 0), (64, 2), (128, 2)])
def test_mixed_kernels_match_fp64_kernels_to_fp32_precision(nside, spin):
    """One analysis and one synthesis pass with FP32 contraction against the FP64 kernels:
    the recurrence is FP64 in both, so only the FP32 products and sums differ."""
    from gmaster_amd import sht

    print(f"mixed kernels nside={nside} spin={spin}", flush=True)
    rng = np.random.default_rng(nside + spin)
    lmax = 3 * nside - 1
    L = lmax + 1
    nelem = L * (L + 1) // 2
    p = sht.plan(nside, lmax, spin, "cuda")
    nring = 4 * nside - 1
    if spin == 0:
        spec = torch.as_tensor(rng.standard_normal((2, nring, L)) + 1j * rng.standard_normal((2, nring, L)),
                               device="cuda")
        a64 = torch.empty((2, nelem), dtype=torch.complex128, device="cuda")
        a32 = torch.empty_like(a64)
        sht._analysis_s0(p, spec, a64, 0)
        sht._analysis_s0(p, spec, a32, 0, mixed=True)
        s64, s32 = torch.empty_like(spec), torch.empty_like(spec)
        sht._synthesis_s0(p, a64, s64)
        sht._synthesis_s0(p, a64, s32, mixed=True)
        # the single-map mixed entry (the package synthesises single maps in FP64)
        from gmaster_amd import _lib
        st = _lib.stream_ptr()
        coef = torch.empty(p.tsize, dtype=torch.complex128, device="cuda")
        _lib.call("gm_lat_prep_s0", _lib.ptr(a64[0:1]), 1, a64.stride(0), _lib.ptr(p.P), p.lmax, _lib.ptr(coef), st)
        one = torch.empty_like(spec[0:1])
        _lib.call("gm_lat_syn_s0_mixed", _lib.ptr(coef), 1, p.nside, p.lmax, _lib.ptr(p.alpha), _lib.ptr(p.smant),
                  _lib.ptr(p.sexp), _lib.ptr(p.nact), _lib.ptr(one), one.shape[2], one.stride(0), st)
        e1 = _rel(one.cpu(), s64[0:1].cpu())
        print(f"  single-map mixed synthesis rel {e1:.2e}", flush=True)
        assert e1 < 1e-5
    else:
        W = 2 * L - 1
        spec = torch.as_tensor(rng.standard_normal((nring, W)) + 1j * rng.standard_normal((nring, W)), device="cuda")
        a64 = torch.empty((2, nelem), dtype=torch.complex128, device="cuda")
        a32 = torch.empty_like(a64)
        sht._analysis_s(p, spec, a64[0], a64[1], 0)
        sht._analysis_s(p, spec, a32[0], a32[1], 0, mixed=True)
        s64, s32 = torch.empty_like(spec), torch.empty_like(spec)
        sht._synthesis_s(p, a64[0], a64[1], s64)
        sht._synthesis_s(p, a64[0], a64[1], s32, mixed=True)
    ea, es = _rel(a32.cpu(), a64.cpu()), _rel(s32.cpu(), s64.cpu())
    print(f"  analysis rel {ea:.2e} synthesis rel {es:.2e}", flush=True)
    assert ea < 1e-5 and es < 1e-5


@pytest.mark.parametrize("nside,spin", [(32, 0), (64, 0), (64, 2), (256, 2), (512, 0)])
def test_mixed_refinement_matches_fp64_refinement(nside, spin):
    from gmaster_amd import sht, workload

    print(f"mixed refinement nside={nside} spin={spin}", flush=True)
    lmax = 3 * nside - 1
    maps = workload.tqu_maps(nside, lmax, seed=nside)
    mask = workload.hole_mask(nside, device="cuda")
    m = maps[0:1] * mask if spin == 0 else maps[1:3] * mask
    a64 = sht.map2alm(m, spin, lmax, n_iter=3)
    a32 = sht.map2alm(m, spin, lmax, n_iter=3, precision="mixed")
    a0 = sht.map2alm(m, spin, lmax, n_iter=0)
    e, step = _rel(a32.cpu(), a64.cpu()), _rel(a0.cpu(), a64.cpu())
    print(f"  mixed vs fp64 rel {e:.2e} (refinement moves alm by {step:.2e})", flush=True)
    assert e < 1e-6 and e < 1e-4 * step


def test_fast_form_pipeline_matches_fixture():
    import gmaster_amd as nmt
    from gmaster_amd import field, workspaces

    fx = np.load(os.path.join(GOLDEN, "pipeline_nside64.npz"))
    workspaces.clear_coupling_cache()
    field.clear_mask_cache()
    nmt.set_form("fast")
    try:
        mask = torch.as_tensor(fx["mask"], device="cuda")
        tqu = torch.as_tensor(fx["tqu"], device="cuda")
        fT = nmt.NmtField(mask, tqu[0:1], n_iter=3)
        fP = nmt.NmtField(mask, tqu[1:3], n_iter=3)
        b = nmt.NmtBin.from_nside_linear(64, 16)
        for name, (f1, f2) in {"00": (fT, fT), "02": (fT, fP), "22": (fP, fP)}.items():
            w = nmt.NmtWorkspace.from_fields(f1, f2, b)
            bp = np.asarray(w.decouple_cell(nmt.compute_coupled_cell(f1, f2)))
            e = _rel(bp, fx["bp_" + name])
            print(f"fast form bp_{name} rel {e:.2e} coupling {w.coupling_info}", flush=True)
            assert e < 1e-6
    finally:
        nmt.set_form("fp64")
        field.clear_mask_cache()
        workspaces.clear_coupling_cache()


def test_fast_coupling_spot_check_falls_back_when_tolerance_is_tiny():
    from gmaster_amd import workload
    from gmaster_amd.learned import mask_window
    from gmaster_amd.workspaces import clear_coupling_cache, coupling_blocks, fast_coupling_blocks

    nside = 256
    lmax = 3 * nside - 1
    win = mask_window(workload.hole_mask(nside, device="cuda"), nside, lmax)
    clear_coupling_cache()
    exact = {k: v.clone() for k, v in coupling_blocks(win, lmax).items()}
    clear_coupling_cache()
    approx, info = fast_coupling_blocks(win, lmax, toeplitz=(200, 32, 32), tol=1.0)
    print(f"spot-check info {info}", flush=True)
    assert not info["fallback"] and info["sample_err"] > 0
    assert not torch.equal(approx["00"], exact["00"])
    clear_coupling_cache()
    back, info = fast_coupling_blocks(win, lmax, toeplitz=(200, 32, 32), tol=1e-14)
    assert info["fallback"]
    for k in exact:
        assert torch.equal(back[k], exact[k])


@pytest.mark.parametrize("nside,spin", [(512, 0), (256, 2), (512, 2)])
def test_large_geometry_round_trip(nside, spin):
    """Mid-size geometries, which the numpy oracle is too slow to reach: band-limited alms
    (l <= 2 nside) come back from synthesis + analysis with 3 Jacobi steps to the HEALPix
    quadrature's accuracy (~1e-4 here; an indexing or chunking mistake gives O(1)), and the
    FP64 and mixed configurations agree to FP32 rounding.  Configurations reached:
    (512, 0) analysis 2-wave R 4 (1024 folded rings <= 2048), synthesis 4-wave (large R);
    (256, 2) and (512, 2) analysis 4-wave (large R, 1023 / 2047 rings > 512), synthesis
    2-wave R 4 (<= 2048 rings).  The 4-wave spin-s syntheses and the large analyses at the
    benchmark's size are checked against the C oracle in tests/test_gpu_large.py."""
    from gmaster_amd import sht

    print(f"round trip nside={nside} spin={spin}", flush=True)
    lmax = 3 * nside - 1
    L = lmax + 1
    rng = np.random.default_rng(nside + spin)
    nelem = L * (L + 1) // 2
    ell = np.concatenate([np.arange(m, L) for m in range(L)])
    m = np.repeat(np.arange(L), L - np.arange(L))
    ncomp = 1 if spin == 0 else 2
    a = (rng.standard_normal((ncomp, nelem)) + 1j * rng.standard_normal((ncomp, nelem))) / (1.0 + ell)
    a[:, m == 0] = a[:, m == 0].real
    a[:, (ell > 2 * nside) | (ell < spin)] = 0
    at = torch.as_tensor(a, device="cuda")
    maps = sht.alm2map(at, spin, nside, lmax)
    b64 = sht.map2alm(maps, spin, lmax, n_iter=3)
    b32 = sht.map2alm(maps, spin, lmax, n_iter=3, precision="mixed")
    e64, e32, d = _rel(b64.cpu(), a), _rel(b32.cpu(), a), _rel(b32.cpu(), b64.cpu())
    print(f"  round trip rel fp64 {e64:.2e} mixed {e32:.2e}; mixed vs fp64 {d:.2e}", flush=True)
    assert e64 < 1e-3 and e32 < 1e-3
    assert d < 1e-6


@pytest.mark.parametrize("nside,nmaps", [(64, 13), (512, 12)])
def test_batched_maps_match_single_maps(nside, nmaps):
    """Maps sharing one recurrence in groups of 8, 4, 2, 1 (mixed) against each map alone in
    FP64, for analysis with refinement and for synthesis."""
    from gmaster_amd import sht, workload

    print(f"batched nside={nside} nmaps={nmaps}", flush=True)
    lmax = 3 * nside - 1
    mask = workload.hole_mask(nside, device="cuda")
    maps = torch.stack([workload.tqu_maps(nside, lmax, seed=s)[0] * mask for s in range(nmaps)])
    batch = sht.map2alm_s0(maps, lmax, n_iter=3, precision="mixed", max_group=8)
    single = torch.cat([sht.map2alm_s0(maps[i:i + 1], lmax, n_iter=3) for i in range(nmaps)])
    e = _rel(batch.cpu(), single.cpu())
    back_b = sht.alm2map_s0(single, nside, lmax, precision="mixed", max_group=8)
    back_s = torch.cat([sht.alm2map_s0(single[i:i + 1], nside, lmax) for i in range(nmaps)])
    es = _rel(back_b.cpu(), back_s.cpu())
    print(f"  analysis rel {e:.2e} synthesis rel {es:.2e}", flush=True)
    assert e < 1e-6 and es < 1e-6
