"""Nside 8192 / lmax 24575: the largest configuration of the metric (bins of 32).

* Ring FFTs of rings above the LDS limit (the belt and the caps with q > 4096 have more
  than 8192 packed points and run as two halves) against numpy's FFT on sampled rings,
  forward and inverse, spin 0 and spin 2 (the in-LDS ring kernels are checked ring by ring
  at Nside 8 to 4096 in test_gpu_ringfft.py).
* Multi-pass Legendre analyses (the ring chunks are covered four partial slots at a time
  above Nside 4096) against the C oracle on sampled orders, FP64 and mixed.
* The full T+P MASTER pipeline in the chosen and the plain FP64 form: finite bandpowers
  that scatter around the input spectrum, and agreement within the metric's 1e-4.
"""

import os
import subprocess

import numpy as np
import pytest
import torch

from oracle.geometry import Rings

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NSIDE = 8192
LMAX = 3 * NSIDE - 1
L = LMAX + 1
NRING = 4 * NSIDE - 1
NPIX = 12 * NSIDE * NSIDE


@pytest.fixture(scope="module")
def rings():
    return Rings(NSIDE)


def _sample_rings():
    n = NSIDE
    # equator, a belt ring, the largest cap (q = 8191, prime), caps just above the LDS limit
    # (q = 4097, 4099) and a composite one (q = 6144), small caps, their southern mirrors
    north = [2 * n, n + 1, n - 1, 4097, 4099, 6144, 5, 1]
    r = [q - 1 for q in north] + [NRING - q for q in (n - 1, 4099, 1)]
    return sorted(set(r))


def _phase(m, R, r):
    return np.exp(-1j * np.pi * ((m * R.phi0_num[r]) % (2 * R.phi0_den[r])) / R.phi0_den[r])


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_ring_ffts_above_the_lds_limit(rings):
    from gmaster_amd import _lib
    R = rings
    st = _lib.stream_ptr("cuda")
    g = torch.Generator("cuda").manual_seed(81)
    sample = _sample_rings()
    print(f"rings {sample}", flush=True)
    # spin 0, forward then inverse of the same spectrum
    maps = torch.randn((1, NPIX), dtype=torch.float64, device="cuda", generator=g)
    spec = torch.empty((1, NRING, L), dtype=torch.complex128, device="cuda")
    _lib.call("gm_ring_fft_fwd_s0", _lib.ptr(maps), 1, NPIX, NSIDE, LMAX, _lib.ptr(spec), L, spec.stride(0), st)
    m = np.arange(L)
    for r in sample:
        n, o = int(R.nphi[r]), int(R.offset[r])
        X = np.fft.fft(maps[0, o:o + n].cpu().numpy())
        ref = R.weight * X[m % n] * _phase(m, R, r)
        e = _rel(spec[0, r].cpu().numpy(), ref)
        assert e < 1e-11, ("fwd_s0", r, n, e)  # Horner over the prime 8191: ~p eps
    back = torch.empty_like(maps)
    _lib.call("gm_ring_fft_inv_s0", _lib.ptr(spec), 1, L, spec.stride(0), NSIDE, LMAX, _lib.ptr(back), NPIX, st)
    for r in sample:
        n, o = int(R.nphi[r]), int(R.offset[r])
        F = spec[0, r].cpu().numpy()
        Y = np.zeros(n, dtype=np.complex128)
        np.add.at(Y, m % n, F * np.conj(_phase(m, R, r)))
        np.add.at(Y, (-m[1:]) % n, np.conj(F[1:] * np.conj(_phase(m[1:], R, r))))
        ref = (n * np.fft.ifft(Y)).real
        e = _rel(back[0, o:o + n].cpu().numpy(), ref)
        assert e < 5e-11, ("inv_s0", r, n, e)
    del maps, spec, back
    # spin 2
    qu = torch.randn((2, NPIX), dtype=torch.float64, device="cuda", generator=g)
    W = 2 * L - 1
    spec = torch.empty((NRING, W), dtype=torch.complex128, device="cuda")
    _lib.call("gm_ring_fft_fwd_s", _lib.ptr(qu[0]), _lib.ptr(qu[1]), NSIDE, LMAX, _lib.ptr(spec), W, st)
    ms = np.arange(-(L - 1), L)
    for r in sample:
        n, o = int(R.nphi[r]), int(R.offset[r])
        z = qu[0, o:o + n].cpu().numpy() + 1j * qu[1, o:o + n].cpu().numpy()
        ref = R.weight * np.fft.fft(z)[ms % n] * _phase(ms, R, r)
        e = _rel(spec[r].cpu().numpy(), ref)
        assert e < 1e-11, ("fwd_s", r, n, e)
    q2, u2 = torch.empty_like(qu[0]), torch.empty_like(qu[1])
    _lib.call("gm_ring_fft_inv_s", _lib.ptr(spec), W, NSIDE, LMAX, _lib.ptr(q2), _lib.ptr(u2), st)
    for r in sample:
        n, o = int(R.nphi[r]), int(R.offset[r])
        Y = np.zeros(n, dtype=np.complex128)
        np.add.at(Y, ms % n, spec[r].cpu().numpy() * np.conj(_phase(ms, R, r)))
        ref = n * np.fft.ifft(Y)
        got = q2[o:o + n].cpu().numpy() + 1j * u2[o:o + n].cpu().numpy()
        e = _rel(got, ref)
        assert e < 5e-11, ("inv_s", r, n, e)
    print("  ok", flush=True)


@pytest.fixture(scope="module")
def cb():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle", "c")], check=True)
    from oracle.c import cpu_baseline
    return cpu_baseline


ORDERS = np.array([0, 2, 4097, L // 2, L - 300, L - 1])


def _mstart(m):
    return m * (2 * LMAX + 1 - m) // 2


@pytest.mark.parametrize("mixed", [False, True])
def test_multipass_spin2_analysis(cb, mixed):
    from gmaster_amd import _lib, sht
    print(f"spin-2 analysis nside 8192 mixed={mixed}", flush=True)
    W = 2 * L - 1
    g = torch.Generator("cuda").manual_seed(82)
    qu = torch.randn((2, NPIX), dtype=torch.float64, device="cuda", generator=g)
    spec = torch.empty((NRING, W), dtype=torch.complex128, device="cuda")
    _lib.call("gm_ring_fft_fwd_s", _lib.ptr(qu[0]), _lib.ptr(qu[1]), NSIDE, LMAX, _lib.ptr(spec), W,
              _lib.stream_ptr("cuda"))
    del qu
    p = sht.plan(NSIDE, LMAX, 2, "cuda")
    passes = sht._passes(p, 2, mixed, 1, p.nchunks_mixed if mixed else p.nchunks, 0)
    print(f"  {len(passes)} passes", flush=True)
    assert len(passes) > 1
    nelem = L * (L + 1) // 2
    alm = torch.empty((2, nelem), dtype=torch.complex128, device="cuda")
    sht._analysis_s(p, spec, alm[0], alm[1], 0, mixed)
    oi = torch.as_tensor(L - 1 + ORDERS, device="cuda")
    dcol = spec[:, oi].T.cpu().numpy()
    mcol = spec[:, torch.as_tensor(L - 1 - ORDERS, device="cuda")].T.cpu().numpy()
    del spec
    out = cb.lat_ana_s_cols(NSIDE, LMAX, 2, dcol, mcol, ORDERS)
    ell = np.arange(L)
    for n, m in enumerate(ORDERS):
        plus = out[n, m:, 0] + 1j * out[n, m:, 1]
        if m:
            am = (-1.0) ** ell[m:] * (out[n, m:, 2] + 1j * out[n, m:, 3])
            minus = (-1.0) ** m * np.conj(am)
        else:
            minus = np.conj(plus)
        E = -(plus + minus) / 2
        B = 0.5j * (plus - minus)
        E[ell[m:] < 2] = 0
        B[ell[m:] < 2] = 0
        idx = torch.arange(_mstart(m) + m, _mstart(m) + L, device="cuda")
        for c, ref in enumerate((E, B)):
            got = alm[c, idx].cpu().numpy()
            e = _rel(got, ref)
            assert e < (2e-6 if mixed else 1e-10), (c, m, e)
    print("  ok", flush=True)


def test_multipass_spin0_mixed_analysis(cb):
    from gmaster_amd import _lib, sht
    print("spin-0 mixed analysis nside 8192", flush=True)
    g = torch.Generator("cuda").manual_seed(83)
    maps = torch.randn((1, NPIX), dtype=torch.float64, device="cuda", generator=g)
    spec = torch.empty((1, NRING, L), dtype=torch.complex128, device="cuda")
    _lib.call("gm_ring_fft_fwd_s0", _lib.ptr(maps), 1, NPIX, NSIDE, LMAX, _lib.ptr(spec), L, spec.stride(0),
              _lib.stream_ptr("cuda"))
    del maps
    p = sht.plan(NSIDE, LMAX, 0, "cuda")
    nch = int(_lib.query("gm_lat_nchunks_mixed", NSIDE, 0, 1))
    assert len(sht._passes(p, 0, True, 1, nch, 0)) > 1
    nelem = L * (L + 1) // 2
    alm = torch.empty((1, nelem), dtype=torch.complex128, device="cuda")
    sht._analysis_s0(p, spec, alm, 0, True)
    dcol = spec[0][:, torch.as_tensor(ORDERS, device="cuda")].T.cpu().numpy()
    del spec
    out = cb.lat_ana_s_cols(NSIDE, LMAX, 0, dcol, np.zeros_like(dcol), ORDERS)
    for n, m in enumerate(ORDERS):
        ref = out[n, m:, 0] + 1j * out[n, m:, 1]
        got = alm[0, torch.arange(_mstart(m) + m, _mstart(m) + L, device="cuda")].cpu().numpy()
        e = _rel(got, ref)
        assert e < 2e-6, (m, e)
    print("  ok", flush=True)


def test_full_master_nside8192_bins32():
    import gmaster_amd as nmt
    from gmaster_amd import field, sht, workload, workspaces
    print("T+P MASTER Nside 8192, lmax 24575, bins of 32", flush=True)
    mask = workload.hole_mask(NSIDE, nholes=20, radius_deg=5.0, apo_deg=2.0, kind="C1", seed=1234)
    tqu = workload.tqu_maps(NSIDE, seed=4321)
    b = nmt.NmtBin.from_nside_linear(NSIDE, 32)
    res = {}
    try:
        for form in ("fast", "fp64"):
            nmt.set_form(form)
            sht.clear_plans()
            workspaces.clear_coupling_cache()
            field.clear_mask_cache()
            torch.cuda.empty_cache()
            fT = nmt.NmtField(mask, tqu[0:1], n_iter=3)
            fP = nmt.NmtField(mask, tqu[1:3], n_iter=3)
            out = []
            for f1, f2 in ((fT, fT), (fT, fP), (fP, fP)):
                w = nmt.NmtWorkspace.from_fields(f1, f2, b)
                out.append(np.asarray(w.decouple_cell(nmt.compute_coupled_cell(f1, f2))))
            res[form] = out
            del fT, fP
            print(f"  {form} done", flush=True)
    finally:
        nmt.set_form("fp64")
        sht.clear_plans()
        workspaces.clear_coupling_cache()
        field.clear_mask_cache()
    tt, ee, bb, te = workload.theory_cls(LMAX)
    ttb = np.asarray(b.bin_cell(tt[None]))[0]
    leff = np.asarray(b.get_effective_ells())
    sel = (leff > 100) & (leff < 2 * NSIDE)
    for form, out in res.items():
        assert all(np.all(np.isfinite(o)) for o in out)
        dev = np.mean(np.abs(out[0][0][sel] / ttb[sel] - 1))
        print(f"  {form}: mean |TT/theory - 1| = {dev:.4f}", flush=True)
        assert dev < 0.03
    f, r = res["fast"], res["fp64"]
    err = {"TT": np.max(np.abs(f[0][0] / r[0][0] - 1)), "TE": np.max(np.abs(f[1][0] / r[1][0] - 1)),
           "EE": np.max(np.abs(f[2][0] / r[2][0] - 1)), "BB": np.max(np.abs(f[2][3] / r[2][3] - 1))}
    print("  fast vs fp64:", {k: f"{v:.2e}" for k, v in err.items()}, flush=True)
    assert max(err.values()) < 1e-4
