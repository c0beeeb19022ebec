"""Batched transforms (T, mask and Q/U through one spin-0 recurrence, ``gm_bat_*``) against
the oracle: the numpy transforms at small Nside (every order, with and without Jacobi
steps), the C oracle on sampled orders at the benchmark geometry, and the per-spin FP64
kernels over whole alms.  Tolerances are those of the mixed form (FP32 contraction over
the FP64 recurrence): 2e-6 of the largest value per order, except the spin-2 synthesis at
the benchmark geometry, 5e-6: there the spin-2 rows are sums of a 1/sin^2 part and a
plain part that partly cancel, and at the FP32 region's edge (sin^2 = 1/16) the 1/sin^2
part's FP32 rounding is amplified up to 16-fold (measured 2.6e-6 at Nside 4096, m = 999)."""

import os
import subprocess

import numpy as np
import pytest
import torch

from oracle import sht as osht

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOL = 2e-6
TOL_SYN2 = 5e-6


@pytest.fixture(scope="module")
def cb():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle", "c")], check=True)
    from oracle.c import cpu_baseline

    return cpu_baseline


def _maps(nside, seed, k0=2):
    g = torch.Generator("cuda").manual_seed(seed)
    npix = 12 * nside * nside
    m0 = torch.randn((k0, npix), dtype=torch.float64, device="cuda", generator=g)
    m2 = torch.randn((1, 2, npix), dtype=torch.float64, device="cuda", generator=g)
    return m0, m2


def _per_order_err(got, ref, ai, where=False, band=0):
    """Largest |got - ref| of each order relative to the largest |ref| of that order (of the
    orders m - band .. m with ``band``: the last orders hold one or two alms, whose own size
    is a noisy scale for white-noise maps)."""
    L = ai.lmax + 1
    worst, at = 0.0, None
    scale = np.array([np.max(np.abs(ref[ai.mstart[m] + np.arange(m, L)])) for m in range(L)])
    for m in range(L):
        idx = ai.mstart[m] + np.arange(m, L)
        sc = max(np.max(scale[max(0, m - band):m + 1]), 1e-300)
        d = np.abs(got[idx] - ref[idx]) / sc
        if d.max() > worst:
            worst, at = float(d.max()), (m, m + int(np.argmax(d)))
    return (worst, at) if where else worst


@pytest.mark.parametrize("nside", [8, 16, 32])
@pytest.mark.parametrize("n_iter", [0, 3])
def test_batched_map2alm_small_vs_oracle(nside, n_iter):
    from gmaster_amd import batched

    print(f"batched map2alm nside={nside} n_iter={n_iter}", flush=True)
    lmax = 3 * nside - 1
    m0, m2 = _maps(nside, 100 + nside)
    a0, a2 = batched.map2alm(m0, m2, lmax, n_iter=n_iter)
    a0, a2 = a0.cpu().numpy(), a2.cpu().numpy()
    ai = osht.AlmInfo(lmax)
    h0, h2 = m0.cpu().numpy(), m2.cpu().numpy()
    errs = []
    for k in range(2):
        ref = osht.map2alm_s0(h0[k], nside, lmax, n_iter)
        errs.append(_per_order_err(a0[k], ref, ai, where=True))
    ref = osht.map2alm_spin(h2[0], nside, lmax, 2, n_iter)
    for c in range(2):
        errs.append(_per_order_err(a2[0, c], ref[c], ai, where=True))
    print("  per-order errors (error, (m, l)):", errs, flush=True)
    for i, (e, at) in enumerate(errs):
        assert e < TOL, (i, e, at)
    print("  ok", flush=True)


@pytest.mark.parametrize("nside", [8, 32])
def test_batched_alm2map_small_vs_oracle(nside):
    from gmaster_amd import batched

    print(f"batched alm2map nside={nside}", flush=True)
    lmax = 3 * nside - 1
    ai = osht.AlmInfo(lmax)
    rng = np.random.default_rng(nside)
    a = rng.normal(size=(4, ai.nelem)) + 1j * rng.normal(size=(4, ai.nelem))
    a[:, : lmax + 1] = a[:, : lmax + 1].real
    a[2:, ai.ell < 2] = 0
    t = torch.as_tensor(a, device="cuda")
    mp0, mp2 = batched.alm2map(t[0:2], t[2:4][None], nside, lmax)
    mp0, mp2 = mp0.cpu().numpy(), mp2.cpu().numpy()
    for k in range(2):
        ref = osht.alm2map_s0(a[k], nside, lmax)
        assert np.max(np.abs(mp0[k] - ref)) < TOL * np.max(np.abs(ref)), k
    ref = osht.alm2map_spin(a[2:4], nside, lmax, 2)
    assert np.max(np.abs(mp2[0] - ref)) < TOL * np.max(np.abs(ref))
    print("  ok", flush=True)


@pytest.mark.parametrize("k0,k2", [(1, 1), (0, 1), (0, 2), (2, 0), (4, 0), (8, 0)])
def test_batched_other_shapes(k0, k2):
    """Every supported batch shape agrees with the (2, 1) batch's channels and the oracle."""
    from gmaster_amd import batched

    nside, lmax = 16, 47
    print(f"batched shape k0={k0} k2={k2}", flush=True)
    g = torch.Generator("cuda").manual_seed(7 + k0 + 10 * k2)
    npix = 12 * nside * nside
    m0 = torch.randn((k0, npix), dtype=torch.float64, device="cuda", generator=g) if k0 else None
    m2 = torch.randn((k2, 2, npix), dtype=torch.float64, device="cuda", generator=g) if k2 else None
    a0, a2 = batched.map2alm(m0, m2, lmax, n_iter=1)
    ai = osht.AlmInfo(lmax)
    for k in range(k0):
        ref = osht.map2alm_s0(m0[k].cpu().numpy(), nside, lmax, 1)
        assert _per_order_err(a0[k].cpu().numpy(), ref, ai) < TOL, k
    for f in range(k2):
        ref = osht.map2alm_spin(m2[f].cpu().numpy(), nside, lmax, 2, 1)
        for c in range(2):
            assert _per_order_err(a2[f, c].cpu().numpy(), ref[c], ai) < TOL, (f, c)
    print("  ok", flush=True)


@pytest.mark.parametrize("nside", [256, 512])
def test_batched_vs_fp64_kernels(nside):
    """Whole alms of white-noise maps with n_iter 3 against the per-spin FP64 kernels
    (Nside 256, 512: the FP32 ring blocks, several FP64 polar blocks, every order), next to
    the per-spin mixed kernels' own error: the batched form must stay within 2e-6 of the
    largest alm of each order and the 16 orders below it, or within twice the per-spin mixed
    form's error."""
    from gmaster_amd import batched, sht

    print(f"batched vs fp64 nside={nside}", flush=True)
    lmax = 3 * nside - 1
    m0, m2 = _maps(nside, 5 + nside)
    a0, a2 = batched.map2alm(m0, m2, lmax, n_iter=3)
    r0 = sht.map2alm_s0(m0, lmax, n_iter=3, precision="fp64")
    r2 = sht.map2alm_spin(m2[0], lmax, 2, n_iter=3, precision="fp64")
    x0 = sht.map2alm_s0(m0, lmax, n_iter=3, precision="mixed")
    x2 = sht.map2alm_spin(m2[0], lmax, 2, n_iter=3, precision="mixed")
    ai = osht.AlmInfo(lmax)
    pairs = [(a0[k], x0[k], r0[k]) for k in range(2)] + [(a2[0, c], x2[c], r2[c]) for c in range(2)]
    errs = []
    for i, (got, mixed, ref) in enumerate(pairs):
        ref = ref.cpu().numpy()
        e_bat, at = _per_order_err(got.cpu().numpy(), ref, ai, where=True, band=16)
        e_mix, at_mix = _per_order_err(mixed.cpu().numpy(), ref, ai, where=True, band=16)
        print(f"  channel {i}: batched {e_bat:.2e} at (m, l) {at}, per-spin mixed {e_mix:.2e} at {at_mix}",
              flush=True)
        errs.append((i, e_bat, e_mix))
    for i, e_bat, e_mix in errs:
        assert e_bat < max(TOL, 2.0 * e_mix), (i, e_bat, e_mix)
    print("  ok", flush=True)


@pytest.mark.parametrize("nside", [256, 1024])
def test_residual_analysis_vs_fp32(nside):
    """The refinement passes' analysis (FP16 hi/lo operands on the matrix unit) against the
    FP32 analysis on the same ring spectra of white-noise maps, every order: within 4e-6 of
    the largest alm of each order and the 16 below it (the split carries each product to
    ~2^-21; the measured error is printed)."""
    from gmaster_amd import batched

    print(f"residual analysis nside={nside}", flush=True)
    lmax = 3 * nside - 1
    m0, m2 = _maps(nside, 11 + nside)
    bp = batched.plan(nside, lmax, "cuda")
    spec0, spec2 = batched._fft_fwd(bp, m0, m2)
    L = lmax + 1
    nelem = L * (L + 1) // 2
    a0 = torch.empty((2, nelem), dtype=torch.complex128, device="cuda")
    a2 = torch.empty((1, 2, nelem), dtype=torch.complex128, device="cuda")
    h0, h2 = torch.empty_like(a0), torch.empty_like(a2)
    batched.analysis(bp, spec0, spec2, a0, a2, 0)
    batched.analysis(bp, spec0, spec2, h0, h2, 0, residual=True)
    ai = osht.AlmInfo(lmax)
    worst = 0.0
    for i, (got, ref) in enumerate([(h0[0], a0[0]), (h0[1], a0[1]), (h2[0, 0], a2[0, 0]), (h2[0, 1], a2[0, 1])]):
        e, at = _per_order_err(got.cpu().numpy(), ref.cpu().numpy(), ai, where=True, band=16)
        print(f"  channel {i}: {e:.2e} at (m, l) {at}", flush=True)
        worst = max(worst, e)
        assert torch.isfinite(got).all(), i
    assert worst < 4e-6, worst
    print("  ok", flush=True)


def _orders(L):
    if L > 4000:
        return np.unique(np.array([0, 1, 2, 3, 999, L // 2, L - 300, L - 1]))
    return np.unique(np.array([0, 1, 2, 5, 64, L // 4, L // 2, L - 40, L - 1]))


@pytest.mark.parametrize("nside", [1024, 4096])
def test_batched_analysis_large_vs_c_oracle(cb, nside):
    """One analysis pass (no refinement) at Nside 1024 and 4096 against the C oracle's
    spin-0 and spin-2 direct sums on sampled orders."""
    from gmaster_amd import batched

    print(f"batched analysis large nside={nside}", flush=True)
    lmax = 3 * nside - 1
    L = lmax + 1
    nring = 4 * nside - 1
    k0 = 2 if nside <= 2048 else 1
    m0, m2 = _maps(nside, 41 + nside, k0=k0)
    bp = batched.plan(nside, lmax, "cuda")
    spec0, spec2 = batched._fft_fwd(bp, m0, m2)
    del m0, m2
    nelem = L * (L + 1) // 2
    a0 = torch.empty((k0, nelem), dtype=torch.complex128, device="cuda")
    a2 = torch.empty((1, 2, nelem), dtype=torch.complex128, device="cuda")
    batched.analysis(bp, spec0, spec2, a0, a2, 0)  # a (1, 1) batch at Nside 4096 bounds host memory
    orders = _orders(L)
    oc = torch.as_tensor(orders, device="cuda")
    s0 = spec0[:, :, oc].cpu().numpy()
    dcol = spec2[0][:, L - 1 + oc].T.contiguous().cpu().numpy()
    mcol = spec2[0][:, L - 1 - oc].T.contiguous().cpu().numpy()
    del spec0, spec2
    ai = osht.AlmInfo(lmax)
    a0h = [a0[k] for k in range(k0)]
    for k in range(k0):
        out = cb.lat_ana_s_cols(nside, lmax, 0, s0[k].T.copy(), np.zeros_like(s0[k].T), orders)
        for n, m in enumerate(orders):
            idx = torch.as_tensor(ai.mstart[m] + np.arange(m, L), device="cuda")
            got = a0h[k][idx].cpu().numpy()
            ref = out[n, m:, 0] + 1j * out[n, m:, 1]
            e = float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300))
            assert e < TOL, (k, m, e)
    out = cb.lat_ana_s_cols(nside, lmax, 2, dcol, mcol, orders)
    ell = np.arange(L)
    for n, m in enumerate(orders):
        plus = out[n, :, 0] + 1j * out[n, :, 1]
        mir = (-1.0) ** ell * (out[n, :, 2] + 1j * out[n, :, 3]) if m else plus
        minus = (-1) ** m * np.conj(mir)
        E = -(plus + minus) / 2
        B = 0.5j * (plus - minus)
        E[: max(m, 2)] = 0
        B[: max(m, 2)] = 0
        idx = torch.as_tensor(ai.mstart[m] + np.arange(m, L), device="cuda")
        for c, ref in ((0, E[m:]), (1, B[m:])):
            got = a2[0, c][idx].cpu().numpy()
            e = float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300))
            assert e < TOL, (c, m, e)
    print("  ok", flush=True)


@pytest.mark.parametrize("nside", [1024, 4096])
def test_batched_synthesis_large_vs_c_oracle(cb, nside):
    """One synthesis at Nside 1024 and 4096 against the C oracle (extended precision) on
    sampled orders: spin-0 direct sums and the spin-2 F+(+-m) of the E/B coefficients."""
    from gmaster_amd import batched

    print(f"batched synthesis large nside={nside}", flush=True)
    lmax = 3 * nside - 1
    L = lmax + 1
    W = 2 * L - 1
    nring = 4 * nside - 1
    nelem = L * (L + 1) // 2
    g = torch.Generator("cuda").manual_seed(nside + 77)

    def rnd(*shape):
        return torch.complex(torch.randn(shape, dtype=torch.float64, device="cuda", generator=g),
                             torch.randn(shape, dtype=torch.float64, device="cuda", generator=g))

    a0 = rnd(1, nelem)
    a2 = rnd(1, 2, nelem)
    a0[:, :L] = a0[:, :L].real.to(torch.complex128)
    a2[:, :, :L] = a2[:, :, :L].real.to(torch.complex128)
    ell_t = torch.as_tensor(osht.AlmInfo(lmax).ell, device="cuda")
    a2[:, :, ell_t < 2] = 0
    bp = batched.plan(nside, lmax, "cuda")
    spec0 = torch.empty((1, nring, L), dtype=torch.complex128, device="cuda")
    spec2 = torch.empty((1, nring, W), dtype=torch.complex128, device="cuda")
    batched.synthesis(bp, a0, a2, spec0, spec2)
    orders = _orders(L)
    oc = torch.as_tensor(orders, device="cuda")
    g0 = spec0[0][:, oc].cpu().numpy()
    gp = spec2[0][:, L - 1 + oc].cpu().numpy()
    gm = spec2[0][:, L - 1 - oc].cpu().numpy()
    del spec0, spec2
    ai = osht.AlmInfo(lmax)

    def gather(a):
        out = np.zeros((len(orders), L), dtype=np.complex128)
        for n, m in enumerate(orders):
            idx = torch.as_tensor(ai.mstart[m] + np.arange(m, L), device="cuda")
            out[n, m:] = a[idx].cpu().numpy()
        return out

    c0 = gather(a0[0])
    out = cb.lat_syn_s(nside, lmax, 0, c0, np.zeros_like(c0), orders, extended=True)
    for n, m in enumerate(orders):
        ref = out[n, :, 0] + 1j * out[n, :, 1]
        e = float(np.max(np.abs(g0[:, n] - ref)) / max(np.max(np.abs(ref)), 1e-300))
        assert e < TOL, (m, e)
    E, B = gather(a2[0, 0]), gather(a2[0, 1])
    ell = np.arange(L)
    sgn_m = ((-1.0) ** orders)[:, None]
    plus = -(E + 1j * B)
    minus = -sgn_m * (np.conj(E) + 1j * np.conj(B))
    cd = plus
    cm = ((-1.0) ** ell)[None, :] * minus
    out = cb.lat_syn_s(nside, lmax, 2, cd, cm, orders, extended=True)
    for n, m in enumerate(orders):
        scale = max(np.max(np.abs(out[n, :, 0:2])), np.max(np.abs(out[n, :, 2:4])))
        ed = float(np.max(np.abs(gp[:, n] - (out[n, :, 0] + 1j * out[n, :, 1]))) / scale)
        assert ed < TOL_SYN2, ("+m", m, ed)
        if m:
            em = float(np.max(np.abs(gm[:, n] - (out[n, :, 2] + 1j * out[n, :, 3]))) / scale)
            assert em < TOL_SYN2, ("-m", m, em)
    print("  ok", flush=True)
