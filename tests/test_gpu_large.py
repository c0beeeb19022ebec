"""The Legendre kernels in their large-geometry configurations (4-wave workgroups, 5-8
rings per lane, several ring chunks; syntheses: 2-wave workgroups up to 2048 rings, 4-wave
above) against the C/OpenMP oracle on a sample of orders, FP64 and mixed, analyses and
syntheses.  These are the configurations the Nside-4096 benchmark runs; the numpy
oracle is too slow here, the C oracle (an independent natural-form recurrence with its own
extended exponent, itself pinned against the numpy oracle) is not."""

import os
import subprocess

import numpy as np
import pytest
import torch

from oracle import sht as osht

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def cb():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle", "c")], check=True)
    from oracle.c import cpu_baseline

    return cpu_baseline


def _orders(L):
    if L > 4000:  # the benchmark size: fewer orders keep the CPU side to seconds
        return np.unique(np.array([0, 2, 999, L // 2, L - 300, L - 1]))
    return np.unique(np.array([0, 1, 2, 3, 7, 64, L // 4, L // 2, (3 * L) // 4, L - 40, L - 2, L - 1]))


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("nside", [512, 1024, 4096])
@pytest.mark.parametrize("mixed", [False, True])
def test_spin0_analysis_large_geometry(cb, nside, mixed):
    from gmaster_amd import _lib, sht

    print(f"spin-0 large geometry nside={nside} mixed={mixed}", flush=True)
    lmax = 3 * nside - 1
    L = lmax + 1
    npix = 12 * nside * nside
    nring = 4 * nside - 1
    K = 1 if nside > 2048 else 2  # one map at the benchmark size bounds host memory
    maps = torch.randn((K, npix), dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(nside))
    spec = torch.empty((K, nring, L), dtype=torch.complex128, device="cuda")
    _lib.call("gm_ring_fft_fwd_s0", _lib.ptr(maps), K, npix, nside, lmax, _lib.ptr(spec), L, spec.stride(0),
              _lib.stream_ptr("cuda"))
    p = sht.plan(nside, lmax, 0, "cuda")
    nelem = L * (L + 1) // 2
    alm = torch.empty((K, nelem), dtype=torch.complex128, device="cuda")
    sht._analysis_s0(p, spec, alm, 0, mixed)
    orders = _orders(L)
    alm = alm.cpu().numpy()
    sp = spec[:, :, torch.as_tensor(orders, device="cuda")].cpu().numpy()
    del spec
    ai = osht.AlmInfo(lmax)
    for k in range(K):
        G = np.zeros((nring, 2 * L - 1), dtype=np.complex128)
        G[:, L - 1 + orders] = sp[k]
        out = cb.lat_ana_s(nside, lmax, 0, G, orders)
        for n, m in enumerate(orders):
            ref = out[n, m:, 0] + 1j * out[n, m:, 1]
            got = alm[k][ai.mstart[m] + np.arange(m, L)]
            e = _rel(got, ref)
            assert e < (2e-6 if mixed else 1e-10), (k, m, e)
    print("  ok", flush=True)


@pytest.mark.parametrize("nside", [256, 1024, 4096])
@pytest.mark.parametrize("mixed", [False, True])
def test_spin2_analysis_large_geometry(cb, nside, mixed):
    from gmaster_amd import _lib, sht

    print(f"spin-2 large geometry nside={nside} mixed={mixed}", flush=True)
    lmax = 3 * nside - 1
    L = lmax + 1
    W = 2 * L - 1
    npix = 12 * nside * nside
    nring = 4 * nside - 1
    qu = torch.randn((2, npix), dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(nside + 2))
    spec = torch.empty((nring, W), dtype=torch.complex128, device="cuda")
    _lib.call("gm_ring_fft_fwd_s", _lib.ptr(qu[0]), _lib.ptr(qu[1]), nside, lmax, _lib.ptr(spec), W,
              _lib.stream_ptr("cuda"))
    p = sht.plan(nside, lmax, 2, "cuda")
    nelem = L * (L + 1) // 2
    alm = torch.empty((2, nelem), dtype=torch.complex128, device="cuda")
    sht._analysis_s(p, spec, alm[0], alm[1], 0, mixed)
    alm = alm.cpu().numpy()
    orders = _orders(L)
    out = cb.lat_ana_s(nside, lmax, 2, spec.cpu().numpy(), orders)
    ap = np.zeros((L, W), dtype=np.complex128)
    ell = np.arange(L)
    for n, m in enumerate(orders):
        ap[:, L - 1 + m] = out[n, :, 0] + 1j * out[n, :, 1]
        if m:
            ap[:, L - 1 - m] = (-1.0) ** (ell + 2) * (out[n, :, 2] + 1j * out[n, :, 3])
    ai = osht.AlmInfo(lmax)
    eb = osht.plus_to_eb(ap, ai, 2)
    for n, m in enumerate(orders):
        idx = ai.mstart[m] + np.arange(m, L)
        for c in range(2):
            e = _rel(alm[c][idx], eb[c][idx])
            assert e < (2e-6 if mixed else 1e-10), (c, m, e)
    print("  ok", flush=True)


def _syn_err_ok(err, nside, mixed):
    """Per-ring relative errors of a synthesis against the extended-precision C oracle:
    below 5e-6 for the mixed form (FP32 accumulation over up to L = 12288 degrees: measured
    2.6e-6 at Nside 4096, m = 0) and below 1e-9 for FP64.  A synthesis reports every
    ring's value, so it sees the FP64 recurrence's own conditioning directly: near the
    poles at low m (x = cos(theta) -> 1, each step changes lambda by ~theta^2) and around
    the turning points.  At Nside 4096 (L = 12288) the natural-form FP64 recurrence of the
    C oracle is itself 1.5e-10 off the extended one at the polar ring for m = 0, and the
    kernels' mu-form, whose tables fold the normalisation P_l = B_l P_{l-2} into alpha,
    measured 4.4e-10 there and 1.2e-10 at m = 6144 (analyses, which sum over rings, stay
    below 1e-10)."""
    return float(np.max(err)) < (5e-6 if mixed else 1e-9)


def _gather_orders(alm, lmax, orders):
    """``(ncomp, norders, L)`` coefficients a_lm (zero below l = m) of the sampled orders,
    gathered on the device."""
    L = lmax + 1
    out = torch.zeros((alm.shape[0], len(orders), L), dtype=torch.complex128, device=alm.device)
    for k, m in enumerate(orders):
        start = m * (2 * lmax + 1 - m) // 2
        out[:, k, m:] = alm[:, start + m:start + L]
    return out.cpu().numpy()


@pytest.mark.parametrize("nside", [1024, 4096])
@pytest.mark.parametrize("mixed", [False, True])
def test_spin0_synthesis_large_geometry(cb, nside, mixed):
    """Spin-0 synthesis of a map pair (the pipeline's field + mask; a pair takes the mixed
    kernel) at Nside 1024 and 4096 (2048 and 8192 folded rings: the 4-wave configurations
    with 4 (FP64) and 5 (mixed) rings per lane) against the C oracle's direct channel."""
    from gmaster_amd import sht

    print(f"spin-0 synthesis large geometry nside={nside} mixed={mixed}", flush=True)
    lmax = 3 * nside - 1
    L = lmax + 1
    nring = 4 * nside - 1
    nelem = L * (L + 1) // 2
    g = torch.Generator("cuda").manual_seed(nside + 7)
    alm = torch.complex(torch.randn((2, nelem), dtype=torch.float64, device="cuda", generator=g),
                        torch.randn((2, nelem), dtype=torch.float64, device="cuda", generator=g))
    alm[:, :L] = alm[:, :L].real.to(torch.complex128)  # m = 0 real
    p = sht.plan(nside, lmax, 0, "cuda")
    spec = torch.empty((2, nring, L), dtype=torch.complex128, device="cuda")
    sht._synthesis_s0(p, alm, spec, mixed)
    orders = _orders(L)
    got = spec[:, :, torch.as_tensor(orders, device="cuda")].cpu().numpy()  # (2, nring, norders)
    coef = _gather_orders(alm, lmax, orders)
    del spec
    for k in range(2):
        out = cb.lat_syn_s(nside, lmax, 0, coef[k], np.zeros_like(coef[k]), orders, extended=True)
        for n, m in enumerate(orders):
            ref = out[n, :, 0] + 1j * out[n, :, 1]
            err = np.abs(got[k, :, n] - ref) / max(np.max(np.abs(ref)), 1e-300)
            assert _syn_err_ok(err, nside, mixed), (k, m, float(err.max()))
    print("  ok", flush=True)


@pytest.mark.parametrize("nside", [512, 1024, 4096])
@pytest.mark.parametrize("mixed", [False, True])
def test_spin2_synthesis_large_geometry(cb, nside, mixed):
    """Spin-2 synthesis at Nside 512 (2047 rings: 2-wave configuration), 1024 and 4096
    (4095 / 16383 rings: the 4-wave R = 6 configurations of the benchmark) against the C
    oracle's direct and mirror channels; the E/B -> a+ combination restates
    ``gmaster/_sht/healpix.py:599-611``."""
    from gmaster_amd import sht

    print(f"spin-2 synthesis large geometry nside={nside} mixed={mixed}", flush=True)
    s = 2
    lmax = 3 * nside - 1
    L = lmax + 1
    W = 2 * L - 1
    nring = 4 * nside - 1
    nelem = L * (L + 1) // 2
    g = torch.Generator("cuda").manual_seed(nside + 9)
    eb = torch.complex(torch.randn((2, nelem), dtype=torch.float64, device="cuda", generator=g),
                       torch.randn((2, nelem), dtype=torch.float64, device="cuda", generator=g))
    eb[:, :L] = eb[:, :L].real.to(torch.complex128)
    p = sht.plan(nside, lmax, s, "cuda")
    spec = torch.empty((nring, W), dtype=torch.complex128, device="cuda")
    sht._synthesis_s(p, eb[0], eb[1], spec, mixed)
    orders = _orders(L)
    cols = torch.as_tensor(np.concatenate([L - 1 + orders, L - 1 - orders]), device="cuda")
    got = spec[:, cols].cpu().numpy()
    del spec
    E, B = _gather_orders(eb, lmax, orders)
    ell = np.arange(L)
    sgn_m = ((-1.0) ** orders)[:, None]
    plus = -(E + 1j * B)                                # a+_{l, m}
    minus = -sgn_m * (np.conj(E) + 1j * np.conj(B))     # a+_{l, -m}
    for n, m in enumerate(orders):
        plus[n, :max(m, s)] = 0
        minus[n, :max(m, s)] = 0
    sg = (-1.0) ** s
    cd = sg * plus
    cm = sg * ((-1.0) ** (ell + s))[None, :] * minus
    out = cb.lat_syn_s(nside, lmax, s, cd, cm, orders, extended=True)
    no = len(orders)
    for n, m in enumerate(orders):
        scale = max(np.max(np.abs(out[n, :, 0:2])), np.max(np.abs(out[n, :, 2:4])))
        ed = np.abs(got[:, n] - (out[n, :, 0] + 1j * out[n, :, 1])) / scale
        assert _syn_err_ok(ed, nside, mixed), ("direct", m, float(ed.max()))
        if m:
            em = np.abs(got[:, no + n] - (out[n, :, 2] + 1j * out[n, :, 3])) / scale
            assert _syn_err_ok(em, nside, mixed), ("mirror", m, float(em.max()))
    print("  ok", flush=True)


@pytest.mark.parametrize("lmax", [3071, 12287])
def test_coupling_rows_large_lmax(cb, lmax):
    """The 3j coupling kernel at the benchmark's lmax against the C oracle's independent
    l3 recursion on a sample of rows (benchmark-like window: a decaying mask spectrum)."""
    from gmaster_amd.workspaces import clear_coupling_cache, coupling_blocks

    print(f"coupling rows lmax={lmax}", flush=True)
    L = lmax + 1
    ell = np.arange(L)
    window = 4 * np.pi * 0.9 / (1.0 + ell) ** 2.5
    clear_coupling_cache()
    blocks = coupling_blocks(torch.as_tensor(window, device="cuda"), lmax)
    rows = np.array([0, 2, 3, 17, L // 3, L // 2, L - 2, L - 1], dtype=np.int32)
    raw = cb.coupling_rows(lmax, window / (4 * np.pi), rows)  # (rows, L, 4), columns l2 >= l1
    col = 2 * ell + 1
    for k, name in enumerate(("00", "02", "22p", "22m")):
        got = blocks[name][torch.as_tensor(rows.astype(np.int64), device="cuda")].cpu().numpy()
        scale = np.max(np.abs(got))
        for n, l1 in enumerate(rows):
            ref = raw[n, l1:, k] * col[l1:]
            err = np.max(np.abs(got[n, l1:] - ref)) / scale
            assert err < 1e-12, (name, l1, err)
    clear_coupling_cache()
    print("  ok", flush=True)
