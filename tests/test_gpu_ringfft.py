"""Ring FFTs (gm_ring_fft_fwd_s0 / _s, gm_ring_fft_inv_s0 / _s) against numpy's FFT, ring by
ring: every ring at Nside 8, 32 and 256, and at Nside 1024 and 4096 the cap rings whose
packed length 2q has a large prime factor (q prime or twice / three times a prime: the
direct DFT pass with its Horner chains), their southern mirrors, small caps and belt rings.

Forward: the phased spectrum G[r][m] = w e^{-i m phi0} X[m mod nphi] (spin 0: m = 0..lmax;
spin 2: m = -lmax..lmax of Q + iU).  Inverse: the map sum_m F[m] e^{i m phi} (spin 0 with the
Hermitian half implied).  Tolerance 1e-11 of the ring's largest value (Horner over a
prime p carries ~p eps, p <= 4093 here)."""

import numpy as np
import pytest
import torch

from oracle.geometry import Rings

pytestmark = pytest.mark.gpu


def _phase(m, R, r):
    return np.exp(-1j * np.pi * ((m * R.phi0_num[r]) % (2 * R.phi0_den[r])) / R.phi0_den[r])


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _is_prime(n):
    return n > 1 and all(n % f for f in range(2, int(n ** 0.5) + 1))


def _sample(nside):
    nring = 4 * nside - 1
    if nside <= 256:
        return list(range(nring))
    caps = [q for q in range(nside - 1, nside // 2, -1)
            if _is_prime(q) or (q % 2 == 0 and _is_prime(q // 2)) or (q % 3 == 0 and _is_prime(q // 3))][:12]
    north = sorted(set(caps + [1, 2, 5, 7, 31, nside // 2 + 1, nside, nside + 1, 2 * nside]))
    south = [nring - q for q in caps[:3]]
    return sorted(set([q - 1 for q in north] + south))


@pytest.mark.parametrize("nside", [8, 32, 256, 1024, 4096])
def test_ring_fft_forward_and_inverse(nside):
    from gmaster_amd import _lib

    print(f"ring ffts nside={nside}", flush=True)
    R = Rings(nside)
    lmax = 3 * nside - 1
    L, W, npix, nring = lmax + 1, 2 * lmax + 1, 12 * nside * nside, 4 * nside - 1
    st = _lib.stream_ptr("cuda")
    g = torch.Generator("cuda").manual_seed(nside)
    sample = _sample(nside)
    m = np.arange(L)
    ms = np.arange(-(L - 1), L)

    maps = torch.randn((1, npix), dtype=torch.float64, device="cuda", generator=g)
    spec = torch.empty((1, nring, L), dtype=torch.complex128, device="cuda")
    _lib.call("gm_ring_fft_fwd_s0", _lib.ptr(maps), 1, npix, nside, lmax, _lib.ptr(spec), L, spec.stride(0), st)
    back = torch.empty_like(maps)
    _lib.call("gm_ring_fft_inv_s0", _lib.ptr(spec), 1, L, spec.stride(0), nside, lmax, _lib.ptr(back), npix, st)
    hm, hs, hb = maps[0].cpu().numpy(), spec[0, sample].cpu().numpy(), back[0].cpu().numpy()
    del maps, spec, back
    worst = {"fwd_s0": 0.0, "inv_s0": 0.0, "fwd_s": 0.0, "inv_s": 0.0}
    for i, r in enumerate(sample):
        n, o = int(R.nphi[r]), int(R.offset[r])
        ref = R.weight * np.fft.fft(hm[o:o + n])[m % n] * _phase(m, R, r)
        e = _rel(hs[i], ref)
        worst["fwd_s0"] = max(worst["fwd_s0"], e)
        assert e < 1e-11, ("fwd_s0", r, n, e)
        F = hs[i]
        Y = np.zeros(n, dtype=np.complex128)
        np.add.at(Y, m % n, F * np.conj(_phase(m, R, r)))
        np.add.at(Y, (-m[1:]) % n, np.conj(F[1:] * np.conj(_phase(m[1:], R, r))))
        ref = (n * np.fft.ifft(Y)).real
        e = _rel(hb[o:o + n], ref)
        worst["inv_s0"] = max(worst["inv_s0"], e)
        assert e < 1e-11, ("inv_s0", r, n, e)

    qu = torch.randn((2, npix), dtype=torch.float64, device="cuda", generator=g)
    spec = torch.empty((nring, W), dtype=torch.complex128, device="cuda")
    _lib.call("gm_ring_fft_fwd_s", _lib.ptr(qu[0]), _lib.ptr(qu[1]), nside, lmax, _lib.ptr(spec), W, st)
    q2, u2 = torch.empty_like(qu[0]), torch.empty_like(qu[1])
    _lib.call("gm_ring_fft_inv_s", _lib.ptr(spec), W, nside, lmax, _lib.ptr(q2), _lib.ptr(u2), st)
    hq, hs = qu.cpu().numpy(), spec[sample].cpu().numpy()
    hz = q2.cpu().numpy() + 1j * u2.cpu().numpy()
    del qu, spec, q2, u2
    for i, r in enumerate(sample):
        n, o = int(R.nphi[r]), int(R.offset[r])
        z = hq[0, o:o + n] + 1j * hq[1, o:o + n]
        ref = R.weight * np.fft.fft(z)[ms % n] * _phase(ms, R, r)
        e = _rel(hs[i], ref)
        worst["fwd_s"] = max(worst["fwd_s"], e)
        assert e < 1e-11, ("fwd_s", r, n, e)
        Y = np.zeros(n, dtype=np.complex128)
        np.add.at(Y, ms % n, hs[i] * np.conj(_phase(ms, R, r)))
        ref = n * np.fft.ifft(Y)
        e = _rel(hz[o:o + n], ref)
        worst["inv_s"] = max(worst["inv_s"], e)
        assert e < 1e-11, ("inv_s", r, n, e)
    print(f"  {len(sample)} rings, worst {worst}", flush=True)
