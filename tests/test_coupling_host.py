"""The coupling recursion (``gmaster_amd/csrc/coupling_rec.h``) built for the host and run on
the CPU: every pair l1 <= l2 through the per-pair driver (the pair-list kernels' code) and
through the one-l1-per-wave driver of the full-matrix kernel (lanes emulated one by one, its
work-item decode checked to cover the triangle exactly once and its LDS window reads checked
to stay inside the 128 staged entries).

Pinned against the oracle's Gauss-Legendre coupling (``oracle/coupling.py``, an independent
quadrature of Wigner-d products, long-double node sums) and the exact Racah 3j fixture at
lmax 12: windows longer than 2 lmax (no truncation) and cut at lmax or below (lanes of one
wave end at different degrees), decaying and white spectra.  The tolerance is the oracle's
own: its node sums cancel by up to ~1e3 for white windows, so 1e-9 of the block's largest
entry there and 1e-10 for decaying windows.  The two drivers agree bit for bit where no lane
is cut short and to 1e-15 elsewhere."""

import os
import subprocess

import numpy as np
import pytest

from oracle.coupling import coupling_blocks, coupling_blocks_exact

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "host", "coupling_host.cpp")


@pytest.fixture(scope="module")
def host_bin(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("coupling_host") / "coupling_host")
    cmd = ["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "--offload-arch=gfx950", "-x", "hip", SRC, "-o", out]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        pytest.fail(res.stderr[-2000:])
    return out


def _run(host_bin, tmp_path, window, lmax):
    L = lmax + 1
    ell = np.arange(len(window))
    wp = window * (2 * ell + 1) / (4 * np.pi)
    wp.tofile(tmp_path / "w")
    res = subprocess.run([host_bin, str(lmax), str(len(window) - 1), str(tmp_path / "w"), str(tmp_path / "p"),
                          str(tmp_path / "q")])
    assert res.returncode == 0, "work-item decode or window bounds check failed"
    col = 2 * np.arange(L) + 1
    pair = np.fromfile(tmp_path / "p").reshape(L, L, 4) * col[None, :, None]
    wave = np.fromfile(tmp_path / "q").reshape(L, L, 4) * col[None, :, None]
    return pair, wave


def _window(kind, n, seed):
    ell = np.arange(n)
    if kind == "decay":
        return 4 * np.pi * 0.9 / (1.0 + ell) ** 2.5
    return np.random.default_rng(seed).uniform(size=n)


@pytest.mark.parametrize("lmax,lw", [(40, 80), (40, 40), (40, 25), (130, 260), (130, 130), (130, 70), (200, 400)])
@pytest.mark.parametrize("kind", ["decay", "white"])
def test_recursion_matches_quadrature_oracle(host_bin, tmp_path, lmax, lw, kind):
    window = _window(kind, lw + 1, lmax + lw)
    pair, wave = _run(host_bin, tmp_path, window, lmax)
    ref = coupling_blocks(window, lmax)
    iu = np.triu_indices(lmax + 1)
    tol = 1e-9 if kind == "white" else 1e-10
    for k, name in enumerate(("00", "02", "22p", "22m")):
        scale = np.max(np.abs(ref[k]))
        err = np.max(np.abs(pair[:, :, k][iu] - ref[k][iu])) / scale
        assert err < tol, (name, err)
        dw = np.max(np.abs(wave[:, :, k][iu] - pair[:, :, k][iu])) / scale
        if lw >= 2 * lmax:
            assert dw == 0.0, (name, dw)
        else:
            assert dw < 1e-15, (name, dw)


def test_recursion_matches_exact_3j(host_bin, tmp_path):
    window = _window("white", 25, 7)
    pair, wave = _run(host_bin, tmp_path, window, 12)
    ref = coupling_blocks_exact(window, 12)
    iu = np.triu_indices(13)
    for k in range(4):
        assert np.max(np.abs(pair[:, :, k][iu] - ref[k][iu])) / np.max(np.abs(ref[k])) < 1e-13, k
        assert np.array_equal(wave[:, :, k][iu], pair[:, :, k][iu]), k
