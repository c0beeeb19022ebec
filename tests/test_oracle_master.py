"""Bandpower windows and ``couple_cell`` of the oracle (no GPU).

Pinned by what the algebra fixes, independently of how the windows are assembled:

* any mask, beam and spin pair: applying the windows to a spectrum equals decoupling its
  coupled expectation (``couple_cell`` then ``decouple``), and the windows reproduce any
  band-constant spectrum exactly (``windows . T = I``);
* full sky (M = I) and no beam: the windows are the binning weights, block-diagonal in the
  spectrum index;
* planted mistakes (a dropped beam, the multipole and spectrum axes exchanged in the
  reshape) break the first identity.
"""

import numpy as np
import pytest

from oracle import master
from oracle.coupling import assemble_mcm, coupling_blocks


def _setup(lmax, nlb, spin1, spin2, seed, full_sky=False):
    rng = np.random.default_rng(seed)
    if full_sky:
        pcl = np.zeros(lmax + 1)
        pcl[0] = 4 * np.pi
    else:
        pcl = 4 * np.pi * rng.uniform(0.5, 1.0) / (1.0 + np.arange(lmax + 1)) ** 2.2
    mcm, ncls = assemble_mcm(coupling_blocks(pcl, lmax), spin1, spin2)
    ells, bpws, w = master.linear_bins(lmax, nlb)
    out_op, th_op = master.binning_operators(lmax, ells, bpws, w)
    return rng, mcm, ncls, out_op, th_op, (ells, bpws, w)


@pytest.mark.parametrize("spins", [(0, 0), (0, 2), (2, 2)])
def test_windows_equal_decoupled_expectation(spins):
    lmax, nlb = 47, 6
    rng, mcm, ncls, out_op, th_op, (ells, bpws, w) = _setup(lmax, nlb, *spins, seed=sum(spins) + 1)
    beam = np.exp(-0.5 * (np.arange(lmax + 1) / 30.0) ** 2)
    F = master.bandpower_windows(mcm, ncls, lmax, out_op, th_op, beam)
    nb = out_op.shape[0]
    assert F.shape == (ncls, nb, ncls, lmax + 1)
    cl = rng.normal(size=(ncls, lmax + 1))
    coupled = master.couple_cell(mcm, cl, ncls, lmax, beam)
    # the binned MCM decouple() uses carries the same beam
    eye = np.eye(ncls)
    mb = np.kron(out_op, eye) @ mcm @ np.diag(np.repeat(beam, ncls)) @ np.kron(th_op, eye)
    bp = master.decouple(mb, coupled, lmax, ells, bpws, w, ncls)
    assert np.allclose(np.einsum("cbdl,dl->cb", F, cl), bp, rtol=0, atol=1e-12 * np.max(np.abs(bp)))
    # band-constant spectra come back exactly
    c_b = rng.normal(size=(ncls, nb))
    spread = (th_op @ c_b.T).T
    assert np.allclose(np.einsum("cbdl,dl->cb", F, spread), c_b, atol=1e-11)
    # planted mistakes are caught
    no_beam = master.bandpower_windows(mcm, ncls, lmax, out_op, th_op)
    assert not np.allclose(np.einsum("cbdl,dl->cb", no_beam, cl), bp, atol=1e-6 * np.max(np.abs(bp)))
    if ncls > 1:
        flat = F.transpose(1, 0, 3, 2).reshape(nb * ncls, (lmax + 1) * ncls)
        bad = flat.reshape(nb, ncls, ncls, lmax + 1).transpose(1, 0, 2, 3)
        assert not np.allclose(np.einsum("cbdl,dl->cb", bad, cl), bp, atol=1e-6 * np.max(np.abs(bp)))


@pytest.mark.parametrize("spins", [(0, 0), (0, 2), (2, 2)])
def test_full_sky_windows_are_the_binning_weights(spins):
    lmax, nlb = 40, 5
    _, mcm, ncls, out_op, th_op, _ = _setup(lmax, nlb, *spins, seed=3, full_sky=True)
    F = master.bandpower_windows(mcm, ncls, lmax, out_op, th_op)
    for c in range(ncls):
        for d in range(ncls):
            # bands start at l = 2, so spin-2 spectra (no l < 2 modes) share the weights too
            ref = out_op if c == d else np.zeros_like(out_op)
            assert np.allclose(F[c, :, d], ref, atol=1e-12), (c, d)

def test_couple_cell_is_the_mcm_product():
    lmax = 30
    rng, mcm, ncls, *_ = _setup(lmax, 5, 2, 2, seed=9)
    cl = rng.normal(size=(ncls, lmax + 5))  # longer inputs are cut at lmax
    got = master.couple_cell(mcm, cl, ncls, lmax)
    for c in range(ncls):
        for l in (2, 17, lmax):
            ref = sum(mcm[l * ncls + c, lp * ncls + d] * cl[d, lp] for lp in range(lmax + 1) for d in range(ncls))
            assert abs(got[c, l] - ref) < 1e-12 * max(1.0, abs(ref))
