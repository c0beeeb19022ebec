"""Coupled C_ell, binning and decoupling (oracle, FP64).

Restates ``gmaster/workspaces.py:86-95`` (coupled C_ell), ``gmaster/bins.py:30-37,
65-101, 228-250`` (linear bands, weights, ``bin_cell``), ``gmaster/workspaces.py:
1452-1470, 1978-2003`` (binning operators, binned MCM) and ``:2106-2147``
(``decouple_cell``), ``:2074-2104`` (``couple_cell``) and ``:1628-1634, 1970-1990,
2149-2160`` (bandpower windows).  Pinned by the direct definitions (explicit loops over m and
bands in ``tests/test_oracle_master.py``) and, end to end, by the full-sky identity
(decoupled bandpowers == binned input pseudo-C_ell when the mask is 1).
"""

import numpy as np

from .coupling import assemble_mcm, coupling_blocks
from .geometry import AlmInfo


def coupled_cell(alm1, alm2, lmax):
    """``C_l^{ab} = sum_m c_m Re(a_lm conj b_lm) / (2l + 1)``, c_0 = 1, c_{m>0} = 2; rows a*n2 + b."""
    ainfo = AlmInfo(lmax)
    alm1 = np.atleast_2d(alm1)
    alm2 = np.atleast_2d(alm2)
    wt = np.where(ainfo.m == 0, 1.0, 2.0)
    out = []
    for a in alm1:
        for b in alm2:
            prod = (a * np.conj(b)).real * wt
            out.append(np.bincount(ainfo.ell, weights=prod, minlength=lmax + 1) / (2 * np.arange(lmax + 1) + 1))
    return np.array(out)


def linear_bins(lmax, nlb):
    """``(ells, bpws, weights)``: bands of ``nlb`` from l = 2, incomplete last band dropped."""
    ells = np.arange(lmax + 1)
    bpws = (ells - 2) // nlb
    bpws[:2] = -1
    if np.sum(bpws == bpws[-1]) != nlb:
        bpws[bpws == bpws[-1]] = -1
    keep = bpws >= 0
    ells, bpws = ells[keep], bpws[keep]
    nb = bpws.max() + 1
    w = np.ones(len(ells))
    w /= np.bincount(bpws, weights=w, minlength=nb)[bpws]
    return ells, bpws, w


def binning_operators(lmax, ells, bpws, weights, f_ell=None):
    nb = bpws.max() + 1
    f = np.ones(len(ells)) if f_ell is None else f_ell
    out = np.zeros((nb, lmax + 1))
    out[bpws, ells] = weights * f
    th = np.zeros((lmax + 1, nb))
    th[ells, bpws] = 1.0 / f
    return out, th


def bin_cell(cls, lmax, ells, bpws, weights):
    nb = bpws.max() + 1
    cls = np.atleast_2d(cls)
    return np.array([np.bincount(bpws, weights=c[ells] * weights, minlength=nb) for c in cls])


def binned_mcm(mcm, ncls, out_op, th_op):
    eye = np.eye(ncls)
    W = np.kron(out_op, eye)
    T = np.kron(th_op, eye)
    return W @ mcm @ T


def decouple(mcm_binned, cl_coupled, lmax, ells, bpws, weights, ncls):
    nb = bpws.max() + 1
    binned = bin_cell(cl_coupled, lmax, ells, bpws, weights).T.reshape(-1)
    return np.linalg.solve(mcm_binned, binned).reshape((nb, ncls)).T


def _beam_diag(beam, ncls, lmax):
    b = np.ones(lmax + 1) if beam is None else np.asarray(beam, dtype=np.float64)[: lmax + 1]
    return np.repeat(b, ncls)  # MCM index l * ncls + c


def couple_cell(mcm, cl_in, ncls, lmax, beam=None):
    """Expected coupled pseudo-spectra of full-sky spectra ``cl_in`` (``(ncls, >= L)``):
    ``(M diag(b1 b2) C)`` in the ``l * ncls + c`` ordering, back as ``(ncls, L)``
    (``gmaster/workspaces.py:2074-2104``)."""
    L = lmax + 1
    flat = np.asarray(cl_in, dtype=np.float64)[:, :L].T.reshape(-1) * _beam_diag(beam, ncls, lmax)
    return (mcm @ flat).reshape(L, ncls).T


def bandpower_windows(mcm, ncls, lmax, out_op, th_op, beam=None):
    """``(ncls, nb, ncls, L)`` windows ``F`` with ``<C_b^c> = sum_{c', l} F[c, b, c', l] C^{c'}_l``:
    the binned MCM ``W M diag(b1 b2) T`` solved against the one-sided ``W M diag(b1 b2)``
    (``gmaster/workspaces.py:1628-1634, 1970-1990, 2149-2160``)."""
    eye = np.eye(ncls)
    one_sided = np.kron(out_op, eye) @ mcm * _beam_diag(beam, ncls, lmax)[None, :]
    binned = one_sided @ np.kron(th_op, eye)
    nb = out_op.shape[0]
    return np.linalg.solve(binned, one_sided).reshape(nb, ncls, lmax + 1, ncls).transpose(1, 0, 3, 2)


def workspace(pcl_mask, lmax, spin1, spin2, nlb):
    """Binned MCM for a spin pair from the mask pseudo-C_ell (window zero above lmax)."""
    blocks = coupling_blocks(pcl_mask, lmax)
    mcm, ncls = assemble_mcm(blocks, spin1, spin2)
    ells, bpws, w = linear_bins(lmax, nlb)
    out_op, th_op = binning_operators(lmax, ells, bpws, w)
    return mcm, binned_mcm(mcm, ncls, out_op, th_op), ncls, (ells, bpws, w)


def pipeline_tp(mask, tqu, nside, lmax, nlb, n_iter):
    """Full T+P MASTER on the oracle: fields (n_iter Jacobi steps), mask pseudo-C_ell,
    00 / 02 / 22 workspaces, coupled spectra and decoupled bandpowers
    (``gmaster/field.py:78-315``, ``gmaster/workspaces.py:1668-2147``)."""
    from .sht import map2alm

    mask = np.asarray(mask, dtype=np.float64)
    tqu = np.asarray(tqu, dtype=np.float64)
    aT = map2alm(mask * tqu[0], 0, nside, lmax, n_iter)
    aM = map2alm(mask, 0, nside, lmax, n_iter)
    aP = map2alm(mask[None] * tqu[1:3], 2, nside, lmax, n_iter)
    pcl = coupled_cell(aM, aM, lmax)[0]
    blocks = coupling_blocks(pcl, lmax)
    ells, bpws, w = linear_bins(lmax, nlb)
    out_op, th_op = binning_operators(lmax, ells, bpws, w)
    res = {"alm_T": aT[0], "alm_mask": aM[0], "alm_E": aP[0], "alm_B": aP[1], "pcl_mask": pcl}
    for name, (s1, s2, a, b) in {"00": (0, 0, aT, aT), "02": (0, 2, aT, aP), "22": (2, 2, aP, aP)}.items():
        mcm, ncls = assemble_mcm(blocks, s1, s2)
        mb = binned_mcm(mcm, ncls, out_op, th_op)
        cl = coupled_cell(a, b, lmax)
        res["cl_" + name] = cl
        res["mcm_binned_" + name] = mb
        res["bp_" + name] = decouple(mb, cl, lmax, ells, bpws, w, ncls)
    return res
