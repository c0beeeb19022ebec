"""CPU baseline for bench.py: the C/OpenMP oracle timed on a bounded sample at the
benchmark's own Nside, extrapolated to the whole T+P pipeline (oracle / test
infrastructure -- never part of the product).

Extrapolated work (the two stages that carry >95% of the FLOP):

* Legendre: 4 analyses + 3 syntheses per field, for the T + mask pair (folded rings,
  ``T0 = L(L+1)/2 * 2 nside`` (m, l, ring) triples per pass, 4 channels) and for the
  spin-2 field (all rings, ``T2 = L(L+1)/2 * (4 nside - 1)`` triples, 4 channels).
  Time per triple comes from the spin-2 analysis over a uniform sample of orders m
  (the full triangle, no polar skip), so both pass types cost the same per triple.
* Coupling: ``N3j = sum_{l1 <= l2} (min(l1 + l2, lmax) - (l2 - l1) + 1)`` l3 terms, timed on
  a uniform sample of rows l1.

Ring FFTs, coupled spectra, binning and the LU solve are not included, so the estimate is
a lower bound on the CPU pipeline time.
"""

import ctypes
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libgmoracle.so")


def load():
    if not os.path.exists(LIB):
        raise RuntimeError(f"{LIB} missing: run `make -C oracle/c`")
    lib = ctypes.CDLL(LIB)
    p = ctypes.c_void_p
    lib.gmo_lat_ana_s.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, p, ctypes.c_int64, p, ctypes.c_int, p]
    lib.gmo_lat_ana_s.restype = None
    lib.gmo_lat_ana_s_cols.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, p, p, p, ctypes.c_int, p]
    lib.gmo_lat_ana_s_cols.restype = None
    lib.gmo_lat_syn_s.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, p, p, p, ctypes.c_int, p]
    lib.gmo_lat_syn_s.restype = None
    lib.gmo_lat_syn_s_ext.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, p, p, p, ctypes.c_int, p]
    lib.gmo_lat_syn_s_ext.restype = None
    lib.gmo_coupling.argtypes = [ctypes.c_int, p, ctypes.c_int, p, p, ctypes.c_int, p]
    lib.gmo_coupling.restype = None
    return lib


def lat_ana_s(nside, lmax, spin, spec, orders):
    """``(norders, L, 4)``: direct (re, im) and mirror (re, im) sums per degree."""
    lib = load()
    spec = np.ascontiguousarray(spec, dtype=np.complex128)
    orders = np.ascontiguousarray(orders, dtype=np.int32)
    L = lmax + 1
    out = np.zeros((len(orders), L, 4))
    lib.gmo_lat_ana_s(nside, lmax, spin, spec.ctypes.data, spec.shape[1], orders.ctypes.data, len(orders),
                      out.ctypes.data)
    return out


def lat_ana_s_cols(nside, lmax, spin, dcol, mcol, orders):
    """As ``lat_ana_s`` from per-order spectrum columns ``dcol[k][r]`` (order +m) and
    ``mcol[k][r]`` (order -m), each ``(len(orders), nring)`` complex."""
    lib = load()
    dcol = np.ascontiguousarray(dcol, dtype=np.complex128)
    mcol = np.ascontiguousarray(mcol, dtype=np.complex128)
    orders = np.ascontiguousarray(orders, dtype=np.int32)
    if dcol.shape != (len(orders), 4 * nside - 1) or mcol.shape != dcol.shape:
        raise ValueError("columns must be (len(orders), nring)")
    L = lmax + 1
    out = np.zeros((len(orders), L, 4))
    lib.gmo_lat_ana_s_cols(nside, lmax, spin, dcol.ctypes.data, mcol.ctypes.data, orders.ctypes.data, len(orders),
                           out.ctypes.data)
    return out


def lat_syn_s(nside, lmax, spin, cd, cm, orders, extended=False):
    """``(norders, nring, 4)``: per ring, direct ``sum_l cd[k][l] lambda_lm(theta_r)`` and
    mirror ``sum_l cm[k][l] lambda_lm(pi - theta_r)`` (re, im); ``cd``, ``cm`` are
    ``(len(orders), L)`` complex.  ``extended=True`` runs the recurrence and sums in x87
    extended precision (a reference for judging FP64 results at the polar rings, where
    the FP64 recurrence's own rounding reaches ~1e-10 relative at L = 12288)."""
    lib = load()
    L = lmax + 1
    cd = np.ascontiguousarray(cd, dtype=np.complex128)
    cm = np.ascontiguousarray(cm, dtype=np.complex128)
    orders = np.ascontiguousarray(orders, dtype=np.int32)
    if cd.shape != (len(orders), L) or cm.shape != cd.shape:
        raise ValueError("coefficients must be (len(orders), lmax + 1)")
    out = np.zeros((len(orders), 4 * nside - 1, 4))
    fn = lib.gmo_lat_syn_s_ext if extended else lib.gmo_lat_syn_s
    fn(nside, lmax, spin, cd.ctypes.data, cm.ctypes.data, orders.ctypes.data, len(orders), out.ctypes.data)
    return out


def coupling_rows(lmax, window, rows):
    """``(nrows, L, 4)`` raw sums (00, 02, 22 even, 22 odd) for rows l1, columns l2 >= l1."""
    from ..wigner3j import g_table

    lib = load()
    window = np.ascontiguousarray(window, dtype=np.float64)
    g = np.ascontiguousarray(g_table(lmax + 2))
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    out = np.zeros((len(rows), lmax + 1, 4))
    lib.gmo_coupling(lmax, window.ctypes.data, len(window) - 1, g.ctypes.data, rows.ctypes.data, len(rows),
                     out.ctypes.data)
    return out


def n3j(lmax):
    L = lmax + 1
    l1 = np.arange(L, dtype=np.int64)[:, None]
    l2 = np.arange(L, dtype=np.int64)[None, :]
    cnt = np.minimum(l1 + l2, lmax) - (l2 - l1) + 1
    return int(np.sum(np.where(l2 >= l1, np.maximum(cnt, 0), 0)))


def estimate_pipeline(nside, lmax, target_seconds=20.0, log=print):
    cores = len(os.sched_getaffinity(0))
    os.environ["OMP_NUM_THREADS"] = str(cores)
    L = lmax + 1
    nring = 4 * nside - 1
    rng = np.random.default_rng(0)
    W = 2 * L - 1
    spec = (rng.standard_normal((nring, W)) + 1j * rng.standard_normal((nring, W))) * 1e-3
    # Legendre: calibrate with a few orders, then size the sample to ~60% of the budget
    probe = np.array([L // 2], dtype=np.int32)
    t = time.perf_counter()
    lat_ana_s(nside, lmax, 2, spec, probe)
    tp = time.perf_counter() - t
    per_triple_single = tp / ((L - probe[0]) * nring)
    n = max(cores, int(0.6 * target_seconds * cores / max(tp, 1e-3)))
    n = min(n, L)
    orders = np.unique(np.linspace(0, L - 1, n).astype(np.int32))
    t = time.perf_counter()
    lat_ana_s(nside, lmax, 2, spec, orders)
    tl = time.perf_counter() - t
    triples = float(np.sum(L - orders)) * nring
    t_triple = tl / triples
    tri = L * (L + 1) / 2
    pipeline_triples = 7 * tri * 2 * nside + 7 * tri * nring
    t_leg = t_triple * pipeline_triples
    log(f"cpu baseline: legendre sample {len(orders)} orders in {tl:.1f}s on {cores} threads "
        f"({t_triple * 1e9:.2f} ns/triple; single-order probe {per_triple_single * 1e9:.2f} ns/triple)")
    # coupling rows
    window = 1.0 / (1.0 + np.arange(L)) ** 2
    probe_rows = np.array([L // 2], dtype=np.int32)
    t = time.perf_counter()
    coupling_rows(lmax, window, probe_rows)
    tpc = time.perf_counter() - t
    nr = max(cores, int(0.3 * target_seconds * cores / max(tpc, 1e-3)))
    rows = np.unique(np.linspace(0, L - 1, min(nr, L)).astype(np.int32))
    t = time.perf_counter()
    coupling_rows(lmax, window, rows)
    tc = time.perf_counter() - t
    r = rows[:, None].astype(np.int64)
    l2 = np.arange(L, dtype=np.int64)[None, :]
    cnt = np.minimum(r + l2, lmax) - (l2 - r) + 1
    sample_terms = float(np.sum(np.where(l2 >= r, np.maximum(cnt, 0), 0)))
    t_cpl = tc / sample_terms * n3j(lmax)
    log(f"cpu baseline: coupling sample {len(rows)} rows in {tc:.1f}s")
    return {
        "value": t_leg + t_cpl,
        "unit": "s",
        "cores": cores,
        "kind": "port",
        "sample": (f"C/OpenMP FP64 oracle at Nside {nside}: spin-2 Legendre analysis of {len(orders)} "
                   f"uniformly spaced orders ({tl:.1f}s) and 3j coupling of {len(rows)} rows ({tc:.1f}s), "
                   f"extrapolated by triple / term counts to the pipeline's 14 Legendre passes "
                   f"({t_leg:.0f}s) plus the coupling pass ({t_cpl:.0f}s); FFT, spectra and LU excluded"),
        "legendre_s": t_leg,
        "coupling_s": t_cpl,
        "ns_per_triple": t_triple * 1e9,
    }


def ladder(nsides=(64, 256), log=print):
    """Unsampled CPU timings at small Nside: one complete spin-2 Legendre analysis pass (every
    order, every ring) and every row of the 3j coupling, scaled to the pipeline's 14 passes by
    triple count as ``estimate_pipeline`` does.  Returns a list of dicts (seconds per pipeline)."""
    cores = len(os.sched_getaffinity(0))
    os.environ["OMP_NUM_THREADS"] = str(cores)
    out = []
    for nside in nsides:
        lmax = 3 * nside - 1
        L = lmax + 1
        nring = 4 * nside - 1
        rng = np.random.default_rng(nside)
        W = 2 * L - 1
        spec = (rng.standard_normal((nring, W)) + 1j * rng.standard_normal((nring, W))) * 1e-3
        orders = np.arange(L, dtype=np.int32)
        t = time.perf_counter()
        lat_ana_s(nside, lmax, 2, spec, orders)
        tl = time.perf_counter() - t
        tri = L * (L + 1) / 2
        t_leg = tl / (tri * nring) * (7 * tri * 2 * nside + 7 * tri * nring)
        window = 1.0 / (1.0 + np.arange(L)) ** 2
        t = time.perf_counter()
        coupling_rows(lmax, window, orders)
        tc = time.perf_counter() - t
        out.append({"nside": nside, "lmax": lmax, "seconds": t_leg + tc, "legendre_pass_s": tl,
                    "legendre_s": t_leg, "coupling_s": tc, "cores": cores})
        log(f"cpu ladder: Nside {nside}: full spin-2 pass {tl:.2f}s, full coupling {tc:.2f}s "
            f"-> {t_leg + tc:.1f}s per pipeline")
    return out

