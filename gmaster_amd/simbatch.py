"""Many T+P realisations of mixed Nside through the batched transforms, as one HIP graph.

A job holds ``counts[i]`` realisations (T, Q, U maps) at ``nsides[i]``.  Per Nside the T maps
run through the transforms in batches of spin-0 maps and the Q/U pairs in batches of
spin-2 fields, each batch sharing one spin-0 recurrence (all 16 analysis channels are in
use at 8 T maps or 2 Q/U pairs).  Every batch is ``map2alm`` with ``n_iter`` Jacobi steps
into the job's alm buffers, then ``alm2map`` of those alms.  ``capture`` records the whole
job -- every batch of every Nside -- into one HIP graph (``torch.cuda.CUDAGraph`` over the
library's launches on the capturing stream) and ``replay`` runs it as one launch.  The
graph reads the job's input buffers, so maps copied in between replays are transformed
afresh.

Batch sizes per Nside are chosen by timing (``tune_batches``): each candidate number of T
maps per batch (8, 4, 2) and of Q/U pairs per batch (2, 1) is timed on one batch of that
size (map2alm + alm2map) and the fastest per map is kept, cached per (device, Nside,
n_iter).  A remainder takes one batch of the smallest size that holds it, padded with
zero maps.

Reference: the batched march entry points and their driver (``gmaster/_native/cuda/
march_v2.cu:551-556``, ``gmaster/_sht/healpix.py:1269-1341``), which pair two maps per
recurrence; the job structure (many maps of uneven Nside in one launch) is this package's.
"""

import numpy as np
import torch

from . import _lib, batched, workload

T_SIZES = (8, 4, 2)  # spin-0 maps per batch (supported batch shapes (8, 0), (4, 0), (2, 0))
P_SIZES = (2, 1)  # spin-2 fields per batch ((0, 2), (0, 1))
_TUNED = {}


def _time_batch(nside, lmax, n_iter, k0, k2, dev, reps=2):
    """Seconds per map of one (k0 spin-0, k2 spin-2) batch: map2alm + alm2map."""
    npix = 12 * nside * nside
    g = torch.Generator(dev).manual_seed(17)
    m0 = torch.randn((k0, npix), dtype=torch.float64, device=dev, generator=g) if k0 else None
    m2 = torch.randn((k2, 2, npix), dtype=torch.float64, device=dev, generator=g) if k2 else None

    def once():
        a0, a2 = batched.map2alm(m0, m2, lmax, n_iter=n_iter)
        batched.alm2map(a0, a2, nside, lmax)

    once()
    torch.cuda.synchronize(dev)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for a, b in ev:
        a.record()
        once()
        b.record()
    torch.cuda.synchronize(dev)
    return min(a.elapsed_time(b) for a, b in ev) * 1e-3 / (k0 + k2)


def tune_batches(nside, n_iter=3, device="cuda", log=None):
    """``(t_size, p_size)``: the T maps and Q/U pairs per batch that run fastest per map at
    this Nside, timed once and cached per (device, Nside, n_iter)."""
    dev = torch.device(device)
    key = (dev.index if dev.index is not None else torch.cuda.current_device(), int(nside), int(n_iter))
    if key in _TUNED:
        return _TUNED[key]
    lmax = 3 * int(nside) - 1
    tt = {k: _time_batch(nside, lmax, n_iter, k, 0, dev) for k in T_SIZES}
    tp = {k: _time_batch(nside, lmax, n_iter, 0, k, dev) for k in P_SIZES}
    best = (min(tt, key=tt.get), min(tp, key=tp.get))
    if log is not None:
        log(f"tuned batches at Nside {nside}: T {best[0]} ({', '.join(f'{k}: {v * 1e3:.2f}' for k, v in tt.items())} "
            f"ms/map), Q/U {best[1]} ({', '.join(f'{k}: {v * 1e3:.2f}' for k, v in tp.items())} ms/map)")
    _TUNED[key] = best
    return best


def _split(n, sizes):
    """Batch sizes covering n items: as many of the largest of ``sizes`` as fit, then one
    batch of the smallest size that holds the remainder (padded; a batch costs about the
    same at any size below the largest, so one padded batch beats several small ones)."""
    big = max(sizes)
    out = [big] * (n // big)
    r = n - big * len(out)
    if r:
        out.append(min(s for s in sizes if s >= r))
    return out


def _cov_root(lmax):
    """Per-l root of the T, E, B covariance of the benchmark spectra, ``(L, 3, 3)``."""
    tt, ee, bb, te = workload.theory_cls(lmax)
    L = lmax + 1
    cov = np.zeros((L, 3, 3))
    cov[:, 0, 0], cov[:, 1, 1], cov[:, 2, 2] = tt[:L], ee[:L], bb[:L]
    cov[:, 0, 1] = cov[:, 1, 0] = te[:L]
    w, v = np.linalg.eigh(cov)
    return np.ascontiguousarray(v * np.sqrt(np.maximum(w, 0.0))[:, None, :])


class SimJob:
    """``counts[i]`` T+P realisations at ``nsides[i]`` (lmax = 3 Nside - 1), transformed with
    ``n_iter`` Jacobi steps.  Buffers per Nside: ``maps_t`` (n (+1 zero map), npix),
    ``maps_p`` (n, 2, npix) inputs; ``alm_t`` (n (+1), nelem), ``alm_p`` (n, 2, nelem)
    (E, B); the alm2map outputs in ``out_t`` / ``out_p`` (every map with ``keep_maps``,
    else one batch's worth, overwritten batch by batch)."""

    def __init__(self, nsides, counts, n_iter=3, device="cuda", sizes=None, keep_maps=False, log=None):
        if len(nsides) != len(counts):
            raise ValueError("one count per Nside")
        self.device = torch.device(device)
        self.n_iter = int(n_iter)
        self.keep_maps = bool(keep_maps)
        self.graph = None
        self.parts = []
        self.batches = []
        for nside, n in zip(nsides, counts):
            nside, n = int(nside), int(n)
            if n <= 0:
                continue
            lmax = 3 * nside - 1
            t_size, p_size = sizes[nside] if sizes is not None and nside in sizes else \
                tune_batches(nside, self.n_iter, self.device, log)
            tb = _split(n, [s for s in T_SIZES if s <= t_size])
            pb = _split(n, [s for s in P_SIZES if s <= p_size])
            pad = sum(tb) - n
            npix, nelem = 12 * nside * nside, (lmax + 1) * (lmax + 2) // 2
            f64, c128 = dict(dtype=torch.float64, device=self.device), dict(dtype=torch.complex128, device=self.device)
            part = {
                "nside": nside, "lmax": lmax, "n": n,
                "maps_t": torch.zeros((n + pad, npix), **f64),
                "maps_p": torch.zeros((n, 2, npix), **f64),
                "alm_t": torch.zeros((n + pad, nelem), **c128),
                "alm_p": torch.zeros((n, 2, nelem), **c128),
                "out_t": torch.zeros((n + pad if keep_maps else max(tb), npix), **f64),
                "out_p": torch.zeros((n if keep_maps else max(pb), 2, npix), **f64),
            }
            self.parts.append(part)
            start = 0
            for k in tb:
                self.batches.append((part, "t", start, k))
                start += k
            start = 0
            for k in pb:
                self.batches.append((part, "p", start, k))
                start += k

    def fill(self, seed=0):
        """Draw every realisation on the device (the benchmark spectra; seeds seed, seed + 1,
        ... in job order) and make its maps with the job's batches of alm2map."""
        st = _lib.stream_ptr(self.device)
        s = int(seed)
        for part in self.parts:
            lmax, n = part["lmax"], part["n"]
            root = torch.as_tensor(_cov_root(lmax), device=self.device)
            draw = torch.empty((3, part["alm_t"].shape[1]), dtype=torch.complex128, device=self.device)
            for i in range(n):
                _lib.call("gm_gaussian_alms", _lib.ptr(root), lmax, 3, s & 0xFFFFFFFFFFFFFFFF, _lib.ptr(draw), st)
                part["alm_t"][i].copy_(draw[0])
                part["alm_p"][i].copy_(draw[1:3])
                s += 1
            part["alm_t"][n:].zero_()  # the padding map
        for part, kind, start, k in self.batches:
            nside, lmax = part["nside"], part["lmax"]
            if kind == "t":
                batched.alm2map(part["alm_t"][start:start + k], None, nside, lmax,
                                out=(part["maps_t"][start:start + k], None))
            else:
                batched.alm2map(None, part["alm_p"][start:start + k], nside, lmax,
                                out=(None, part["maps_p"][start:start + k]))

    def run(self):
        """Every batch eagerly (or into a graph under capture): map2alm, then alm2map."""
        for part, kind, start, k in self.batches:
            nside, lmax = part["nside"], part["lmax"]
            if kind == "t":
                a0 = part["alm_t"][start:start + k]
                batched.map2alm(part["maps_t"][start:start + k], None, lmax, n_iter=self.n_iter, out=(a0, None))
                o = part["out_t"][start:start + k] if self.keep_maps else part["out_t"][:k]
                batched.alm2map(a0, None, nside, lmax, out=(o, None))
            else:
                a2 = part["alm_p"][start:start + k]
                batched.map2alm(None, part["maps_p"][start:start + k], lmax, n_iter=self.n_iter, out=(None, a2))
                o = part["out_p"][start:start + k] if self.keep_maps else part["out_p"][:k]
                batched.alm2map(None, a2, nside, lmax, out=(None, o))

    def capture(self):
        """Record the whole job into one HIP graph (after one eager run on a side stream,
        which builds the recurrence tables and sets the kernels' LDS limits)."""
        if _lib.TIMERS is not None:
            raise RuntimeError("disable the entry-point timers before capturing")
        side = torch.cuda.Stream(device=self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            self.run()
        torch.cuda.current_stream(self.device).wait_stream(side)
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run()
        self.graph = g
        return g

    def replay(self):
        if self.graph is None:
            raise RuntimeError("capture() first")
        self.graph.replay()

    @property
    def nsims(self):
        return sum(p["n"] for p in self.parts)
