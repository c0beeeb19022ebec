"""Learned coupling: a dense network that predicts the residual of the Toeplitz approximation.

The Toeplitz approximation (Louis et al. 2020; ``gmaster/workspaces.py:1141-1194``) reads
the correlation ``rho(l, l') = S_ll' / sqrt(S_ll S_l'l')`` of the master coefficients
``S = M / (2 l' + 1)`` from one column and extends it along diagonals.  Its error, in
correlation units,

    e(j, d) = [S(j, j + d) - S_T(j, j + d)] / sqrt(S_jj S_{j+d,j+d}),

is smooth in ``(log j, log d)``.  The network maps features of the mask pseudo-C_ell
(``log10 W_l / W_0`` at 256 multipoles up to l = 767) to ``e`` on a fixed grid of physical
multipoles for each of the four blocks (00, 02, 22+, 22-); ``gm_toeplitz_correct`` adds the
interpolated residual to every element the fill did not keep exact.  Grids, features and
Toeplitz parameters are in physical multipoles, so a network trained on masks at Nside
256-1024 applies at any Nside; above the grid the correction is zero.

Only the exact entries the fill reads are evaluated (``gm_coupling_pairs``), so the
learned form costs O(L^2) instead of the exact coupling's O(L^3).  The exact 3j kernel is
the referee: ``approximate`` evaluates a random sample of non-exact entries exactly
(``gm_coupling_pair_values``) and falls back to the full exact coupling when the
surrogate's error there exceeds the tolerance.

Training data, training and inference run on the GPU through this package's kernels
(``gmaster_amd.mlp``).
"""

import json
import math
import os

import numpy as np
import torch

from . import _lib, sht, workload
from ._array import device_tensor
from .mlp import MLP

BLOCKS = ("00", "02", "22p", "22m")
FEATURE_LMAX = 767


def feature_multipoles():
    """The 256 multipoles whose log window values feed the network."""
    head = np.arange(64)
    tail = np.round(np.geomspace(64, FEATURE_LMAX, 193)[1:]).astype(np.int64)
    ell = np.concatenate([head, tail])
    for k in range(1, len(ell)):  # strictly increasing
        ell[k] = max(ell[k], ell[k - 1] + 1)
    return ell


def _grid(lo, hi, n):
    g = np.round(np.geomspace(lo, hi, n)).astype(np.int64)
    for k in range(1, n):
        g[k] = max(g[k], g[k - 1] + 1)
    return g.astype(np.int32)


def features(window):
    """``(256,)`` FP32 features of a mask pseudo-C_ell (device tensor)."""
    w = device_tensor(window)
    ell = torch.as_tensor(feature_multipoles(), device=w.device)
    pad = torch.zeros(FEATURE_LMAX + 1, dtype=torch.float64, device=w.device)
    n = min(len(w), FEATURE_LMAX + 1)
    pad[:n] = w[:n]
    r = torch.clamp(torch.abs(pad[ell]) / torch.abs(pad[0]), min=1e-16)
    return torch.log10(r).to(torch.float32)


class ResidualGrid:
    """Toeplitz parameters and the (j, d) grid of the residual, in physical multipoles."""

    def __init__(self, l_toeplitz=384, l_exact=24, dl_band=24, l_top=3071, d_top=512, n_l=16, n_d=32):
        self.l_toeplitz, self.l_exact, self.dl_band = int(l_toeplitz), int(l_exact), int(dl_band)
        self.l_grid = _grid(self.l_exact + 1, l_top, n_l)
        self.d_grid = _grid(1, d_top, n_d)

    def to_json(self):
        return {"l_toeplitz": self.l_toeplitz, "l_exact": self.l_exact, "dl_band": self.dl_band,
                "l_grid": self.l_grid.tolist(), "d_grid": self.d_grid.tolist()}

    @classmethod
    def from_json(cls, d):
        g = cls(d["l_toeplitz"], d["l_exact"], d["dl_band"])
        g.l_grid = np.asarray(d["l_grid"], np.int32)
        g.d_grid = np.asarray(d["d_grid"], np.int32)
        return g

    @property
    def n_out(self):
        return len(BLOCKS) * len(self.l_grid) * len(self.d_grid)

    def toeplitz(self):
        return (self.l_toeplitz, self.l_exact, self.dl_band)


def _toeplitz_blocks(exact, lmax, grid):
    out = {}
    for k in BLOCKS:
        out[k] = torch.empty_like(exact[k])
        _lib.call("gm_toeplitz", _lib.ptr(exact[k]), lmax, grid.l_toeplitz, grid.l_exact, grid.dl_band,
                  _lib.ptr(out[k]), _lib.stream_ptr(exact[k].device))
    return out


def residual_targets(exact, approx, lmax, grid):
    """``(targets, weights)``, each ``(4 * n_l * n_d,)`` FP32: e(j, d) on the grid, weight 0
    where ``j + d > lmax``."""
    dev = exact["00"].device
    lg = torch.as_tensor(grid.l_grid, dtype=torch.long, device=dev)
    dg = torch.as_tensor(grid.d_grid, dtype=torch.long, device=dev)
    J = lg[:, None].expand(len(lg), len(dg))
    I = J + dg[None, :]
    ok = I <= lmax  # implies J <= lmax
    # grid points beyond this lmax get weight 0; their indices are clamped into the matrix
    Jc = torch.clamp(J, max=lmax)
    Ic = torch.clamp(I, max=lmax)
    col = (2 * torch.arange(lmax + 1, dtype=torch.float64, device=dev) + 1)
    t = []
    for k in BLOCKS:
        S = exact[k] / col[None, :]
        ST = approx[k] / col[None, :]
        D = torch.abs(torch.diagonal(S))
        den = torch.sqrt(D[Jc] * D[Ic])
        e = torch.where(ok & (den > 0), (S[Jc, Ic] - ST[Jc, Ic]) / torch.where(den > 0, den, 1.0), 0.0)
        t.append(e)
    targets = torch.stack(t).reshape(-1).to(torch.float32)
    weights = ok[None].expand(len(BLOCKS), -1, -1).reshape(-1).to(torch.float32)
    return targets, weights


def random_mask(nside, rng, device="cuda"):
    """A training mask: 10-200 holes of radius 0.1-10 deg, a |b| cut of 0-30 deg, C1 or C2
    apodization of 0.5-3 deg (uniform draws from ``rng``)."""
    nholes = int(rng.integers(10, 201))
    radii = rng.uniform(0.1, 10.0, nholes)
    bcut = float(rng.uniform(0.0, 30.0))
    apo = float(rng.uniform(0.5, 3.0))
    kind = "C1" if rng.uniform() < 0.5 else "C2"
    return workload.hole_mask(nside, nholes=nholes, radius_deg=radii, apo_deg=apo, kind=kind,
                              seed=int(rng.integers(1 << 31)), bcut_deg=bcut if bcut > 0.5 else None, device=device)


def mask_window(mask, nside, lmax, n_iter=3):
    """Mask pseudo-C_ell (``lmax + 1``) through this engine's transform and spectra kernels."""
    from .workspaces import _coupled_cell

    alm = sht.map2alm_s0(mask[None], lmax, n_iter=n_iter)
    return _coupled_cell(alm, alm, lmax)[0]


def make_dataset(n_masks, nsides=(256, 512, 1024), seed=0, grid=None, device="cuda", log=None):
    """``(X, Y, Wt)``: features, residual targets and weights of ``n_masks`` random masks
    (Nside cycling through ``nsides``), from exact couplings at lmax = 3 Nside - 1."""
    from .workspaces import clear_coupling_cache, coupling_blocks

    grid = grid or ResidualGrid()
    rng = np.random.default_rng(seed)
    X = torch.empty((n_masks, len(feature_multipoles())), dtype=torch.float32, device=device)
    Y = torch.empty((n_masks, grid.n_out), dtype=torch.float32, device=device)
    W = torch.empty_like(Y)
    for i in range(n_masks):
        nside = int(nsides[i % len(nsides)])
        lmax = 3 * nside - 1
        mask = random_mask(nside, rng, device)
        win = mask_window(mask, nside, lmax)
        clear_coupling_cache()
        exact = coupling_blocks(win, lmax)
        approx = _toeplitz_blocks(exact, lmax, grid)
        X[i] = features(win)
        Y[i], W[i] = residual_targets(exact, approx, lmax, grid)
        del exact, approx
        if log and (i + 1) % 64 == 0:
            log(f"dataset {i + 1}/{n_masks}")
    clear_coupling_cache()
    return X, Y, W


class Surrogate:
    """Trained network + grid + normalisation; ``approximate`` builds coupling blocks."""

    def __init__(self, grid, net, x_mean, x_std, y_scale, meta=None):
        self.grid, self.net = grid, net
        self.x_mean, self.x_std, self.y_scale = x_mean, x_std, y_scale
        self.meta = dict(meta or {})

    # ---- training
    @classmethod
    def fit(cls, X, Y, W, grid, hidden=(512, 512, 512), steps=3000, batch=128, lr=1e-3, weight_decay=1e-4, seed=0,
            log=None):
        dev = X.device
        x_mean, x_std = X.mean(0), X.std(0).clamp(min=1e-3)
        wsum = W.sum(0).clamp(min=1.0)
        y_scale = torch.sqrt((W * Y * Y).sum(0) / wsum).clamp(min=1e-8)
        Xn = ((X - x_mean) / x_std).contiguous()
        Yn = (Y / y_scale).contiguous()
        net = MLP([X.shape[1], *hidden, Y.shape[1]], seed=seed, device=dev)
        g = torch.Generator(device="cpu").manual_seed(seed)
        n = X.shape[0]
        hist = []
        for step in range(steps):
            idx = torch.randint(0, n, (min(batch, n),), generator=g).to(dev)
            rate = lr * 0.5 * (1 + math.cos(math.pi * step / steps))  # cosine decay
            sq = net.train_step(Xn[idx].contiguous(), Yn[idx].contiguous(), W[idx].contiguous(), lr=rate,
                                weight_decay=weight_decay)
            if step % 100 == 0 or step == steps - 1:
                loss = float(sq.sum() / W[idx].sum().clamp(min=1))
                hist.append((step, loss))
                if log:
                    log(f"step {step}: loss {loss:.4e}")
        return cls(grid, net, x_mean, x_std, y_scale, {"steps": steps, "history": hist})

    # ---- inference
    def predict(self, window):
        """``(4, n_l, n_d)`` FP32 residual on the grid for a mask pseudo-C_ell."""
        x = ((features(window).to(self.x_mean.device) - self.x_mean) / self.x_std)[None].contiguous()
        y = self.net.forward(x)[0] * self.y_scale
        return y.reshape(len(BLOCKS), len(self.grid.l_grid), len(self.grid.d_grid))

    def correct(self, blocks, window, lmax):
        """Add the predicted residual to Toeplitz blocks in place."""
        g = self.grid
        dev = blocks["00"].device
        e = self.predict(window).to(dev).contiguous()
        lg = torch.as_tensor(g.l_grid, device=dev)
        dg = torch.as_tensor(g.d_grid, device=dev)
        for b, k in enumerate(BLOCKS):
            eb = e[b].contiguous()
            _lib.call("gm_toeplitz_correct", _lib.ptr(blocks[k]), lmax, g.l_toeplitz, g.l_exact, g.dl_band,
                      _lib.ptr(lg), len(g.l_grid), _lib.ptr(dg), len(g.d_grid), _lib.ptr(eb), _lib.stream_ptr(dev))
        return blocks

    def approximate(self, window, lmax, tol=None, n_check=2048, seed=0):
        """Coupling blocks for a mask window: Toeplitz from the exact entries it reads, plus
        the learned residual; then ``n_check`` random non-exact entries are evaluated exactly.
        Returns ``(blocks, info)``; with ``tol`` set and the sampled error (correlation units)
        above it, the blocks are the exact coupling and ``info["fallback"]`` is True."""
        from .workspaces import coupling_blocks

        g = self.grid
        if g.l_toeplitz >= lmax:
            return coupling_blocks(window, lmax), {"fallback": True, "reason": "lmax below the Toeplitz column"}
        toe = coupling_blocks(window, lmax, g.toeplitz())
        blocks = {k: v.clone() for k, v in toe.items()}
        self.correct(blocks, window, lmax)
        err_t, err_l = _spot_check(window, lmax, g.toeplitz(), toe, blocks, n_check, seed)
        info = {"fallback": False, "sample_err_toeplitz": err_t, "sample_err_learned": err_l, "n_check": n_check}
        if tol is not None and err_l > tol:
            info["fallback"] = True
            from .workspaces import clear_coupling_cache

            clear_coupling_cache()
            blocks = coupling_blocks(window, lmax)
        return blocks, info

    # ---- persistence (safetensors + JSON header, nothing executable)
    def save(self, path):
        from safetensors.torch import save_file

        # the network runs on bf16 weights, so those are what is stored
        t = {f"w{i}": w.to(torch.bfloat16).cpu().contiguous() for i, w in enumerate(self.net.w)}
        t.update({f"b{i}": b.cpu().contiguous() for i, b in enumerate(self.net.b)})
        t.update({"x_mean": self.x_mean.cpu(), "x_std": self.x_std.cpu(), "y_scale": self.y_scale.cpu()})
        meta = {"sizes": json.dumps(self.net.sizes), "grid": json.dumps(self.grid.to_json()),
                "meta": json.dumps({k: v for k, v in self.meta.items() if k != "history"})}
        save_file(t, path, metadata=meta)

    @classmethod
    def load(cls, path, device="cuda"):
        from safetensors import safe_open

        with safe_open(path, framework="pt") as f:
            md = f.metadata()
            t = {k: f.get_tensor(k) for k in f.keys()}
        sizes = json.loads(md["sizes"])
        nl = len(sizes) - 1
        net = MLP.from_state({"sizes": sizes, "w": [t[f"w{i}"].float() for i in range(nl)],
                              "b": [t[f"b{i}"] for i in range(nl)]}, device=device)
        dev = torch.device(device)
        return cls(ResidualGrid.from_json(json.loads(md["grid"])), net, t["x_mean"].to(dev), t["x_std"].to(dev),
                   t["y_scale"].to(dev), json.loads(md.get("meta", "{}")))


def spot_check(window, lmax, toeplitz, blocks, n_check=4096, seed=0):
    """Max |error| in correlation units of approximate blocks over random entries outside
    the exact columns and band of ``toeplitz``, against exact 3j values."""
    return _spot_check(window, lmax, toeplitz, blocks, blocks, n_check, seed)[1]


_SAMPLES = {}


def _spot_sample(lmax, toeplitz, n_check, seed, dev):
    """The random non-exact upper-triangle pairs ``(lo, hi)`` (int32 device tensors) of a spot
    check: drawn once per (lmax, Toeplitz parameters, count, seed) and kept on the device."""
    lt, le, dl = (int(v) for v in toeplitz)
    key = (int(lmax), lt, le, dl, int(n_check), int(seed), str(dev))
    hit = _SAMPLES.get(key)
    if hit is None:
        L = lmax + 1
        rng = np.random.default_rng(seed)
        lo = rng.integers(le + 1, L - 1, 4 * n_check)
        hi = lo + rng.integers(1, L, 4 * n_check)
        keep = (hi < L) & (hi - lo > dl)
        lo, hi = lo[keep][:n_check], hi[keep][:n_check]
        hit = (torch.as_tensor(lo.astype(np.int32), device=dev), torch.as_tensor(hi.astype(np.int32), device=dev))
        if len(_SAMPLES) >= 8:
            _SAMPLES.clear()
        _SAMPLES[key] = hit
    return hit


def _spot_check(window, lmax, toeplitz, toe, learned, n_check, seed):
    """Max |error| in correlation units of the Toeplitz and learned blocks over random
    non-exact upper-triangle entries, against the exact 3j values (one readback)."""
    from .workspaces import _g_table

    dev = toe["00"].device
    L = lmax + 1
    lo_t, hi_t = _spot_sample(lmax, toeplitz, n_check, seed, dev)
    if lo_t.numel() == 0:
        return 0.0, 0.0
    w = device_tensor(window)[: 2 * lmax + 1].contiguous()
    wp = (w * (2 * torch.arange(len(w), dtype=torch.float64, device=dev) + 1) / (4 * np.pi)).contiguous()
    vals = torch.empty((lo_t.numel(), 4), dtype=torch.float64, device=dev)
    _lib.call("gm_coupling_pair_values", _lib.ptr(wp), len(w) - 1, _lib.ptr(_g_table(L + 1, dev)), lmax,
              _lib.ptr(lo_t), _lib.ptr(hi_t), lo_t.numel(), _lib.ptr(vals), _lib.stream_ptr(dev))
    col = (2 * torch.arange(L, dtype=torch.float64, device=dev) + 1)
    li, hj = lo_t.long(), hi_t.long()
    errs = []
    for b, k in enumerate(BLOCKS):
        D = torch.abs(torch.diagonal(toe[k])) / col
        den = torch.sqrt(D[li] * D[hj])
        ok = den > 0
        safe = torch.where(ok, den, torch.ones_like(den))
        ex = vals[:, b]
        at = toe[k][li, hj] / col[hj]
        al = learned[k][li, hj] / col[hj]
        zero = torch.zeros_like(den)
        errs.append(torch.stack([torch.where(ok, torch.abs(at - ex) / safe, zero).max(),
                                 torch.where(ok, torch.abs(al - ex) / safe, zero).max()]))
    e = torch.stack(errs).max(0).values.cpu()
    return float(e[0]), float(e[1])


def default_path():
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "coupling_surrogate.safetensors")


_SURROGATE = {}
_LEARNED_CACHE = {}


def committed_surrogate(device="cuda"):
    """The committed network (``default_path()``), loaded once per device."""
    dev = torch.device(device)
    key = str(dev)
    if key not in _SURROGATE:
        _SURROGATE[key] = Surrogate.load(default_path(), device=dev)
    return _SURROGATE[key]


def learned_coupling_blocks(window, lmax, tol=None, n_check=4096, seed=0):
    """The ``"learned"`` form's coupling blocks: the committed surrogate's Toeplitz
    approximation plus its predicted residual, spot-checked against exact 3j entries with
    the exact fallback above ``tol`` (default: the fast form's).  Cached per (lmax, window)
    so that the 00 / 02 / 22 workspaces of one mask share one pass."""
    import hashlib

    from ._config import nmt_params

    w = device_tensor(window)
    key = (int(lmax), str(w.device), hashlib.sha1(w.cpu().numpy().tobytes()).hexdigest())
    if key in _LEARNED_CACHE:
        return _LEARNED_CACHE[key]
    sur = committed_surrogate(w.device)
    blocks, info = sur.approximate(w, lmax, tol=nmt_params.fast_tol if tol is None else tol, n_check=n_check,
                                   seed=seed)
    info = dict(info, form="learned", toeplitz=sur.grid.toeplitz(), tol=nmt_params.fast_tol if tol is None else tol)
    if len(_LEARNED_CACHE) >= 2:
        _LEARNED_CACHE.clear()
    _LEARNED_CACHE[key] = (blocks, info)
    return blocks, info


def clear_learned_cache():
    _LEARNED_CACHE.clear()


def main(argv=None):
    """Train the surrogate on random masks: ``python -m gmaster_amd.learned --masks 4096 --out PATH``."""
    import argparse
    import sys
    import time

    ap = argparse.ArgumentParser()
    ap.add_argument("--masks", type=int, default=4096)
    ap.add_argument("--holdout", type=int, default=256)
    ap.add_argument("--nsides", default="256,512,1024")
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=default_path())
    args = ap.parse_args(argv)

    def log(msg):
        print(f"[learned] {msg}", file=sys.stderr, flush=True)

    nsides = tuple(int(v) for v in args.nsides.split(","))
    grid = ResidualGrid()
    t0 = time.time()
    X, Y, W = make_dataset(args.masks + args.holdout, nsides, seed=args.seed, grid=grid, log=log)
    log(f"dataset ready in {time.time() - t0:.1f}s")
    n = args.masks
    sur = Surrogate.fit(X[:n], Y[:n], W[:n], grid, steps=args.steps, seed=args.seed, log=log)
    yh = torch.cat([sur.net.forward(((X[n + i:n + i + 1] - sur.x_mean) / sur.x_std).contiguous()) * sur.y_scale
                    for i in range(args.holdout)])
    Wh = W[n:]
    base = float(torch.sqrt((Wh * Y[n:] ** 2).sum() / Wh.sum()))
    resid = float(torch.sqrt((Wh * (Y[n:] - yh) ** 2).sum() / Wh.sum()))
    sur.meta.update({"holdout_rms_toeplitz": base, "holdout_rms_learned": resid, "train_masks": n,
                     "nsides": list(nsides), "seed": args.seed})
    log(f"held-out rms residual: Toeplitz {base:.3e} -> learned {resid:.3e}")
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    sur.save(args.out)
    print(json.dumps(sur.meta))


if __name__ == "__main__":
    main()
