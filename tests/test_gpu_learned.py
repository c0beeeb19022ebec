"""The learned coupling surrogate: training on the GPU with this package's kernels, and what
the trained network gains over the Toeplitz approximation alone on held-out masks."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _frob_err(blocks, exact, lmax):
    """Relative Frobenius error of the master coefficients, all four blocks together."""
    col = 2 * torch.arange(lmax + 1, dtype=torch.float64, device="cuda") + 1
    num = sum(float(torch.sum(((blocks[k] - exact[k]) / col) ** 2)) for k in exact)
    den = sum(float(torch.sum((exact[k] / col) ** 2)) for k in exact)
    return (num / den) ** 0.5


def test_trained_network_reduces_the_toeplitz_error_on_held_out_masks():
    from gmaster_amd import learned
    from gmaster_amd.workspaces import clear_coupling_cache, coupling_blocks

    grid = learned.ResidualGrid(l_toeplitz=96, l_exact=8, dl_band=8, l_top=383, d_top=128, n_l=12, n_d=16)
    X, Y, W = learned.make_dataset(160, nsides=(64, 128), seed=3, grid=grid)
    ntr = 128
    sur = learned.Surrogate.fit(X[:ntr], Y[:ntr], W[:ntr], grid, hidden=(256, 256, 256), steps=800, batch=64,
                                lr=2e-3, seed=1)
    hist = sur.meta["history"]
    print(f"training loss {hist[0][1]:.3e} -> {hist[-1][1]:.3e}", flush=True)
    assert hist[-1][1] < 0.2 * hist[0][1]
    # held-out grid residual (masks the network never saw)
    yh = torch.cat([sur.net.forward(((X[i:i + 1] - sur.x_mean) / sur.x_std).contiguous()) * sur.y_scale
                    for i in range(ntr, X.shape[0])])
    Wh = W[ntr:]
    base = float(torch.sqrt((Wh * Y[ntr:] ** 2).sum() / Wh.sum()))
    resid = float(torch.sqrt((Wh * (Y[ntr:] - yh) ** 2).sum() / Wh.sum()))
    print(f"held-out grid rms: Toeplitz {base:.3e}, learned {resid:.3e}", flush=True)
    assert resid < 0.5 * base
    # whole matrices on two fresh masks: the correction lowers the error of every block set
    rng = np.random.default_rng(99)
    for nside in (64, 128):
        lmax = 3 * nside - 1
        win = learned.mask_window(learned.random_mask(nside, rng), nside, lmax)
        clear_coupling_cache()
        exact = {k: v.clone() for k, v in coupling_blocks(win, lmax).items()}
        clear_coupling_cache()
        toe = {k: v.clone() for k, v in coupling_blocks(win, lmax, grid.toeplitz()).items()}
        clear_coupling_cache()
        lb, info = sur.approximate(win, lmax)
        et, el = _frob_err(toe, exact, lmax), _frob_err(lb, exact, lmax)
        print(f"nside {nside}: Frobenius error Toeplitz {et:.3e} learned {el:.3e}; spot check {info}", flush=True)
        assert el < et


def test_surrogate_round_trips_through_safetensors(tmp_path):
    from gmaster_amd import learned
    from gmaster_amd.mlp import MLP

    grid = learned.ResidualGrid(l_toeplitz=96, l_exact=8, dl_band=8, l_top=383, d_top=128, n_l=12, n_d=16)
    net = MLP([256, 32, grid.n_out], seed=5)
    sur = learned.Surrogate(grid, net, torch.zeros(256, device="cuda"), torch.ones(256, device="cuda"),
                            torch.full((grid.n_out,), 1e-3, device="cuda"), {"note": "test"})
    p = str(tmp_path / "s.safetensors")
    sur.save(p)
    back = learned.Surrogate.load(p)
    w = torch.rand(300, dtype=torch.float64, device="cuda") + 0.1
    assert torch.equal(sur.predict(w), back.predict(w))
    assert back.grid.to_json() == grid.to_json() and back.meta["note"] == "test"


def test_committed_surrogate_beats_toeplitz_on_fresh_masks():
    """The network trained on 4096 random masks at Nside 256-1024 (gmaster_amd/data),
    evaluated on masks drawn from a seed the training never used."""
    from gmaster_amd import learned
    from gmaster_amd.workspaces import clear_coupling_cache, coupling_blocks

    sur = learned.Surrogate.load(learned.default_path())
    print(f"surrogate meta {sur.meta}", flush=True)
    assert sur.meta["train_masks"] == 4096
    rng = np.random.default_rng(2024)
    gains = []
    for nside in (256, 512, 256, 512):
        lmax = 3 * nside - 1
        win = learned.mask_window(learned.random_mask(nside, rng), nside, lmax)
        clear_coupling_cache()
        exact = {k: v.clone() for k, v in coupling_blocks(win, lmax).items()}
        clear_coupling_cache()
        toe = {k: v.clone() for k, v in coupling_blocks(win, lmax, sur.grid.toeplitz()).items()}
        clear_coupling_cache()
        lb, _ = sur.approximate(win, lmax)
        et, el = _frob_err(toe, exact, lmax), _frob_err(lb, exact, lmax)
        print(f"nside {nside}: Frobenius error Toeplitz {et:.3e} learned {el:.3e}", flush=True)
        gains.append(et / el)
    assert min(gains) > 1.0 and float(np.median(gains)) > 2.0


def test_learned_form_workspaces_and_fallback():
    """``set_form("learned")``: workspaces take the committed surrogate's blocks (narrow
    Toeplitz + learned residual) when the exact spot check passes, and the exact coupling
    when it does not (a tolerance of 1e-12 forces the fallback); Nside 256, one random mask."""
    import gmaster_amd as nmt
    from gmaster_amd import learned, workload
    from gmaster_amd.workspaces import clear_coupling_cache, coupling_blocks

    nside = 256
    lmax = 3 * nside - 1
    rng = np.random.default_rng(31)
    mask = learned.random_mask(nside, rng)
    tqu = workload.tqu_maps(nside, lmax, seed=5)
    b = nmt.NmtBin.from_nside_linear(nside, 16)
    try:
        nmt.set_form("learned")
        f = nmt.NmtField(mask, tqu[1:3], n_iter=3)
        clear_coupling_cache()
        w = nmt.NmtWorkspace.from_fields(f, f, b)
        info = w.coupling_info
        print(f"learned form info {info}", flush=True)
        assert info["form"] == "learned"
        win = w.pcl_mask
        clear_coupling_cache()
        exact = {k: v.clone() for k, v in coupling_blocks(win, lmax).items()}
        if not info["fallback"]:
            sur = learned.committed_surrogate()
            clear_coupling_cache()
            toe = {k: v.clone() for k, v in coupling_blocks(win, lmax, sur.grid.toeplitz()).items()}
            el, et = _frob_err(w.blocks, exact, lmax), _frob_err(toe, exact, lmax)
            print(f"  Frobenius error: learned {el:.3e}, Toeplitz {et:.3e}", flush=True)
            assert el < et
        clear_coupling_cache()
        blocks, info2 = learned.learned_coupling_blocks(win, lmax, tol=1e-12)
        assert info2["fallback"]
        for k in exact:
            assert torch.equal(blocks[k], exact[k]), k
    finally:
        nmt.set_form("fp64")
        clear_coupling_cache()
