"""GPU parity of the dense-network kernels (bf16 MFMA, FP32 accumulation) against the
oracle with the same bf16 rounding points, and a check that training lowers the loss."""

import numpy as np
import pytest
import torch

from oracle import mlp as omlp

pytestmark = pytest.mark.gpu


def _bits_to_f32(t):
    return (t.cpu().numpy().view(np.uint16).astype(np.uint32) << 16).view(np.float32)


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


@pytest.mark.parametrize("B,n_in,n_out,act", [(37, 70, 45, 1), (64, 64, 64, 0), (130, 257, 96, 1), (1, 16, 1, 0)])
def test_layer_forward_and_backward_match_oracle(B, n_in, n_out, act):
    from gmaster_amd import _lib
    from gmaster_amd.mlp import _bf16

    print(f"layer B={B} in={n_in} out={n_out} act={act}", flush=True)
    rng = np.random.default_rng(B + n_in)
    x = torch.as_tensor(rng.standard_normal((B, n_in)), dtype=torch.float32, device="cuda")
    w = torch.as_tensor(rng.standard_normal((n_out, n_in)) / np.sqrt(n_in), dtype=torch.float32, device="cuda")
    b = torch.as_tensor(rng.standard_normal(n_out), dtype=torch.float32, device="cuda")
    xb, wb = _bf16(x), _bf16(w)
    z = torch.empty((B, n_out), dtype=torch.float32, device="cuda")
    y = torch.empty_like(z)
    st = _lib.stream_ptr("cuda")
    _lib.call("gm_mlp_fwd", _lib.ptr(xb), _lib.ptr(wb), _lib.ptr(b), B, n_in, n_out, act, _lib.ptr(z), _lib.ptr(y),
              None, st)
    xn, wn = x.cpu().numpy(), w.cpu().numpy()
    np.testing.assert_array_equal(_bits_to_f32(xb), omlp.bf16_round(xn))
    zr, yr = omlp.layer_forward(xn, wn, b.cpu().numpy(), act)
    assert _rel(z.cpu(), zr) < 2e-6
    assert _rel(y.cpu(), yr) < 2e-6

    dy = torch.as_tensor(rng.standard_normal((B, n_out)), dtype=torch.float32, device="cuda")
    dz = torch.empty_like(z)
    dzb = torch.empty((B, n_out), dtype=torch.int16, device="cuda")
    dw = torch.empty((n_out, n_in), dtype=torch.float32, device="cuda")
    db = torch.empty(n_out, dtype=torch.float32, device="cuda")
    dx = torch.empty((B, n_in), dtype=torch.float32, device="cuda")
    _lib.call("gm_mlp_bwd", _lib.ptr(dy), _lib.ptr(z), act, _lib.ptr(xb), _lib.ptr(wb), B, n_in, n_out, _lib.ptr(dz),
              _lib.ptr(dzb), _lib.ptr(dw), _lib.ptr(db), _lib.ptr(dx), None, st)
    # the oracle backward from the GPU's own z (bf16 rounding of dz is then identical)
    dwr, dbr, dxr = omlp.layer_backward(dy.cpu().numpy(), z.cpu().numpy().astype(np.float64), act, xn, wn)
    print(f"  dw {_rel(dw.cpu(), dwr):.1e} db {_rel(db.cpu(), dbr):.1e} dx {_rel(dx.cpu(), dxr):.1e}", flush=True)
    assert _rel(dw.cpu(), dwr) < 1e-5
    assert _rel(db.cpu(), dbr) < 1e-5
    assert _rel(dx.cpu(), dxr) < 1e-5


def test_network_gradients_match_oracle():
    from gmaster_amd.mlp import MLP

    net = MLP([48, 64, 64, 64, 20], seed=3)
    rng = np.random.default_rng(4)
    x = torch.as_tensor(rng.standard_normal((33, 48)), dtype=torch.float32, device="cuda")
    t = torch.as_tensor(rng.standard_normal((33, 20)), dtype=torch.float32, device="cuda")
    y, cache = net.forward(x, keep=True)
    dy, _ = net.mse_grad(y, t)
    dws, dbs = net.backward(cache, dy)
    params = [(w.cpu().numpy(), b.cpu().numpy()) for w, b in zip(net.w, net.b)]
    yr, cr = omlp.forward(params, x.cpu().numpy(), net.acts)
    assert _rel(y.cpu(), yr) < 2e-3  # bf16 re-rounding of hidden activations may differ by one ulp
    gr, _ = omlp.backward(params, cr, net.acts, 2 * (yr - t.cpu().numpy()) / t.numel())
    for k in range(len(params)):
        e_w, e_b = _rel(dws[k].cpu(), gr[k][0]), _rel(dbs[k].cpu(), gr[k][1])
        print(f"layer {k}: dw rel {e_w:.1e} db rel {e_b:.1e}", flush=True)
        assert e_w < 2e-2 and e_b < 2e-2


def test_adamw_matches_oracle():
    from gmaster_amd import _lib

    rng = np.random.default_rng(5)
    n = 1000
    w0 = rng.standard_normal(n).astype(np.float32)
    w = torch.as_tensor(w0, device="cuda")
    m = torch.zeros_like(w)
    v = torch.zeros_like(w)
    wr, mr, vr = w0.astype(np.float64), np.zeros(n), np.zeros(n)
    for step in range(1, 6):
        g = rng.standard_normal(n).astype(np.float32)
        gt = torch.as_tensor(g, device="cuda")
        _lib.call("gm_adamw", _lib.ptr(w), None, _lib.ptr(gt), _lib.ptr(m), _lib.ptr(v), n, 1e-2, 0.9, 0.999, 1e-8,
                  0.1, step, _lib.stream_ptr("cuda"))
        wr, mr, vr = omlp.adamw(wr, g.astype(np.float64), mr, vr, step, 1e-2, wd=0.1)
    assert _rel(w.cpu(), wr) < 1e-5


def test_training_lowers_the_loss():
    from gmaster_amd.mlp import MLP

    rng = np.random.default_rng(6)
    x = torch.as_tensor(rng.uniform(-1, 1, (256, 8)), dtype=torch.float32, device="cuda")
    t = torch.stack([torch.sin(3 * x[:, 0]) * x[:, 1], x[:, 2] ** 2 - x[:, 3]], dim=1).contiguous()
    net = MLP([8, 64, 64, 2], seed=7)
    first = float(net.train_step(x, t, lr=3e-3).mean())
    for _ in range(400):
        sq = net.train_step(x, t, lr=3e-3)
    last = float(sq.mean())
    print(f"mse {first:.3e} -> {last:.3e}", flush=True)
    assert last < 0.05 * first
