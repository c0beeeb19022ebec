"""The decoupling LU (gm_lu_factor / gm_lu_solve through ``workspaces.lu_factor`` /
``lu_solve``) against torch's pivoted FP64 solve: near-identity matrices of sizes that are
and are not multiples of the 64-column panel, one and many right-hand sides; the pivot flag
for matrices on which partial pivoting exchanges rows (then the pivoted path must be taken
and agree); and binned coupling matrices of a real mask workspace.  The flag's condition
(|L_ij| > 1 or a zero pivot) is checked against an explicit partial-pivoting elimination on
the CPU."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _mat(n, seed, off=0.3):
    g = torch.Generator().manual_seed(seed)
    a = torch.rand((n, n), generator=g, dtype=torch.float64) * (off / n)
    a += torch.eye(n, dtype=torch.float64) * (1.0 + torch.rand(n, generator=g, dtype=torch.float64))
    return a


def _pp_swaps(a):
    """Row exchanges partial pivoting makes (first maximum on ties), on the CPU."""
    a = a.clone().numpy()
    n = len(a)
    swaps = 0
    for j in range(n):
        p = j + int(np.argmax(np.abs(a[j:, j])))
        if p != j:
            swaps += 1
            a[[j, p]] = a[[p, j]]
        a[j + 1:, j] /= a[j, j]
        a[j + 1:, j + 1:] -= np.outer(a[j + 1:, j], a[j, j + 1:])
    return swaps


@pytest.mark.parametrize("n", [1, 5, 64, 100, 129, 768, 1000, 3072])
@pytest.mark.parametrize("k", [1, 3, 7, 200])
def test_lu_solve_matches_torch(n, k):
    from gmaster_amd.workspaces import lu_factor, lu_solve

    print(f"lu n={n} k={k}", flush=True)
    a = _mat(n, n).cuda()
    b = torch.randn((n, k), dtype=torch.float64, generator=torch.Generator().manual_seed(k)).cuda()
    fac = lu_factor(a)
    assert fac[0] == "plain"
    x = lu_solve(fac, b)
    ref = torch.linalg.solve(a, b)
    err = float((x - ref).abs().max() / ref.abs().max())
    print(f"  rel={err:.2e}", flush=True)
    assert err < 1e-13
    # the factors themselves: L U == A
    L = torch.tril(fac[1], -1) + torch.eye(n, dtype=torch.float64, device="cuda")
    U = torch.triu(fac[1])
    assert float((L @ U - a).abs().max()) < 1e-13 * n


@pytest.mark.parametrize("n", [7, 130, 500])
def test_pivot_flag_falls_back(n):
    from gmaster_amd.workspaces import lu_factor, lu_solve

    a = _mat(n, 3 * n)
    # a later row with a larger entry in an early column forces an exchange
    a[n - 1, 1] = 5.0
    assert _pp_swaps(a) > 0
    fac = lu_factor(a.cuda())
    assert fac[0] == "pivoted"
    b = torch.randn((n, 2), dtype=torch.float64).cuda()
    x = lu_solve(fac, b)
    ref = torch.linalg.solve(a.cuda(), b)
    assert float((x - ref).abs().max() / ref.abs().max()) < 1e-12


@pytest.mark.parametrize("n", [50, 300])
def test_flag_clear_exactly_when_no_exchange(n):
    from gmaster_amd.workspaces import lu_factor

    a = _mat(n, 7 * n, off=0.9)
    assert _pp_swaps(a) == 0
    assert lu_factor(a.cuda())[0] == "plain"


def test_workspace_decouple_uses_plain_lu():
    import gmaster_amd as nmt
    from gmaster_amd import workload
    from gmaster_amd.workspaces import clear_coupling_cache

    clear_coupling_cache()
    nside = 64
    mask = workload.hole_mask(nside, nholes=20, radius_deg=5.0, apo_deg=2.0, kind="C1", seed=1, device="cuda")
    rng = np.random.default_rng(0)
    maps = torch.as_tensor(rng.normal(size=(3, 12 * nside * nside)), device="cuda")
    f0 = nmt.NmtField(mask, maps[0:1], n_iter=1)
    f2 = nmt.NmtField(mask, maps[1:3], n_iter=1)
    b = nmt.NmtBin.from_nside_linear(nside, 4)
    for fa, fb in ((f0, f0), (f0, f2), (f2, f2)):
        w = nmt.NmtWorkspace.from_fields(fa, fb, b)
        cl = nmt.compute_coupled_cell(fa, fb)
        bp = torch.as_tensor(np.asarray(w.decouple_cell(cl)), device="cuda")
        assert w._lu[0] == "plain"
        binned = b.bin_cell(torch.as_tensor(np.asarray(cl), device="cuda")).as_subclass(torch.Tensor)
        ref = torch.linalg.solve(w.mcm_binned, binned.T.reshape(-1)).reshape(-1, w.ncls).T
        err = float((bp - ref).abs().max() / ref.abs().max())
        print(f"decouple ncls={w.ncls} rel={err:.2e}", flush=True)
        assert err < 1e-12
