"""The no-allocation coupled-C_ell entry (gm_coupled_cell_noalloc, partial sums in the
library's own device array) against gm_coupled_cell with a caller workspace, through the
C-ABI: two calls in a row, the second with a smaller input, bit for bit (both run the same
kernels in the same order)."""

import pytest
import torch

from gmaster_amd import _lib

pytestmark = pytest.mark.gpu


def _alms(ncomp, lmax, seed):
    L = lmax + 1
    nelem = L * (L + 1) // 2
    g = torch.Generator("cuda").manual_seed(seed)
    a = torch.complex(torch.randn((ncomp, nelem), dtype=torch.float64, device="cuda", generator=g),
                      torch.randn((ncomp, nelem), dtype=torch.float64, device="cuda", generator=g))
    return a, nelem


def _both(lmax, n1, n2, seed):
    a, nelem = _alms(n1, lmax, seed)
    b, _ = _alms(n2, lmax, seed + 1)
    L = lmax + 1
    st = _lib.stream_ptr("cuda")
    ref = torch.full((n1 * n2, L), float("nan"), dtype=torch.float64, device="cuda")
    work = torch.empty(_lib.query("gm_cl_work_size", lmax, n1 * n2), dtype=torch.float64, device="cuda")
    _lib.call("gm_coupled_cell", _lib.ptr(a), n1, _lib.ptr(b), n2, nelem, lmax, _lib.ptr(ref), _lib.ptr(work), st)
    got = torch.full((n1 * n2, L), float("nan"), dtype=torch.float64, device="cuda")
    _lib.call("gm_coupled_cell_noalloc", _lib.ptr(a), n1, _lib.ptr(b), n2, nelem, lmax, _lib.ptr(got), st)
    torch.cuda.synchronize()
    return got, ref


def test_noalloc_coupled_cell_matches_workspace_entry():
    print("no-alloc coupled cell: lmax 3071 2x2, then lmax 191 1x2", flush=True)
    got, ref = _both(3071, 2, 2, 11)
    assert torch.isfinite(ref).all()
    assert torch.equal(got, ref)
    # a second, smaller call reuses the same array (larger stale partials stay behind)
    got, ref = _both(191, 1, 2, 12)
    assert torch.isfinite(ref).all()
    assert torch.equal(got, ref)


def test_noalloc_rejects_oversized_request():
    a, nelem = _alms(1, 15, 3)
    out = torch.empty((4, 8193), dtype=torch.float64, device="cuda")
    with pytest.raises(_lib.GmError, match="maximum"):
        _lib.call("gm_coupled_cell_noalloc", _lib.ptr(a), 2, _lib.ptr(a), 2, 1 << 26, 8192, _lib.ptr(out),
                  _lib.stream_ptr("cuda"))
