"""C-ABI checks that need no GPU: the library loads, exports every symbol the header
declares, and rejects bad arguments before launching anything."""

import ctypes
import os
import re

import pytest

from gmaster_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_symbols():
    text = open(os.path.join(ROOT, "include", "gmaster_amd.h")).read()
    decl = r"^\s*(?:const\s+char\s*\*|int64_t|int)\s+(gm_[a-z0-9_]+)\s*\("
    return sorted(set(re.findall(decl, text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = _header_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    # and the Python binding declares every one of them
    assert set(names) <= set(_lib.SIGNATURES)


def test_abi_version():
    assert _lib.query("gm_abi_version") == 2


def test_noalloc_scratch_limit():
    """The no-allocation coupled-C_ell entry holds 2^23 doubles (64 MiB) of partial sums:
    enough for lmax 8191 with 2 x 2 components, not for lmax 8192."""
    assert _lib.query("gm_coupled_cell_noalloc_max_work") == 1 << 23
    assert _lib.query("gm_cl_work_size", 8191, 4) <= 1 << 23 < _lib.query("gm_cl_work_size", 8192, 4)
    assert _lib.query("gm_cl_work_size", 16383, 1) <= 1 << 23


def test_table_size_matches_layout():
    for lmax in (0, 5, 191, 12287):
        L = lmax + 1
        expect = L * (L + 8) - L * (L - 1) // 2
        assert _lib.query("gm_table_size", lmax) == expect


def test_nchunks():
    assert _lib.query("gm_lat_nchunks", 64, 0) == 1
    assert _lib.query("gm_lat_nchunks", 4096, 0) == 6
    assert _lib.query("gm_lat_nchunks", 4096, 2) == 8
    # one analysis pass over every ring chunk up to Nside 4096 ...
    for spin in (0, 2):
        for mixed in (0, 1):
            slots = _lib.query("gm_lat_nchunks_mixed" if mixed else "gm_lat_nchunks", *((4096, spin, 1) if mixed else (4096, spin)))
            assert _lib.query("gm_lat_ring_chunks", 4096, spin, mixed, 1) == slots
    # ... and passes of 4 slots above (Nside 8192: 32767 rings in chunks of 2048 / 1536)
    assert _lib.query("gm_lat_ring_chunks", 8192, 2, 0, 1) == 16
    assert _lib.query("gm_lat_nchunks", 8192, 2) == 4
    assert _lib.query("gm_lat_ring_chunks", 8192, 2, 1, 1) == 22
    assert _lib.query("gm_lat_nchunks_mixed", 8192, 2, 1) == 4
    assert _lib.query("gm_lat_ring_chunks", 8192, 0, 0, 2) == 11
    assert _lib.query("gm_lat_nchunks", 8192, 0) == 4


@pytest.fixture(scope="module")
def dummy_pointer():
    import torch

    if not torch.cuda.is_available():
        yield 1  # never dereferenced: no launch can happen without a device
        return
    buf = torch.zeros(1 << 23, dtype=torch.float64, device="cuda")  # 64 MiB
    yield buf.data_ptr()
    torch.cuda.synchronize()
    del buf


@pytest.mark.parametrize(
    "name,args,code",
    [
        ("gm_ring_fft_fwd_s0", (None, 1, 12, 3, 2, None, 3, 0, None), 1),  # nside not a power of 2
        ("gm_ring_fft_fwd_s0", (1, 1, 768, 8, 10, 1, 5, 0, None), 2),     # ld < lmax + 1
        ("gm_ring_fft_fwd_s0", (1, 1, 0, 16384, 10, 1, 11, 0, None), 3),  # nside above the FFT limit
        ("gm_ring_fft_fwd_s", (1, 1, 8, 10, 1, 20, None), 2),            # ld < 2 lmax + 1
        ("gm_lat_ana_s0", (1, 3, 11, 0, 8, 10, 1, 1, 1, 1, 1, 1, 0, None), 1),  # nmaps = 3
        ("gm_lat_ana_s0", (1, 1, 11, 0, 8, 10, 1, 1, 1, 1, 1, 7, 0, None), 2),  # wrong chunk count
        ("gm_lat_ana_s0", (1, 1, 11, 0, 8, 10, 1, 1, 1, 1, 1, 1, 1, None), 1),  # chunk0 past the last chunk
        ("gm_lat_ana_s", (1, 21, 8, 10, 0, 1, 1, 1, 1, 1, 1, 1, 0, None), 1),   # spin 0 on the spin-s entry
        ("gm_lat_finish_s0", (1, 1, 8, 1, 1, 10, 1, 66, 3, 0, None), 1),       # mode 3
        ("gm_lat_tables", (10, 11, 1, 1, 1, None), 1),                      # spin > lmax
        ("gm_coupling", (None, 10, 1, 10, 1, None, None, None, None), 1),   # null window
        ("gm_coupling", (1, 21, 1, 10, 1, None, None, None, None), 1),      # window beyond 2 lmax
        ("gm_coupled_cell", (1, 3, 1, 1, 66, 10, 1, 1, None), 1),           # 3 components
        ("gm_coupled_cell_noalloc", (None, 1, 1, 1, 66, 10, 1, None), 1),   # null alm
        ("gm_coupled_cell_noalloc", (1, 2, 1, 2, 1 << 26, 9000, 1, None), 2),  # partials beyond the 64 MiB array
        ("gm_coupling_pairs", (1, 20, 1, 10, None, None, 5, 1, None, None, None, None), 1),  # null pair list
        ("gm_coupling_pairs", (1, 20, 1, 10, 1, 1, -1, 1, None, None, None, None), 1),      # negative count
        ("gm_coupling_rows", (1, 20, 1, 10, None, 3, 1, None, None, None, None), 1),   # null item list
        ("gm_coupling_rows", (1, 20, 1, 10, 1, -1, 1, None, None, None, None), 1),     # negative count
        ("gm_coupling_rows", (1, 21, 1, 10, 1, 1, 1, None, None, None, None), 1),      # window beyond 2 lmax
        ("gm_coupling_rows", (1, 20, 1, 10, 1, 1, None, None, None, None, None), 1),   # no output
        ("gm_lu_factor", (None, 10, 1, None), 1),                           # null matrix
        ("gm_lu_factor", (1, 0, 1, None), 1),                               # n = 0
        ("gm_lu_solve", (1, None, 10, 1, 1, None), 1),                      # null transpose
        ("gm_lu_solve", (1, 1, 10, 1, 0, None), 1),                         # no right-hand side
        ("gm_toeplitz", (1, 10, 10, 3, 2, 2, None), 1),                     # l_toeplitz = lmax
        ("gm_toeplitz", (1, 10, 6, 7, 2, 2, None), 1),                      # l_exact > l_toeplitz
        ("gm_toeplitz", (1, 10, 6, 3, 2, 1, None), 1),                      # out aliases M
        ("gm_gauss_legendre", (0, 1, 1, None), 1),                          # order 0
        ("gm_wigner_d_table", (1, 10, 17, 0, 5, None, 1, None), 1),         # |m| > 16
        ("gm_wigner_d_corr", (1, 10, 0, 0, -1, 1, 1, None, None), 1),       # negative lmax
        ("gm_table_combine", (9, 1, None, None, 1, None, 4, 4, 1, None), 1),  # more than 8 terms
        ("gm_parity_split", (1, 1, 0, 1, 1, None), 1),                      # L = 0
        ("gm_mask_apodize", (1, 12, 0.1, 1, 1, 1, None), 1),                # nside not a power of 2
        ("gm_mask_apodize", (1, 16, 0.1, 4, 1, 1, None), 1),                # unknown kind
        ("gm_mask_apodize", (1, 16, 4.0, 1, 1, 1, None), 1),                # scale beyond pi
        ("gm_gaussian_alms", (1, 10, 9, 1, 1, None), 1),                    # 9 components
        ("gm_gauss_draws", (1, 0, 1, 1, None), 1),                          # no elements
    ],
)
def test_argument_validation(name, args, code, dummy_pointer):
    # pointer arguments that are not under test point at a real (zeroed) device buffer when a
    # GPU is visible, so a regressed check launches on valid memory instead of address 0x1
    conv = [ctypes.c_void_p(dummy_pointer) if (a is not None and isinstance(a, int) and t is ctypes.c_void_p) else a
            for a, t in zip(args, _lib.SIGNATURES[name])]
    lib = _lib.load()
    got = getattr(lib, name)(*conv)
    assert got == code, lib.gm_last_error().decode()
    assert lib.gm_last_error().decode()
    with pytest.raises(_lib.GmError):
        _lib.call(name, *conv)
