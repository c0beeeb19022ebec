"""ctypes binding of ``libgmaster_amd.so`` (the C-ABI declared in ``include/gmaster_amd.h``).

The library is built in-tree by ``gmaster_amd.build``.  There is no fallback: if the
library is missing every transform raises, naming the build command.
"""

import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
# GMASTER_AMD_LIB selects another build of the same library (e.g. one compiled with
# different tuning macros); the default is the in-tree build.
LIB_PATH = os.environ.get("GMASTER_AMD_LIB") or os.path.join(_HERE, "libgmaster_amd.so")

_i64 = ctypes.c_int64
_i32 = ctypes.c_int
_p = ctypes.c_void_p
_d = ctypes.c_double

# name -> argument types (all return int status unless listed in _RET_I64)
SIGNATURES = {
    "gm_ring_fft_fwd_s0": [_p, _i64, _i64, _i64, _i64, _p, _i64, _i64, _p],
    "gm_ring_fft_fwd_s": [_p, _p, _i64, _i64, _p, _i64, _p],
    "gm_ring_fft_inv_s0": [_p, _i64, _i64, _i64, _i64, _i64, _p, _i64, _p],
    "gm_ring_fft_inv_s": [_p, _i64, _i64, _i64, _p, _p, _p],
    "gm_ring_fold_residual": [_p, _p, _p, _i64, _i64, _i64, _i64, _i64, _p],
    "gm_ring_fold_residual_c": [_p, _p, _p, _i64, _i64, _i64, _p],
    "gm_lat_tables": [_i64, _i64, _p, _p, _p, _p],
    "gm_lat_ana_s0": [_p, _i64, _i64, _i64, _i64, _i64, _p, _p, _p, _p, _p, _i64, _i64, _p],
    "gm_lat_ana_s0_mixed": [_p, _i64, _i64, _i64, _i64, _i64, _p, _p, _p, _p, _p, _i64, _i64, _p],
    "gm_lat_finish_s0": [_p, _i64, _i64, _p, _p, _i64, _p, _i64, _i32, _i64, _p],
    "gm_lat_finish_s0_mixed": [_p, _i64, _i64, _p, _p, _i64, _p, _i64, _i32, _i64, _p],
    "gm_lat_prep_s0": [_p, _i64, _i64, _p, _i64, _p, _p],
    "gm_lat_syn_s0": [_p, _i64, _i64, _i64, _p, _p, _p, _p, _p, _i64, _i64, _p],
    "gm_lat_syn_s0_mixed": [_p, _i64, _i64, _i64, _p, _p, _p, _p, _p, _i64, _i64, _p],
    "gm_lat_ana_s": [_p, _i64, _i64, _i64, _i64, _p, _p, _p, _p, _p, _p, _i64, _i64, _p],
    "gm_lat_ana_s_mixed": [_p, _i64, _i64, _i64, _i64, _p, _p, _p, _p, _p, _p, _i64, _i64, _p],
    "gm_lat_finish_s": [_p, _i64, _p, _p, _i64, _i64, _p, _p, _i32, _i64, _p],
    "gm_lat_finish_s_mixed": [_p, _i64, _p, _p, _i64, _i64, _p, _p, _i32, _i64, _p],
    "gm_lat_prep_s": [_p, _p, _p, _i64, _i64, _p, _p],
    "gm_lat_syn_s": [_p, _i64, _i64, _i64, _p, _p, _p, _p, _p, _p, _i64, _p],
    "gm_lat_syn_s_mixed": [_p, _i64, _i64, _i64, _p, _p, _p, _p, _p, _p, _i64, _p],
    "gm_coupled_cell": [_p, _i64, _p, _i64, _i64, _i64, _p, _p, _p],
    "gm_bat_blocks": [_i64, _i64, _i64],
    "gm_bat_stride": [_i64, _i64, _i64],
    "gm_bat_ana": [_p, _i64, _i64, _i64, _p, _i64, _i64, _i64, _i64, _i64, _p, _p, _p, _p, _i64, _i64, _p, _p, _p],
    "gm_bat_ana_res": [_p, _i64, _i64, _i64, _p, _i64, _i64, _i64, _i64, _i64, _p, _p, _p, _p, _i64, _i64, _p, _p, _p],
    "gm_bat_finish": [_p, _p, _i64, _i64, _i64, _i64, _p, _i64, _i64, _p, _p, _i64, _p, _i64, _i32, _p],
    "gm_bat_prep": [_p, _i64, _i64, _p, _i64, _i64, _p, _i64, _p, _p],
    "gm_bat_syn": [_p, _i64, _i64, _i64, _i64, _p, _p, _p, _p, _i64, _p, _i64, _i64, _p, _i64, _i64, _p],
    "gm_coupled_cell_noalloc": [_p, _i64, _p, _i64, _i64, _i64, _p, _p],
    "gm_coupled_cell_noalloc_max_work": [],
    "gm_coupling": [_p, _i64, _p, _i64, _p, _p, _p, _p, _p],
    "gm_coupling_pairs": [_p, _i64, _p, _i64, _p, _p, _i64, _p, _p, _p, _p, _p],
    "gm_coupling_rows": [_p, _i64, _p, _i64, _p, _i64, _p, _p, _p, _p, _p],
    "gm_mlp_fwd": [_p, _p, _p, _i64, _i64, _i64, _i32, _p, _p, _p, _p],
    "gm_mlp_bwd": [_p, _p, _i32, _p, _p, _i64, _i64, _i64, _p, _p, _p, _p, _p, _p, _p],
    "gm_to_bf16": [_p, _i64, _p, _p],
    "gm_adamw": [_p, _p, _p, _p, _p, _i64, _d, _d, _d, _d, _d, _i64, _p],
    "gm_mse_grad": [_p, _p, _p, _i64, _d, _p, _p, _p],
    "gm_coupling_pair_values": [_p, _i64, _p, _i64, _p, _p, _i64, _p, _p],
    "gm_toeplitz_correct": [_p, _i64, _i64, _i64, _i64, _p, _i64, _p, _i64, _p, _p],
    "gm_toeplitz": [_p, _i64, _i64, _i64, _i64, _p, _p],
    "gm_lu_factor": [_p, _i64, _p, _p],
    "gm_lu_solve": [_p, _p, _i64, _p, _i64, _p],
    "gm_bin_block": [_p, _i64, _p, _p, _i64, _p, _p, _p, _p],
    "gm_bin_cell": [_p, _i64, _i64, _p, _p, _i64, _p, _p, _p],
    "gm_fingerprint": [_p, _i64, _p, _p],
    "gm_map_dots_work_size": [_i64, _i64, _i64],
    "gm_map_dots": [_p, _i64, _i64, _p, _i64, _i64, _i64, _p, _i64, _d, _p, _p, _p],
    "gm_map_deproject": [_p, _i64, _p, _i64, _i64, _p, _p],
    "gm_alm_filter": [_p, _i64, _i64, _p, _i64, _i64, _p, _i64, _p],
    "gm_spin_mul": [_p, _p, _i64, _p, _p],
    "gm_gauss_legendre": [_i64, _p, _p, _p],
    "gm_wigner_d_table": [_p, _i64, _i64, _i64, _i64, _p, _p, _p],
    "gm_wigner_d_corr": [_p, _i64, _i64, _i64, _i64, _p, _p, _p, _p],
    "gm_table_combine": [_i64, _p, _p, _p, _p, _p, _i64, _i64, _p, _p],
    "gm_parity_split": [_p, _p, _i64, _p, _p, _p],
    "gm_apodize_work_size": [_i64],
    "gm_mask_apodize": [_p, _i64, _d, _i64, _p, _p, _p],
    "gm_gaussian_alms": [_p, _i64, _i64, ctypes.c_uint64, _p, _p],
    "gm_gauss_draws": [ctypes.c_uint64, _i64, _i64, _p, _p],
    "gm_table_size": [_i64],
    "gm_lat_nchunks": [_i64, _i64],
    "gm_lat_nchunks_mixed": [_i64, _i64, _i64],
    "gm_lat_ring_chunks": [_i64, _i64, _i64, _i64],
    "gm_cl_work_size": [_i64, _i64],
    "gm_abi_version": [],
    "gm_last_error": [],
}
_RET_I64 = {"gm_bat_blocks", "gm_bat_stride", "gm_coupled_cell_noalloc_max_work", "gm_apodize_work_size", "gm_lat_ring_chunks", "gm_map_dots_work_size", "gm_lat_nchunks_mixed", "gm_table_size", "gm_lat_nchunks", "gm_cl_work_size"}

_lib = None


class GmError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, name, code, msg):
        super().__init__(f"{name} failed (code {code}): {msg}")
        self.code = code


def load():
    """Load the library (once) and declare every exported symbol."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"gmaster_amd native library not found at {LIB_PATH}; build it with "
            "`python -m gmaster_amd.build` (hipcc, gfx950)"
        )
    lib = ctypes.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        if name == "gm_last_error":
            fn.restype = ctypes.c_char_p
        elif name in _RET_I64:
            fn.restype = _i64
        else:
            fn.restype = ctypes.c_int
    _lib = lib
    return lib


# Optional per-entry-point timing with events on the launching stream (used by bench.py):
# TIMERS[name] collects (start, end) torch.cuda.Event pairs while enabled.
TIMERS = None


def enable_timers(on=True):
    global TIMERS
    TIMERS = {} if on else None
    return TIMERS


def call(name, *args):
    lib = load()
    if TIMERS is not None and not name.startswith(("gm_table", "gm_last", "gm_abi")):
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        start.record()
        code = getattr(lib, name)(*args)
        end.record()
        TIMERS.setdefault(name, []).append((start, end))
    else:
        code = getattr(lib, name)(*args)
    if code != 0:
        raise GmError(name, code, lib.gm_last_error().decode())
    return code


def query(name, *args):
    return getattr(load(), name)(*args)


def ptr(t):
    """Device pointer of a tensor (or None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream_ptr(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
