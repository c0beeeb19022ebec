"""Timings of the stages outside the MASTER pipeline's Legendre passes, one JSON line each.

    python -m gmaster_amd.stagebench pure LMAX [REPS]        pure-E/B coupling blocks
    python -m gmaster_amd.stagebench apodize NSIDE DEG [REPS] C1 apodization of a 20-hole mask
    python -m gmaster_amd.stagebench synfast NSIDE [REPS]     T+P Gaussian fields

Times are means over REPS runs after one warm-up, bracketed by ``torch.cuda.synchronize``;
inputs are built on the device before the timed runs.
"""

import json
import sys
import time

import numpy as np
import torch


def _timed(fn, reps):
    fn()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t) / reps


def pure(lmax, reps):
    from .purecoupling import clear_cache, pure_coupling_blocks
    ell = np.arange(2 * lmax + 1)
    W = torch.as_tensor(4 * np.pi / (1 + ell / 3.0) ** 3, device="cuda")

    def run():
        clear_cache()
        pure_coupling_blocks(W, lmax)

    dt = _timed(run, reps)
    L, N = lmax + 1, 2 * lmax + 1
    flop = 5 * 2.0 * L * L * N
    return {"stage": "pure_coupling_blocks", "lmax": lmax, "seconds": dt, "gemm_flop": flop,
            "tflops_over_stage": flop / dt / 1e12}


def apodize(nside, deg, reps):
    from . import mask_apodization
    from .workload import hole_mask
    m = hole_mask(nside, nholes=20, radius_deg=5.0, apo_deg=0.0, device="cuda")
    dt = _timed(lambda: mask_apodization(m, deg, "C1"), reps)
    return {"stage": "mask_apodization", "nside": nside, "aposize_deg": deg, "seconds": dt}


def synfast(nside, reps):
    from . import synfast_spherical
    lmax = 3 * nside - 1
    ell = np.arange(lmax + 1)
    tt = 400.0 / (ell + 10.0) ** 2
    z = np.zeros(lmax + 1)
    cls = np.array([tt, 0.3 * np.sqrt(0.1) * tt, z, 0.1 * tt, z, 1e-3 * tt])
    dt = _timed(lambda: synfast_spherical(nside, cls, [0, 2], seed=1), reps)
    return {"stage": "synfast_spherical_TP", "nside": nside, "seconds": dt}


def main(argv):
    what = argv[0]
    if what == "pure":
        out = pure(int(argv[1]), int(argv[2]) if len(argv) > 2 else 2)
    elif what == "apodize":
        out = apodize(int(argv[1]), float(argv[2]), int(argv[3]) if len(argv) > 3 else 2)
    elif what == "synfast":
        out = synfast(int(argv[1]), int(argv[2]) if len(argv) > 2 else 2)
    else:
        raise SystemExit(__doc__)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
