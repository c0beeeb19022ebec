"""Write the oracle fixtures under tests/golden/ (run from the repo root).

The reference's own tests hold no fixtures (they compare live against pymaster); these
are outputs of this repository's FP64 oracle, itself pinned by closed forms, brute force
and exact 3j values in tests/test_oracle_*.py.  Inputs are regenerated from seeds by
gmaster_amd.workload, so only outputs are stored.

    python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from gmaster_amd import workload  # noqa: E402
from oracle import master, sht as osht  # noqa: E402
from oracle.coupling import coupling_blocks_exact  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def inputs(nside, lmax, seed_mask=1234, seed_map=4321):
    """Mask and T, Q, U on the host: mask from the workload formula (torch CPU), maps from
    the oracle's own alm2map so the fixture needs no GPU."""
    mask = workload.hole_mask(nside, nholes=20, radius_deg=5.0, apo_deg=2.0, kind="C1", seed=seed_mask,
                              device="cpu").numpy()
    alm = workload.gaussian_alms(lmax, seed_map)
    t = osht.alm2map(alm[0:1], 0, nside, lmax)[0]
    qu = osht.alm2map(alm[1:3], 2, nside, lmax)
    return mask, np.concatenate([t[None], qu])


def pipeline_fixture(nside, lmax, nlb, n_iter, name):
    mask, tqu = inputs(nside, lmax)
    res = master.pipeline_tp(mask, tqu, nside, lmax, nlb, n_iter)
    keep = {k: v for k, v in res.items() if k.startswith(("bp_", "cl_", "pcl_mask", "mcm_binned"))}
    np.savez_compressed(os.path.join(HERE, name), mask=mask, tqu=tqu, **keep)
    print("wrote", name, {k: np.asarray(v).shape for k, v in keep.items()})


def threej_fixture():
    rng = np.random.default_rng(99)
    lmax = 12
    window = rng.uniform(size=2 * lmax + 1)
    blocks = coupling_blocks_exact(window, lmax)
    np.savez_compressed(os.path.join(HERE, "coupling_exact_lmax12.npz"), window=window,
                        m00=blocks[0], m02=blocks[1], m22p=blocks[2], m22m=blocks[3])
    print("wrote coupling_exact_lmax12.npz")


if __name__ == "__main__":
    torch.set_num_threads(8)
    threej_fixture()
    pipeline_fixture(16, 47, 8, 3, "pipeline_nside16.npz")
    pipeline_fixture(64, 191, 16, 3, "pipeline_nside64.npz")
