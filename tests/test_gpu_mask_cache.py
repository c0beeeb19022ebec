"""The mask-alms cache is keyed on mask content.

Fields built from successive masks, each dropped before the next is made (so the caching
allocator can hand the next mask the freed mask's address), must report the alms of their
own mask.  No cache clearing happens between the cases.
"""

import gc

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_fingerprint_distinguishes_content():
    from gmaster_amd.field import mask_fingerprint

    rng = np.random.default_rng(3)
    a = torch.as_tensor(rng.uniform(size=12 * 64 * 64), device="cuda")
    b = a.clone()
    assert mask_fingerprint(a) == mask_fingerprint(b)
    b[12345] = np.nextafter(float(b[12345]), 2.0)  # one ulp in one word
    assert mask_fingerprint(a) != mask_fingerprint(b)
    c = a.clone()
    c[[7, 8]] = c[[8, 7]]  # two words exchanged: position matters
    assert mask_fingerprint(a) != mask_fingerprint(c)
    # odd byte counts take the tail path
    raw = torch.arange(13, dtype=torch.uint8, device="cuda")
    raw2 = raw.clone()
    raw2[12] = 99
    f1 = mask_fingerprint(raw)
    assert f1 == mask_fingerprint(raw.clone()) and f1 != mask_fingerprint(raw2)


def test_successive_masks_get_their_own_alms():
    import gmaster_amd as nmt
    from gmaster_amd import sht

    nside, lmax = 32, 95
    npix = 12 * nside * nside
    rng = np.random.default_rng(11)
    for k in range(3):
        mask_np = (rng.uniform(size=npix) > 0.2 + 0.2 * k).astype(np.float64)
        tmap = rng.normal(size=npix)
        f = nmt.NmtField(mask_np, [tmap], n_iter=3)
        got = f.get_mask_alms().as_subclass(torch.Tensor)
        fresh = sht.map2alm_s0(torch.as_tensor(mask_np, device="cuda")[None], lmax, n_iter=3)[0]
        err = float(torch.max(torch.abs(got - fresh)) / torch.max(torch.abs(fresh)))
        print(f"mask {k}: mask alms vs fresh analysis rel={err:.2e}", flush=True)
        assert err < 1e-12
        fP = nmt.NmtField(mask_np, [tmap, -tmap], n_iter=3)  # lazy spin-2 mask alms
        err2 = float(torch.max(torch.abs(fP.get_mask_alms().as_subclass(torch.Tensor) - fresh))
                     / torch.max(torch.abs(fresh)))
        assert err2 < 1e-12
        del f, fP, got, fresh
        gc.collect()
