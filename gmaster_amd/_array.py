"""Device results that also behave as host arrays.

pymaster users write ``np.asarray(result)`` (the reference returns JAX arrays, which
allow it).  ``GmArray`` is a ``torch.Tensor`` subclass that keeps every torch operation
on the device but converts to a NumPy array when NumPy asks for one.
"""

import numpy as np
import torch


class GmArray(torch.Tensor):
    def __array__(self, dtype=None, copy=None):
        out = self.as_subclass(torch.Tensor).detach().cpu().numpy()
        return out if dtype is None else out.astype(dtype)

    def numpy(self, *args, **kwargs):  # noqa: D401 - torch signature
        return self.as_subclass(torch.Tensor).detach().cpu().numpy()


def wrap(t):
    if t is None:
        return None
    if isinstance(t, GmArray):
        return t
    return t.as_subclass(GmArray)


def device_tensor(x, dtype=torch.float64, device=None):
    """Any array-like -> contiguous tensor on the GPU (no copy if already there)."""
    if isinstance(x, torch.Tensor):
        t = x.as_subclass(torch.Tensor)
    else:
        t = torch.as_tensor(np.asarray(x))
    dev = device if device is not None else (t.device if t.is_cuda else torch.device("cuda"))
    return t.to(device=dev, dtype=dtype).contiguous()
