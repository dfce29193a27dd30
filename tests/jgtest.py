"""Shared helpers for the tests that go through the C ABI."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "fluid-engine-dev_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

import jetgpu  # noqa: E402  (raises if lib/libjetgpu.so is not built)


def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        return None
    return torch


def dev(a, dtype=None):
    """Host numpy array -> device tensor (synchronised, contiguous)."""
    import torch

    t = torch.from_numpy(np.ascontiguousarray(a if dtype is None else a.astype(dtype))).cuda()
    torch.cuda.synchronize()
    return t


def empty(shape, dtype=None):
    import torch

    t = torch.empty(shape, dtype=dtype or torch.float64, device="cuda")
    torch.cuda.synchronize()
    return t


def host(t):
    import torch

    torch.cuda.synchronize()
    return t.cpu().numpy()


def rel_err(x, ref):
    return float(np.linalg.norm((x - ref).ravel()) / max(np.linalg.norm(ref.ravel()), 1e-300))


def rel_err_mean_free(x, ref, mask=None):
    """Relative error modulo the constant null vector of an all-Neumann system."""
    if mask is None:
        mask = np.ones(x.shape, bool)
    a = x[mask] - x[mask].mean()
    b = ref[mask] - ref[mask].mean()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
