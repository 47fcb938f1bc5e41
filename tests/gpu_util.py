"""Helpers shared by the GPU parity tests."""
import numpy as np
import torch


def dev():
    return torch.device("cuda:0")


def t(a, dtype=None):
    x = torch.as_tensor(np.ascontiguousarray(a))
    if dtype is not None:
        x = x.to(dtype)
    return x.to(dev())


def n(x):
    return x.detach().cpu().numpy()


def scaled_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if b.size == 0:
        return 0.0
    scale = max(float(np.abs(b).max()), 1e-300)
    return float(np.abs(a - b).max()) / scale
