"""Helpers for the GPU parity tests."""
import sys

import numpy as np


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def to_dev(a):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a))
    return t.to("cuda")


def bits_equal_nan(a, b):
    """Bitwise equality of float arrays with every NaN treated as equal."""
    a = np.asarray(a)
    b = np.asarray(b)
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return False
    return np.array_equal(a[~na].view(np.uint32) if a.dtype == np.float32 else a[~na],
                          b[~nb].view(np.uint32) if b.dtype == np.float32 else b[~nb])
