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


def host_threads(cap=16):
    """Worker threads for CPU reference work: the host cores this process
    may use, at most `cap`."""
    import os
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(cap, n))


def oracle_blocks_frame(orc, items, tf, W, H, init=None, rows=None, threads=None):
    """The oracle's per-block 8-bit frame: items = [(oracle params, brick)]
    back to front, each ray-cast over the viewport with the per-block 8-bit
    composite (orc_render_block), on `threads` threads, each owning a band of
    rows and walking every block in order.  rows: (y0, y1) limits the work to
    those viewport rows (the others stay as init)."""
    import ctypes as C
    import threading
    vp = np.zeros((H, W), np.uint32) if init is None else np.ascontiguousarray(init, dtype=np.uint32).copy()
    y0, y1 = rows if rows is not None else (0, H)
    L = orc.lib()
    tfc = np.ascontiguousarray(tf, dtype=np.float32)
    tp = tfc.ctypes.data_as(C.POINTER(C.c_float))
    vpp = vp.ctypes.data_as(C.POINTER(C.c_uint32))
    keep = [np.ascontiguousarray(b, dtype=np.float32) for _, b in items]
    bps = [k.ctypes.data_as(C.POINTER(C.c_float)) for k in keep]
    nt = threads or host_threads()
    step = max(1, (y1 - y0 + nt - 1) // nt)

    def band(r0, r1):
        for (po, _), bp in zip(items, bps):
            # params rows are viewport rows; the oracle takes [r0, r1)
            L.orc_render_block(C.byref(po), bp, tp, tfc.shape[0], vpp, r0, r1)

    ths = [threading.Thread(target=band, args=(r, min(y1, r + step))) for r in range(y0, y1, step)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    return vp


def tiled4_expected(lin, n):
    """The micro-brick (TILED4) brick holding the x-fastest brick `lin` of
    n^3 nodes, as the interpolation writes it: G^3 micro-bricks of 16 rows
    (y, then z) of 5 x-positions, micro-brick m holding x-positions 4m ..
    4m + 4; node x at x-position x + 1, the x-positions beyond the brick
    holding the edge nodes; rows beyond n - 1 along y or z NaN."""
    G = ((n + 1) >> 2) + 1
    L = np.asarray(lin, np.float32).reshape(n, n, n)
    P = np.full((4 * G, 4 * G, n), np.nan, np.float32)
    P[:n, :n, :] = L
    zz = 4 * np.arange(G)[:, None] + np.arange(4)[None, :]
    node = np.clip(4 * np.arange(G)[:, None] + np.arange(5)[None, :] - 1, 0, n - 1)
    E = P[zz[:, None, None, :, None, None], zz[None, :, None, None, :, None], node[None, None, :, None, None, :]]
    return E.reshape(-1)
