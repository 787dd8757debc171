"""Synthetic data generators (CPU): the per-axis factored field of the large
share equals the direct evaluation of the same Gaussians + value noise at
every node, for levels where a block spans one or many noise cells; the
large share selects the blocks with the largest centre values and stores
its nodes in n2i order."""
import numpy as np
import pytest

from revise_amd import configs, datasets


def _direct(F, u):
    f = np.zeros(u.shape[:-1])
    for c, s in zip(F.c, F.s):
        f += np.exp(-((u - c) ** 2).sum(-1) / (2 * s * s))
    for o, L in enumerate(F.lat):
        cells = L.shape[0] - 1
        x = np.clip(u, 0, 1) * cells
        i = np.minimum(np.floor(x).astype(np.int64), cells - 1)
        t = x - i
        t = t * t * (3 - 2 * t)
        v = 0.0
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    w = ((t[..., 0] if dx else 1 - t[..., 0]) * (t[..., 1] if dy else 1 - t[..., 1]) *
                         (t[..., 2] if dz else 1 - t[..., 2]))
                    v = v + w * L[i[..., 0] + dx, i[..., 1] + dy, i[..., 2] + dz]
        f += F.amp * v / 2 ** o
    return f


@pytest.mark.parametrize("level,depth", [(0, 3), (2, 4), (4, 4), (8, 4), (3, 6)])
def test_separable_field_matches_direct(level, depth):
    import torch
    F = datasets.SeparableField(44, n_gauss=5, noise_octaves=4)
    m = 2 ** level
    idx = torch.tensor([[0, 0, 0], [m - 1, m - 1, m - 1], [m // 2, m // 3, m - 1]])
    v = F.blocks(torch, level, depth, idx).numpy()
    g, n = 2 ** (level + depth), 2 ** depth + 1
    zz, yy, xx = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    for b in range(3):
        u = np.stack([idx[b, 0].item() * 2 ** depth + xx, idx[b, 1].item() * 2 ** depth + yy,
                      idx[b, 2].item() * 2 ** depth + zz], -1) / g
        assert np.abs(v[b] - _direct(F, u)).max() < 1e-14


def test_large_share_selection_and_node_order():
    import torch
    w = configs.large_share(torch, 5, level=4, depth=2, occupancy=0.12, n_gauss=8, octaves=4,
                            brick_dtype=torch.float32, device="cpu", chunk=100)
    assert w.keys.shape == (round(0.12 * 4 ** 6), 3)
    # the selected blocks have the largest centre values
    allk = np.stack(np.meshgrid(np.arange(16), np.arange(16), np.arange(16), indexing="ij"), -1).reshape(-1, 3)[:, ::-1]
    cen = _direct(w.field, (allk * 4 + 2) / 64.0)
    thr = np.sort(cen)[::-1][w.keys.shape[0] - 1]
    sel = _direct(w.field, (w.keys * 4 + 2) / 64.0)
    assert sel.min() >= thr
    for b in (0, 17, w.keys.shape[0] - 1):
        u = (w.keys[b][None, :] * 4 + w.n2i.astype(np.int64)) / 64.0
        assert np.array_equal(w.sparse[b, :w.n2i.shape[0]].numpy(), _direct(w.field, u).astype(np.float32))
