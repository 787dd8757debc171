"""Seeded synthetic block-tree datasets for the benchmark configurations.

A dataset is a set of subtree roots ("blocks": level + index in [-5,5]^3,
the reference's world cube) each with a local block tree (children array,
node 0 = root, child local index bit0 = x), its n2i node list and fp32
sparse values sampled from an analytic field at the node positions.  Nodes
outside a block's occupied region carry the reference's "no field" value
1e30.

Field: sum of isotropic Gaussians (seeded centres/sigmas) plus optional
multi-octave value noise, evaluated in float64 on the unit cube and rounded
to float32, like the reference's exact-field providers sample a function at
tree nodes (BlockTreeFieldFromFunc.hpp:51-87).
"""
import math

import numpy as np

from . import api

NO_FIELD = np.float32(1e30)


# ---------------------------------------------------------------- fields

class Field:
    def __init__(self, seed, n_gauss=3, centre=(0.2, 0.8), sigma=(0.05, 0.15), noise_octaves=0,
                 noise_amp=0.1):
        rng = np.random.default_rng(seed)
        self.c = rng.uniform(centre[0], centre[1], size=(n_gauss, 3))
        self.s = rng.uniform(sigma[0], sigma[1], size=n_gauss)
        self.octaves = noise_octaves
        self.amp = noise_amp
        # value-noise lattices of 4, 8, 16, ... cells per axis
        self.lat = [rng.uniform(-1.0, 1.0, size=(4 * 2 ** o + 1,) * 3) for o in range(noise_octaves)]

    def __call__(self, u):
        """u: (..., 3) float64 unit-cube coordinates -> float64 values."""
        u = np.asarray(u, dtype=np.float64)
        f = np.zeros(u.shape[:-1])
        for c, s in zip(self.c, self.s):
            d2 = ((u - c) ** 2).sum(-1)
            f += np.exp(-d2 / (2 * s * s))
        for o, L in enumerate(self.lat):
            cells = L.shape[0] - 1
            x = np.clip(u, 0.0, 1.0) * cells
            i = np.minimum(np.floor(x).astype(np.int64), cells - 1)
            t = x - i
            t = t * t * (3 - 2 * t)  # smoothstep
            v = 0.0
            for dz in (0, 1):
                for dy in (0, 1):
                    for dx in (0, 1):
                        w = ((t[..., 0] if dx else 1 - t[..., 0]) *
                             (t[..., 1] if dy else 1 - t[..., 1]) *
                             (t[..., 2] if dz else 1 - t[..., 2]))
                        v = v + w * L[i[..., 0] + dx, i[..., 1] + dy, i[..., 2] + dz]
            f += self.amp * v / (2 ** o)
        return f

    def eval_torch(self, u):
        """Same field on a torch float64 tensor (any device)."""
        import torch
        f = torch.zeros(u.shape[:-1], dtype=torch.float64, device=u.device)
        for c, s in zip(self.c, self.s):
            cc = torch.tensor(c, dtype=torch.float64, device=u.device)
            f += torch.exp(-((u - cc) ** 2).sum(-1) / (2 * s * s))
        for o, L in enumerate(self.lat):
            Lt = torch.tensor(L, dtype=torch.float64, device=u.device)
            cells = L.shape[0] - 1
            x = u.clamp(0.0, 1.0) * cells
            i = torch.clamp(torch.floor(x).long(), max=cells - 1)
            t = x - i
            t = t * t * (3 - 2 * t)
            v = torch.zeros_like(f)
            for dz in (0, 1):
                for dy in (0, 1):
                    for dx in (0, 1):
                        w = ((t[..., 0] if dx else 1 - t[..., 0]) *
                             (t[..., 1] if dy else 1 - t[..., 1]) *
                             (t[..., 2] if dz else 1 - t[..., 2]))
                        v = v + w * Lt[i[..., 0] + dx, i[..., 1] + dy, i[..., 2] + dz]
            f += self.amp * v / (2 ** o)
        return f

    def grad_norm(self, u, h=1e-4):
        g = np.zeros(np.asarray(u).shape[:-1])
        for a in range(3):
            e = np.zeros(3)
            e[a] = h
            g += ((self(u + e) - self(u - e)) / (2 * h)) ** 2
        return np.sqrt(g)


class SeparableField:
    """Sum of seeded isotropic Gaussians plus value noise, evaluated on the
    node lattice of a level (unit cube, 2^(level+depth) cells per axis) with
    each term factored per axis, for shares of millions of blocks.  A
    Gaussian is exp(-(x-cx)^2/2s^2) exp(-(y-cy)^2/2s^2) exp(-(z-cz)^2/2s^2);
    an octave of value noise is the trilinear, smoothstep-weighted
    interpolation of a seeded (4*2^o + 1)^3 lattice in [-1, 1] at amplitude
    amp / 2^o.  Float64 on the device, rounded to float32 by the caller."""

    def __init__(self, seed, n_gauss=32, centre=(0.2, 0.8), sigma=(0.05, 0.15), noise_octaves=4,
                 noise_amp=0.1):
        rng = np.random.default_rng(seed)
        self.c = rng.uniform(centre[0], centre[1], size=(n_gauss, 3))
        self.s = rng.uniform(sigma[0], sigma[1], size=n_gauss)
        self.amp = noise_amp
        self.lat = [rng.uniform(-1.0, 1.0, size=(4 * 2 ** o + 1,) * 3) for o in range(noise_octaves)]

    def blocks(self, torch, level, depth, idx, nodes=None):
        """Values at the nodes of blocks idx ((B, 3) int64 tensor, x y z) of
        the level: (B, m, m, m) float64 indexed [z, y, x], m = 2^depth + 1.
        nodes: optional node offsets per axis (default 0..2^depth)."""
        dev = idx.device
        g = float(2 ** (level + depth))
        loc = torch.arange(2 ** depth + 1, device=dev, dtype=torch.int64) if nodes is None else nodes
        m = loc.shape[0]
        gi = idx[:, :, None] * (2 ** depth) + loc[None, None, :]      # (B, 3, m) global node index
        u = gi.double() / g
        c = torch.tensor(self.c, dtype=torch.float64, device=dev)       # (K, 3)
        s = torch.tensor(self.s, dtype=torch.float64, device=dev)
        e = torch.exp(-(u[None] - c[:, None, :, None]) ** 2 / (2.0 * s[:, None, None, None] ** 2))  # (K,B,3,m)
        ex, ey, ez = e[:, :, 0], e[:, :, 1], e[:, :, 2]                # (K, B, m)
        B = idx.shape[0]
        K = c.shape[0]
        zy = (ez[:, :, :, None] * ey[:, :, None, :]).reshape(K, B, m * m).permute(1, 2, 0).contiguous()
        f = torch.bmm(zy, ex.permute(1, 0, 2).contiguous()).reshape(B, m, m, m)
        for o, L in enumerate(self.lat):
            cells = L.shape[0] - 1
            x = u * cells                                               # (B, 3, m)
            i0 = torch.clamp(torch.floor(x).long(), max=cells - 1)
            t = x - i0
            w1 = t * t * (3.0 - 2.0 * t)
            base = i0.min(dim=2).values                                 # (B, 3): first cell of the block
            rel = i0 - base[:, :, None]                                 # cell of a node, from the first
            # lattice points a block can touch: the cells its 2^depth node
            # steps can span, plus the far corner of the last one
            P = -(-(int(loc.max()) - int(loc.min())) * cells // int(g)) + 2
            assert int(rel.max()) + 1 < P
            W = torch.zeros((B, 3, m, P), dtype=torch.float64, device=dev)
            W.scatter_add_(3, rel[..., None], (1.0 - w1)[..., None])
            W.scatter_add_(3, rel[..., None] + 1, w1[..., None])
            Lt = torch.tensor(L, dtype=torch.float64, device=dev)       # [x, y, z]
            rp = torch.arange(P, device=dev)
            ix = torch.clamp(base[:, 0, None] + rp, max=cells)          # (B, P)
            iy = torch.clamp(base[:, 1, None] + rp, max=cells)
            iz = torch.clamp(base[:, 2, None] + rp, max=cells)
            lb = Lt[ix[:, :, None, None], iy[:, None, :, None], iz[:, None, None, :]]  # (B, Px, Py, Pz)
            # contract x, y, z with the per-axis weights -> (B, z, y, x)
            v = torch.einsum("bijk,bxi->bxjk", lb, W[:, 0])
            v = torch.einsum("bxjk,byj->bxyk", v, W[:, 1])
            v = torch.einsum("bxyk,bzk->bzyx", v, W[:, 2])
            f = f + (self.amp / 2 ** o) * v
        return f


# ---------------------------------------------------------------- trees

def full_subtree(depth):
    """Children array of a fully refined subtree of the given depth
    (breadth-first numbering)."""
    counts = [8 ** l for l in range(depth + 1)]
    total = sum(counts)
    ch = np.zeros((total, 8), dtype=np.uint32)
    start = 0
    for l in range(depth):
        n = counts[l]
        nxt = start + n
        ch[start:start + n] = (nxt + np.arange(n * 8, dtype=np.uint32)).reshape(n, 8)
        start = nxt
    return ch


def refined_subtree(depth, refine):
    """Children array of a subtree refined where refine(level, locs) is True.
    refine gets the relative level and an (n,3) int array of block locations
    at that level and returns a bool array."""
    nodes_loc = [np.zeros((1, 3), dtype=np.int64)]
    level_start = [0]
    ch_rows = []
    total = 1
    for l in range(depth):
        locs = nodes_loc[-1]
        mask = np.asarray(refine(l, locs), dtype=bool)
        rows = np.zeros((locs.shape[0], 8), dtype=np.uint32)
        k = np.arange(8)
        bits = np.stack([k & 1, (k >> 1) & 1, (k >> 2) & 1], axis=1)
        sel = np.nonzero(mask)[0]
        nchild = sel.size * 8
        rows[sel] = (total + np.arange(nchild, dtype=np.uint32)).reshape(sel.size, 8)
        ch_rows.append(rows)
        child_locs = (locs[sel][:, None, :] * 2 + bits[None, :, :]).reshape(-1, 3)
        nodes_loc.append(child_locs)
        level_start.append(total)
        total += nchild
        if nchild == 0:
            break
    ch_rows.append(np.zeros((nodes_loc[-1].shape[0], 8), dtype=np.uint32))
    ch = np.concatenate(ch_rows[: len(nodes_loc)], axis=0)
    assert ch.shape[0] == total
    return ch


# ---------------------------------------------------------------- blocks

class Block:
    __slots__ = ("level", "index", "depth", "n2i", "sparse", "children")

    def __init__(self, level, index, depth, n2i, sparse, children=None):
        self.level = level
        self.index = tuple(int(v) for v in index)
        self.depth = depth
        self.n2i = n2i
        self.sparse = sparse
        self.children = children


def node_unit_coords(level, index, depth, n2i):
    """Unit-cube coordinates of a block's nodes."""
    scale = 1.0 / (2 ** (level + depth))
    base = np.asarray(index, dtype=np.float64) * (2 ** depth)
    return (base[None, :] + n2i.astype(np.float64)) * scale


def sample(field, level, index, depth, n2i, occupied=None):
    u = node_unit_coords(level, index, depth, n2i)
    v = field(u).astype(np.float32)
    if occupied is not None:
        v = np.where(occupied(u), v, NO_FIELD).astype(np.float32)
    return v


def field_range(blocks):
    lo, hi = np.inf, -np.inf
    for b in blocks:
        v = b.sparse[b.sparse != NO_FIELD]
        if v.size:
            lo = min(lo, float(v.min()))
            hi = max(hi, float(v.max()))
    return np.float32(lo), np.float32(hi)


def uniform_level(level, depth, field, n2i=None):
    """All 8^level blocks of a level, each a fully refined depth-`depth`
    subtree (full fill)."""
    if n2i is None:
        n2i = api.subtree_n2i(full_subtree(depth), depth)
    blocks = []
    m = 2 ** level
    for z in range(m):
        for y in range(m):
            for x in range(m):
                blocks.append(Block(level, (x, y, z), depth, n2i,
                                    sample(field, level, (x, y, z), depth, n2i)))
    return blocks


# ---------------------------------------------------------------- transfer function

def seeded_tf(seed, size=256, points=6):
    """Seeded piecewise-linear RGBA ramps sampled into a size x 4 table with
    makeColorTransferVector (ColorTransferFunction.hpp:27-69)."""
    rng = np.random.default_rng(seed)
    xs = np.sort(rng.uniform(0.0, 1.0, size=points - 2))
    xs = np.concatenate([[0.0], xs, [1.0]]).astype(np.float32)
    cols = rng.uniform(0.0, 1.0, size=(points, 4)).astype(np.float32)
    cols[0, 3] = 0.0  # transparent at the bottom of the range
    return api.color_transfer_vector(xs, cols, size)


# ---------------------------------------------------------------- camera

def look_at(eye, center=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)):
    """Row-major world->eye rigid transform of a right-handed camera looking
    along -Z (the VisualizationLibrary getLookAt convention)."""
    eye = np.asarray(eye, dtype=np.float64)
    f = np.asarray(center, dtype=np.float64) - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, np.asarray(up, dtype=np.float64))
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    R = np.stack([s, u, -f])
    t = -R @ eye
    M = np.eye(4)
    M[:3, :3] = R
    M[:3, 3] = t
    return M.astype(np.float32)


def orbit_camera(angle_deg, fov_y_deg=30.0, elevation_deg=20.0):
    """Camera on the orbit of the reference's default distance
    L = 10 sqrt(3) 1.01 / (2 sin(fovY/2)) (CameraController.cpp:88-101)."""
    L = 10.0 * math.sqrt(3.0) * 1.01 / (2.0 * math.sin(0.5 * math.radians(fov_y_deg)))
    a, e = math.radians(angle_deg), math.radians(elevation_deg)
    eye = L * np.array([math.cos(e) * math.sin(a), math.sin(e), math.cos(e) * math.cos(a)])
    return look_at(eye)


def default_camera(fov_y_deg=30.0):
    """CameraController::resetCameraTransform: eye along (2,1,3)."""
    d = np.array([2.0, 1.0, 3.0])
    d /= np.linalg.norm(d)
    L = 10.0 * math.sqrt(3.0) * 1.01 / (2.0 * math.sin(0.5 * math.radians(fov_y_deg)))
    return look_at(d * L)
