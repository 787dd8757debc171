"""Device-resident synthetic workloads of the benchmark configurations.

* uniform(level, depth): every block of a level, each a fully refined
  depth-`depth` subtree (config "GPU plain drop-in": level 3, depth 6 ->
  512 x 65^3 fp32, full fill).
* mixed(nblocks, ...): disjoint blocks of levels 2..6 produced by seeded
  splitting of the 64 level-2 blocks (splits follow the field gradient),
  each a depth-6 subtree whose occupied part is a slab covering a seeded
  fraction `fill` in [5%, 100%] of the block: there the subtree is fully
  refined and carries field values; elsewhere it is refined only to depth 2
  and its nodes are empty (1e30 -> NaN after interpolation).  Subtree shapes
  are drawn from 16 fill levels, so 16 gather plans serve all blocks.
Sparse values are sampled on the GPU in float64 and rounded to float32.
"""
import numpy as np

from . import api, datasets

NO_FIELD = float(np.float32(1e30))


class Workload:
    def __init__(self):
        self.blocks = []      # dicts: level, index, depth, sparse, plan, brick
        self.fmin = self.fmax = None
        self.n2i = {}         # shape key -> host n2i
        self.plans = {}       # shape key -> device plan
        self.sparse_nodes = 0
        self.dense_nodes = 0


def _coords(torch, level, index, depth, n2i_dev):
    scale = 1.0 / 2 ** (level + depth)
    base = torch.tensor(index, dtype=torch.float64, device=n2i_dev.device) * 2 ** depth
    return (base[None, :] + n2i_dev) * scale


def uniform(torch, field, level=3, depth=6, brick_dtype=None):
    brick_dtype = brick_dtype or torch.float32
    w = Workload()
    n2i = api.subtree_n2i(datasets.full_subtree(depth), depth)
    w.n2i["full"] = n2i
    plan = w.plans["full"] = api.interp_plan(n2i, depth)
    n2i_dev = torch.from_numpy(n2i.astype(np.int64)).cuda().double()
    m = 2 ** level
    N, V = n2i.shape[0], api.brick_nodes(depth)
    sparse = torch.empty((m ** 3, N), dtype=torch.float32, device="cuda")
    bricks = torch.empty((m ** 3, V), dtype=brick_dtype, device="cuda")
    b = 0
    for z in range(m):
        for y in range(m):
            for x in range(m):
                sparse[b] = field.eval_torch(_coords(torch, level, (x, y, z), depth, n2i_dev)).float()
                w.blocks.append({"level": level, "index": (x, y, z), "depth": depth,
                                 "sparse": sparse[b], "plan": plan, "brick": bricks[b],
                                 "shape": "full"})
                b += 1
    w.sparse, w.bricks = sparse, bricks
    w.fmin, w.fmax = np.float32(sparse.min().item()), np.float32(sparse.max().item())
    w.sparse_nodes, w.dense_nodes = m ** 3 * N, m ** 3 * V
    return w


def slab_subtree(depth, fill, coarse_depth=2):
    """Subtree refined fully where the block cell lies in z < fill, to
    coarse_depth elsewhere."""
    def refine(l, locs):
        if l < coarse_depth:
            return np.ones(locs.shape[0], dtype=bool)
        zlo = locs[:, 2] / 2.0 ** l
        return zlo < fill
    return datasets.refined_subtree(depth, refine)


def mixed_keys(field, nblocks=2000, seed=0, min_level=2, max_level=6):
    rng = np.random.default_rng(seed)
    keys = {(min_level, x, y, z) for z in range(4) for y in range(4) for x in range(4)}
    grid = np.stack(np.meshgrid(*[np.linspace(0.05, 0.95, 24)] * 3, indexing="ij"), -1).reshape(-1, 3)
    g = field.grad_norm(grid)
    gmax = g.max()
    level_weight = {2: 1.0, 3: 0.5, 4: 0.35, 5: 0.25}
    while len(keys) < nblocks:
        cand = sorted(k for k in keys if k[0] < max_level)
        cen = np.array([[(k[1] + 0.5) / 2 ** k[0], (k[2] + 0.5) / 2 ** k[0], (k[3] + 0.5) / 2 ** k[0]]
                        for k in cand])
        lev = np.array([k[0] for k in cand])
        # split where the field varies; seeded sampling spreads the levels
        wgt = (field.grad_norm(cen) / gmax + 0.05) * np.array([level_weight.get(int(l), 0.2) for l in lev])
        take = min(len(cand), max(1, min(16, (nblocks - len(keys) + 6) // 7)))
        for j in rng.choice(len(cand), size=take, replace=False, p=wgt / wgt.sum()):
            L, x, y, z = cand[j]
            keys.discard((L, x, y, z))
            for k in range(8):
                keys.add((L + 1, 2 * x + (k & 1), 2 * y + ((k >> 1) & 1), 2 * z + ((k >> 2) & 1)))
    keys = sorted(keys)
    while len(keys) > nblocks:  # sparse set: drop a few seeded blocks
        keys.pop(int(rng.integers(len(keys))))
    return keys


def mixed(torch, field, nblocks=2000, depth=6, seed=0, n_fill=16, brick_dtype=None, keys=None):
    brick_dtype = brick_dtype or torch.float32
    rng = np.random.default_rng(seed + 1)
    w = Workload()
    keys = keys if keys is not None else mixed_keys(field, nblocks, seed)
    fills = 0.05 + 0.95 * np.arange(n_fill) / (n_fill - 1)
    shapes = {}
    for k, f in enumerate(fills):
        ch = slab_subtree(depth, f)
        d = api.subtree_depth(ch, depth)
        n2i = api.subtree_n2i(ch, d)
        plan = api.interp_plan(n2i, d)
        shapes[k] = (d, n2i, torch.from_numpy(n2i.astype(np.int64)).cuda().double(), plan, float(f))
        w.n2i[k] = n2i
        w.plans[k] = plan
    V = api.brick_nodes(depth)
    bricks = torch.empty((len(keys), V), dtype=brick_dtype, device="cuda")
    lo, hi = np.inf, -np.inf
    sparse_all = []
    for i, (L, x, y, z) in enumerate(keys):
        k = int(rng.integers(n_fill))
        d, n2i, n2i_dev, plan, f = shapes[k]
        u = _coords(torch, L, (x, y, z), d, n2i_dev)
        v = field.eval_torch(u)
        # occupied: node inside the slab z <= fill of the block
        zloc = n2i_dev[:, 2] / 2 ** d
        occ = zloc <= f + 1e-12
        lo = min(lo, float(v[occ].min()))
        hi = max(hi, float(v[occ].max()))
        sp = torch.where(occ, v, torch.full_like(v, NO_FIELD)).float()
        sparse_all.append(sp)
        w.blocks.append({"level": L, "index": (x, y, z), "depth": d, "sparse": sp, "plan": plan,
                         "brick": bricks[i], "shape": k, "fill": f})
        w.sparse_nodes += n2i.shape[0]
        w.dense_nodes += V
    w.bricks = bricks
    w.sparse = sparse_all
    w.fmin, w.fmax = np.float32(lo), np.float32(hi)
    return w


class LargeShare:
    """One GPU's share of a deep uniform level: keys (B, 3) int64 block
    indices (x, y, z), sparse (B, N) fp32 node values in n2i order (rows
    padded to sparse_stride), bricks (B, brick_stride) fp16 (or fp32), one
    shared gather plan (every block is a full subtree)."""

    def frame_blocks(self, with_sparse=True):
        """numpy array of api.FRAME_BLOCK_DTYPE (device addresses)."""
        B = self.keys.shape[0]
        a = np.zeros(B, dtype=api.FRAME_BLOCK_DTYPE)
        a["level"] = self.level
        a["index"] = self.keys.astype(np.uint32)
        a["depth"] = self.depth
        rows = np.arange(B, dtype=np.uint64)
        if with_sparse:
            a["sparse"] = np.uint64(self.sparse.data_ptr()) + rows * np.uint64(4 * self.sparse_stride)
            a["plan"] = self.plan.data_ptr()
        es = self.bricks.element_size()
        base = np.uint64(self.bricks.data_ptr()) + rows * np.uint64(es * self.brick_stride)
        if es == 2:
            a["brick16"] = base
        else:
            a["brick"] = base
        return a

    def brick_jobs(self):
        """numpy array of api.BRICK_JOB_DTYPE for api.InterpBatch."""
        B = self.keys.shape[0]
        a = np.zeros(B, dtype=api.BRICK_JOB_DTYPE)
        rows = np.arange(B, dtype=np.uint64)
        a["sparse"] = np.uint64(self.sparse.data_ptr()) + rows * np.uint64(4 * self.sparse_stride)
        a["plan"] = self.plan.data_ptr()
        es = self.bricks.element_size()
        base = np.uint64(self.bricks.data_ptr()) + rows * np.uint64(es * self.brick_stride)
        a["dense16" if es == 2 else "dense"] = base
        a["depth"] = self.depth
        return a


def large_share(torch, seed, level=8, depth=4, occupancy=0.12, n_gauss=32, octaves=4,
                brick_dtype=None, chunk=32768, device="cuda"):
    """The blocks of one level (8^level of them) whose centre values are the
    largest `occupancy` fraction of the field, each a fully refined
    depth-`depth` subtree; field: SeparableField (n_gauss Gaussians plus
    `octaves` of value noise).  Rows are padded to 16-byte multiples."""
    brick_dtype = brick_dtype or torch.float16
    field = datasets.SeparableField(seed, n_gauss=n_gauss, noise_octaves=octaves)
    m = 2 ** level
    dev = device
    # centre value of every block of the level, z slab by z slab
    centre = torch.empty(m ** 3, dtype=torch.float32, device=dev)
    mid = torch.tensor([2 ** (depth - 1)], dtype=torch.int64, device=dev)
    ar = torch.arange(m, dtype=torch.int64, device=dev)
    zc = max(1, (1 << 20) // (m * m))
    for z0 in range(0, m, zc):
        z1 = min(m, z0 + zc)
        zz, yy, xx = torch.meshgrid(torch.arange(z0, z1, device=dev), ar, ar, indexing="ij")
        idx = torch.stack([xx.reshape(-1), yy.reshape(-1), zz.reshape(-1)], 1)
        centre[z0 * m * m:z1 * m * m] = field.blocks(torch, level, depth, idx, mid).reshape(-1).float()
    nsel = int(round(occupancy * m ** 3))
    sel = torch.sort(torch.argsort(centre, descending=True)[:nsel]).values  # linear z-major index
    del centre
    keys = torch.stack([sel % m, (sel // m) % m, sel // (m * m)], 1)
    n2i = api.subtree_n2i(datasets.full_subtree(depth), depth)
    nn = 2 ** depth + 1
    N, V = n2i.shape[0], api.brick_nodes(depth)
    fi = torch.from_numpy(((n2i[:, 2].astype(np.int64) * nn + n2i[:, 1]) * nn + n2i[:, 0])).to(dev)
    w = LargeShare()
    w.level, w.depth, w.field = level, depth, field
    w.n2i = n2i
    w.plan = torch.from_numpy(api.interp_plan_encode(n2i, depth)).to(dev)
    w.sparse_stride = (N + 3) // 4 * 4
    w.brick_stride = (V + 7) // 8 * 8
    w.sparse = torch.empty((nsel, w.sparse_stride), dtype=torch.float32, device=dev)
    w.bricks = torch.empty((nsel, w.brick_stride), dtype=brick_dtype, device=dev)
    lo, hi = float("inf"), float("-inf")
    for c0 in range(0, nsel, chunk):
        c1 = min(nsel, c0 + chunk)
        v = field.blocks(torch, level, depth, keys[c0:c1]).reshape(c1 - c0, nn ** 3)[:, fi].float()
        w.sparse[c0:c1, :N] = v
        lo = min(lo, float(v.min()))
        hi = max(hi, float(v.max()))
    w.sparse[:, N:] = 0.0
    w.keys = keys.cpu().numpy()
    w.fmin, w.fmax = np.float32(lo), np.float32(hi)
    w.sparse_nodes, w.dense_nodes = nsel * N, nsel * V
    return w


class RefinedTree:
    """The block tree of the CPU interpolation configuration: an octree over
    the unit cube refined to level `octree_levels` wherever |grad f| exceeds
    `frac` of its maximum, each octree leaf carrying `block_depth` further
    full levels (a leaf "block" of 16^3 cells for block_depth 4).

    leaf_level[i, j, k]: the octree-leaf level containing level-
    `octree_levels` cell (i, j, k).  A block of the tree at level a is
    refined when a < octree_levels (it lies inside a leaf's cells or is an
    internal octree node) or when a < L + block_depth, L the level of the
    octree leaf containing it."""

    def __init__(self, field, octree_levels=4, block_depth=4, frac=0.5, samples=5, grid=129):
        self.field = field
        self.octree_levels = octree_levels
        self.block_depth = block_depth
        g = np.linspace(0.0, 1.0, grid)
        u = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
        self.gmax = float(field.grad_norm(u).max())
        thr = frac * self.gmax
        m = 2 ** octree_levels
        leaf = np.full((m, m, m), -1, dtype=np.int32)
        self.leaves = []
        s = (np.arange(samples) + 0.5) / samples
        loc_s = np.stack(np.meshgrid(s, s, s, indexing="ij"), -1).reshape(-1, 3)

        def visit(level, idx):
            size = 1.0 / 2 ** level
            refine = False
            if level < octree_levels:
                pts = (np.asarray(idx, dtype=np.float64)[None, :] + loc_s) * size
                refine = bool((field.grad_norm(pts) > thr).any())
            if refine:
                for c in range(8):
                    visit(level + 1, (2 * idx[0] + (c & 1), 2 * idx[1] + ((c >> 1) & 1),
                                      2 * idx[2] + ((c >> 2) & 1)))
                return
            self.leaves.append((level, tuple(idx)))
            sh = octree_levels - level
            x, y, z = (v << sh for v in idx)
            leaf[x:x + (1 << sh), y:y + (1 << sh), z:z + (1 << sh)] = level

        visit(0, (0, 0, 0))
        assert (leaf >= 0).all()
        self.leaf_level = leaf

    def subtree(self, level, index, max_depth=6):
        """Children table of the subtree rooted at block (level, index),
        walked to max_depth levels below it."""
        ol, bd, leaf = self.octree_levels, self.block_depth, self.leaf_level
        index = np.asarray(index, dtype=np.int64)

        def refine(l, locs):
            a = level + l
            if a < ol:
                return np.ones(locs.shape[0], dtype=bool)
            loc = index[None, :] * 2 ** l + locs
            c = loc >> (a - ol)
            return a < leaf[c[:, 0], c[:, 1], c[:, 2]] + bd

        return datasets.refined_subtree(max_depth, refine)


def cpu_interp_brick(seed=2024, root_level=2, max_depth=6, octree_levels=4, block_depth=4, frac=0.5):
    """The CPU interpolation workload: the 3-Gaussian field's refined tree
    and one level-`root_level` subtree of depth `max_depth` (65^3 nodes for
    depth 6): among the roots whose subtree reaches that depth, the one with
    the median node count.  Returns a dict: tree, level, index, depth,
    children, n2i, sparse (fp32 values of f at the nodes)."""
    field = datasets.Field(seed, n_gauss=3)
    tree = RefinedTree(field, octree_levels, block_depth, frac)
    m = 2 ** root_level
    cands = []
    for z in range(m):
        for y in range(m):
            for x in range(m):
                ch = tree.subtree(root_level, (x, y, z), max_depth)
                d = api.subtree_depth(ch, max_depth)
                if d == max_depth:
                    cands.append(((x, y, z), ch, d))
    assert cands, "no level-%d subtree reaches depth %d" % (root_level, max_depth)
    sized = sorted((api.subtree_n2i(ch, d).shape[0], idx, ch, d) for idx, ch, d in cands)
    n, idx, ch, d = sized[len(sized) // 2]
    n2i = api.subtree_n2i(ch, d)
    sparse = datasets.sample(field, root_level, idx, d, n2i)
    return {"tree": tree, "level": root_level, "index": idx, "depth": d, "children": ch, "n2i": n2i,
            "sparse": sparse, "candidates": len(cands)}
