"""HIP interpolation against the CPU oracle (DenseFieldInterpolator.hpp:198-419
restated): bit-exact index maps and bit-exact dense bricks (the bar is 1e-6
relative; the kernels keep the CPU operand order, so results are identical)."""
import numpy as np
import pytest

from revise_amd import api, datasets
from tests.gpu_util import bits_equal_nan, say, tiled4_expected, to_dev

pytestmark = pytest.mark.gpu


def _case(seed, depth, p_refine, p_nofield=0.0):
    rng = np.random.default_rng(seed)
    ch = datasets.refined_subtree(depth, lambda l, locs: rng.random(locs.shape[0]) < p_refine)
    d = api.subtree_depth(ch, depth)
    n2i = api.subtree_n2i(ch, d)
    vals = rng.normal(size=n2i.shape[0]).astype(np.float32)
    if p_nofield:
        vals[rng.random(n2i.shape[0]) < p_nofield] = np.float32(1e30)
    return d, n2i, vals


def test_ordinals_bit_exact(orc):
    import torch
    for seed in range(4):
        d, n2i, _ = _case(seed, 6, 0.3)
        say("ordinals seed", seed, "depth", d, "n", n2i.shape[0])
        out = torch.empty(n2i.shape[0], dtype=torch.int32, device="cuda")
        api.node_ordinals(to_dev(n2i.view(np.int16)), d, out)
        got = out.cpu().numpy().view(np.uint32)
        assert np.array_equal(got, orc.ordinals(n2i, d))


@pytest.mark.parametrize("seed,depth,p,pnf", [
    (0, 0, 0.0, 0.0), (1, 1, 1.0, 0.0), (2, 2, 0.6, 0.0), (3, 3, 0.5, 0.05), (4, 4, 0.4, 0.0),
    (5, 5, 0.35, 0.02), (6, 6, 0.3, 0.0), (7, 6, 0.45, 0.01), (8, 6, 1.0, 0.0), (9, 7, 0.25, 0.0),
    (10, 8, 0.12, 0.0)])
def test_interp_planned_bit_exact(orc, seed, depth, p, pnf):
    import torch
    d, n2i, vals = _case(seed, depth, p, pnf)
    say("interp seed", seed, "depth", d, "nodes", n2i.shape[0])
    V = api.brick_nodes(d)
    plan = api.interp_plan(n2i, d)
    dense = torch.empty(V, dtype=torch.float32, device="cuda")
    api.interp_brick_planned(to_dev(vals), plan, d, dense)
    got = dense.cpu().numpy()
    exp = orc.interpolate(vals, n2i, d)
    assert bits_equal_nan(got, exp)


def test_interp_unplanned(orc):
    import torch
    d, n2i, vals = _case(21, 5, 0.4, 0.02)
    V = api.brick_nodes(d)
    exp = orc.interpolate(vals, n2i, d)
    a = torch.empty(V, dtype=torch.float32, device="cuda")
    api.interp_brick(to_dev(vals), n2i, d, a)  # plan staged inside the call
    assert bits_equal_nan(a.cpu().numpy(), exp)


@pytest.mark.parametrize("depth", [3, 6, 8])
def test_interp_full_tree_with_nan_values(orc, depth):
    """A fully refined subtree whose sparse data holds NaN (undefined) values:
    the reference interpolates those nodes like absent ones, so the kernel's
    all-nodes-present fast path must fall back to the full method."""
    import torch
    rng = np.random.default_rng(depth)
    n2i = api.subtree_n2i(datasets.full_subtree(depth), depth)
    vals = rng.normal(size=n2i.shape[0]).astype(np.float32)
    vals[rng.random(vals.size) < 0.002] = np.nan
    vals[rng.random(vals.size) < 0.001] = np.float32(1e30)
    dense = torch.empty(api.brick_nodes(depth), dtype=torch.float32, device="cuda")
    api.interp_brick_planned(to_dev(vals), api.interp_plan(n2i, depth), depth, dense)
    assert bits_equal_nan(dense.cpu().numpy(), orc.interpolate(vals, n2i, depth))


def test_interp_trilinear_field_reproduced():
    """Closed form: a globally trilinear field sampled at the nodes of any
    refinement pattern is reproduced at every interpolated node (edge, face and
    cell means are exact for trilinear functions)."""
    import torch
    rng = np.random.default_rng(4)
    ch = datasets.refined_subtree(6, lambda l, locs: rng.random(locs.shape[0]) < 0.3)
    d = api.subtree_depth(ch, 6)
    n2i = api.subtree_n2i(ch, d)
    c = rng.uniform(-1, 1, size=8)
    def f(x, y, z):
        return c[0] + c[1] * x + c[2] * y + c[3] * z + c[4] * x * y + c[5] * y * z + c[6] * x * z + c[7] * x * y * z
    u = n2i.astype(np.float64) / 2 ** d
    vals = f(u[:, 0], u[:, 1], u[:, 2]).astype(np.float32)
    V = api.brick_nodes(d)
    dense = torch.empty(V, dtype=torch.float32, device="cuda")
    api.interp_brick(to_dev(vals), n2i, d, dense)
    n = 2 ** d + 1
    g = np.arange(n) / 2 ** d
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    exp = f(x, y, z).reshape(-1)
    assert np.max(np.abs(dense.cpu().numpy() - exp)) < 1e-5


def test_interp_batch_mixed_depths_fp32_fp16(orc):
    import torch
    cases = [_case(100 + i, dep, p, 0.01) for i, (dep, p) in
             enumerate([(6, 0.3), (4, 0.5), (6, 1.0), (2, 0.9), (5, 0.4), (0, 0.0), (7, 0.2)])]
    jobs, outs, exps = [], [], []
    for i, (d, n2i, vals) in enumerate(cases):
        V = api.brick_nodes(d)
        plan = api.interp_plan(n2i, d)
        dt = torch.float16 if i % 2 else torch.float32
        dense = torch.empty(V, dtype=dt, device="cuda")
        jobs.append((to_dev(vals), plan, dense, d))
        outs.append(dense)
        exps.append(orc.interpolate(vals, n2i, d))
    b = api.InterpBatch(jobs)
    tiles, nodes = b.info()
    assert nodes == sum(api.brick_nodes(c[0]) for c in cases)
    say("batch tiles", tiles)
    b.run()
    torch.cuda.synchronize()
    for i, (o, e) in enumerate(zip(outs, exps)):
        got = o.cpu().numpy()
        if got.dtype == np.float16:
            assert bits_equal_nan(got.astype(np.float32), e.astype(np.float16).astype(np.float32))
        else:
            assert bits_equal_nan(got, e)
    b.close()


def test_interp_full_fill_is_scatter(orc):
    """Full-fill subtree: no node is interpolated, the brick is the scatter."""
    import torch
    d = 4
    n2i = api.subtree_n2i(datasets.full_subtree(d), d)
    vals = np.arange(n2i.shape[0], dtype=np.float32)
    dense = torch.empty(api.brick_nodes(d), dtype=torch.float32, device="cuda")
    api.interp_brick(to_dev(vals), n2i, d, dense)
    out = dense.cpu().numpy()
    assert np.array_equal(out[orc.ordinals(n2i, d)], vals)


def test_batch_rejects_plan_of_another_depth():
    """A job whose plan was built for another depth is refused when the batch
    is built (the kernel would otherwise index outside the brick)."""
    import torch
    from revise_amd import RvError
    d, n2i, vals = _case(31, 5, 0.5)
    plan = api.interp_plan(n2i, d)
    dense = torch.empty(api.brick_nodes(d + 1), dtype=torch.float32, device="cuda")
    with pytest.raises(RvError, match="another depth"):
        api.InterpBatch([(to_dev(vals), plan, dense, d + 1)])
    junk = torch.zeros(256, dtype=torch.uint8, device="cuda")
    with pytest.raises(RvError, match="not an interpolation plan"):
        api.InterpBatch([(to_dev(vals), junk, dense, d)])


@pytest.mark.parametrize("depth,p,offset", [(4, 1.0, 0), (4, 1.0, 1), (3, 0.6, 1), (2, 0.8, 0), (4, 0.5, 3)])
def test_interp_fp16_whole_brick_tile_alignment(orc, depth, p, offset):
    """Bricks of depth <= 4 are one tile, stored as one run: fp16 pairs when
    the brick is 4-byte aligned, single values otherwise (offset in fp16
    elements into a larger allocation)."""
    import torch
    d, n2i, vals = _case(200 + 10 * depth + offset, depth, p, 0.02)
    V = api.brick_nodes(d)
    buf = torch.full((V + 8,), -7.0, dtype=torch.float16, device="cuda")
    dense = buf[offset:offset + V]
    plan = api.interp_plan(n2i, d)
    b = api.InterpBatch([(to_dev(vals), plan, dense, d)])
    b.run()
    torch.cuda.synchronize()
    got = buf.cpu().numpy().astype(np.float32)
    e = orc.interpolate(vals, n2i, d).astype(np.float16).astype(np.float32)
    assert bits_equal_nan(got[offset:offset + V], e)
    assert (got[:offset] == -7.0).all() and (got[offset + V:] == -7.0).all()
    b.close()


def test_interp_batch_full_small_bricks(orc):
    """A batch of 3000 whole, fully refined bricks of depths 0-4 (each one
    tile, stored as one run), fp16 and fp32, some holding NaN values (the
    tile is then interpolated as a partial one, as the reference
    interpolates undefined nodes) or 1e30: the oracle's bricks."""
    import torch
    rng = np.random.default_rng(77)
    jobs, outs, exps = [], [], []
    plans = {d: (api.subtree_n2i(datasets.full_subtree(d), d)) for d in range(5)}
    dev_plans = {d: api.interp_plan(plans[d], d) for d in range(5)}
    for i in range(3000):
        d = int(rng.integers(0, 5))
        n2i = plans[d]
        vals = rng.normal(size=n2i.shape[0]).astype(np.float32)
        if i % 97 == 0:
            vals[rng.random(n2i.shape[0]) < 0.05] = np.nan
        if i % 89 == 0:
            vals[rng.random(n2i.shape[0]) < 0.05] = np.float32(1e30)
        dt = torch.float16 if i % 2 else torch.float32
        dense = torch.empty(api.brick_nodes(d), dtype=dt, device="cuda")
        jobs.append((to_dev(vals), dev_plans[d], dense, d))
        outs.append(dense)
        exps.append((vals, n2i, d))
    b = api.InterpBatch(jobs)
    b.run()
    torch.cuda.synchronize()
    for o, (vals, n2i, d) in zip(outs, exps):
        e = orc.interpolate(vals, n2i, d)
        got = o.cpu().numpy()
        if got.dtype == np.float16:
            assert bits_equal_nan(got.astype(np.float32), e.astype(np.float16).astype(np.float32))
        else:
            assert bits_equal_nan(got, e)
    b.close()


@pytest.mark.parametrize("seed,depth,p,pnf", [(30, 0, 0.0, 0.0), (31, 1, 1.0, 0.0), (32, 2, 0.6, 0.05),
                                              (33, 4, 0.4, 0.02), (34, 5, 0.5, 0.0), (35, 6, 0.35, 0.01),
                                              (36, 6, 1.0, 0.0), (37, 7, 0.3, 0.01), (38, 8, 0.1, 0.0)])
def test_batch_tiled_layout_holds_the_linear_values(orc, seed, depth, p, pnf):
    """A batch writing the micro-brick layout stores every float of the
    brick as tiled4_expected lays out the x-fastest brick's values (NaN
    where undefined): node x at x-position x + 1, rows of 5 x-positions
    overlapping by one, the edge nodes repeated beyond the brick, NaN rows
    beyond it along y and z.  Depths 7 and 8 have several x-tiles, which
    share the micro-bricks at their seams."""
    import torch
    d, n2i, vals = _case(seed, depth, p, pnf)
    plan = api.interp_plan(n2i, d)
    sp = to_dev(vals)
    lin = torch.empty(api.brick_nodes(d), dtype=torch.float32, device="cuda")
    til = torch.full((api.brick_floats(d, api.BRICK_TILED4),), 12345.0, dtype=torch.float32, device="cuda")
    api.InterpBatch([(sp, plan, lin, d)]).run()
    api.InterpBatch([(sp, plan, til, d)], api.BRICK_TILED4).run()
    torch.cuda.synchronize()
    n = 2 ** d + 1
    G = ((n + 1) >> 2) + 1
    assert til.numel() == G ** 3 * 80
    assert bits_equal_nan(til.cpu().numpy(), tiled4_expected(lin.cpu().numpy(), n))
    assert bits_equal_nan(lin.cpu().numpy(), orc.interpolate(vals, n2i, d))


def test_cpu_interp_configuration_brick_bit_exact(orc):
    """The level-2, depth-6 brick of the refined 3-Gaussian tree (the CPU
    interpolation workload, 65^3 nodes, partly refined): rv_interp_brick
    and the batched micro-brick path against the oracle, bit-exact."""
    import torch
    from revise_amd import configs
    c = configs.cpu_interp_brick()
    n2i, d = c["n2i"], c["depth"]
    exp = orc.interpolate(c["sparse"], n2i, d)
    sp = torch.from_numpy(c["sparse"]).cuda()
    dense = torch.empty(api.brick_nodes(d), dtype=torch.float32, device="cuda")
    api.interp_brick(sp, n2i, d, dense)
    got = dense.cpu().numpy()
    say("config-1 brick: %d sparse nodes, %d dense" % (n2i.shape[0], exp.size))
    assert not np.isnan(exp).any()
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
    til = torch.empty(api.brick_floats(d, api.BRICK_TILED4), dtype=torch.float32, device="cuda")
    api.InterpBatch([(sp, api.interp_plan(n2i, d), til, d)], api.BRICK_TILED4).run()
    t = til.cpu().numpy()
    assert bits_equal_nan(t, tiled4_expected(exp, 2 ** d + 1))
