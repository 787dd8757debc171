"""Closed-form and invariant checks that pin the oracle's interpolation,
pow replacement and render setup (CPU only)."""
import numpy as np
import pytest

from revise_amd import api, datasets


def test_pow_error_bound(orc):
    rng = np.random.default_rng(0)
    bs = np.concatenate([rng.uniform(0, 1, 4000), 1 - 10 ** rng.uniform(-7, -1, 500),
                         10 ** rng.uniform(-30, -1, 500)]).astype(np.float32)
    es = np.concatenate([rng.uniform(0.001, 2.0, 4500), rng.uniform(2.0, 120.0, 500)]).astype(np.float32)
    worst = {True: 0.0, False: 0.0}
    for b, e in zip(bs, es):
        got = orc.powf(float(b), float(e))
        exact = float(np.float64(b) ** np.float64(e))
        if exact < 2.0 ** -120:
            assert got == 0.0 or abs(got - exact) < 2.0 ** -120
            continue
        small = e <= 2.0
        worst[small] = max(worst[small], abs(got - exact) / exact)
    # exponents of the renderer (100 / (2^(depth+level) quality)) are <= 2 in
    # practice; the error grows with e through e*log2(b)
    assert worst[True] < 1.5e-6, worst
    assert worst[False] < 1e-5, worst
    assert orc.powf(0.0, 0.3) == 0.0 and orc.powf(1.0, 0.3) == 1.0
    assert np.isnan(orc.powf(-0.1, 0.3))


def test_interp_trilinear_reproduction(orc):
    rng = np.random.default_rng(1)
    for trial in range(4):
        ch = datasets.refined_subtree(5, lambda l, locs: rng.random(locs.shape[0]) < 0.35)
        d = orc.subtree_depth(ch, 5)
        n2i = orc.subtree_n2i(ch, d)
        c = rng.uniform(-1, 1, size=8)
        f = lambda x, y, z: (c[0] + c[1] * x + c[2] * y + c[3] * z + c[4] * x * y + c[5] * y * z +
                             c[6] * x * z + c[7] * x * y * z)
        u = n2i.astype(np.float64) / 2 ** d
        dense = orc.interpolate(f(u[:, 0], u[:, 1], u[:, 2]).astype(np.float32), n2i, d)
        n = 2 ** d + 1
        g = np.arange(n) / 2 ** d
        z, y, x = np.meshgrid(g, g, g, indexing="ij")
        assert np.max(np.abs(dense - f(x, y, z).reshape(-1))) < 1e-5


def test_interp_scatter_identity_and_full_fill(orc):
    d = 3
    n2i = orc.subtree_n2i(datasets.full_subtree(d), d)
    vals = np.random.default_rng(0).normal(size=n2i.shape[0]).astype(np.float32)
    dense = orc.interpolate(vals, n2i, d)
    assert np.array_equal(dense[orc.ordinals(n2i, d)], vals)
    assert not np.isnan(dense).any()


def test_interp_nofield_poisons_exactly_its_stencils(orc):
    """Root only (depth 1 subtree of a leaf root): one 1e30 corner makes the
    three adjacent edge midpoints, three faces touching it and the cell NaN."""
    ch = np.zeros((2, 8), np.uint32)
    ch[0, 0] = 1  # one child: depth 1
    n2i = orc.subtree_n2i(ch, 1)
    vals = np.ones(n2i.shape[0], np.float32)
    # node (2,2,2) is the 8th root corner
    k = [tuple(r) for r in n2i.tolist()].index((2, 2, 2))
    vals[k] = np.float32(1e30)
    dense = orc.interpolate(vals, n2i, 1).reshape(3, 3, 3)  # z,y,x
    nan = np.isnan(dense)
    # defined by the child block's corners: all nodes of [0,1]^3 are sparse
    assert not nan[:2, :2, :2].any()
    assert nan[2, 2, 2]
    # edge midpoints adjacent to (2,2,2): (1,2,2),(2,1,2),(2,2,1) (x,y,z)
    for (x, y, z) in [(1, 2, 2), (2, 1, 2), (2, 2, 1)]:
        assert nan[z, y, x]
    # midpoints not touching it stay finite
    for (x, y, z) in [(1, 0, 2), (0, 1, 2), (2, 0, 1)]:
        assert not nan[z, y, x]


def test_sample_count_matches_formula(orc):
    """Closed form of sampleCount for a ray through the block centre."""
    M = datasets.look_at((0.0, 0.0, 30.0))
    live, p = orc.block_params(M, 101, 101, 30.0, 0, (0, 0, 0), 4, 0.0, 1.0, 0.5, 2.0, 2.0)
    assert live
    n = 17
    brick = np.full(n ** 3, 1.0, np.float32)
    tf = np.tile(np.array([[1, 1, 1, 0.01]], np.float32), (8, 1))
    hit, seg = orc.ray_segment(p, brick, tf, 50, 50)
    assert hit
    # ray along -z through the centre: tauEnter/Exit at z=+-5 over |n| = 0.1
    texel = 10.0 / n
    k_expected = int(2.0 * ((35.0 - 25.0) / 0.1) * 0.01 / (texel * 0.1)) - 1
    of, ilc = 0.5, 100.0 / (2 ** 4 * 2.0)
    ap = 1 - (1 - 0.01 * of) ** ilc
    k = np.log(1 - seg[3]) / np.log(1 - ap)
    assert abs(k - k_expected) <= 1
