"""The fast Argb sample sequence (the default; rv_block_params.flags without
RV_RENDER_EXACT) against the oracle's exact sequence.

Colour is compared premultiplied by alpha.  Every later use of an ARGB
pixel weighs its colour by its alpha (blendAbove, blendImagePart, which also
skips alpha-0 pixels, and the final compose over the background), so where
alpha is ~0 the stored colour carries no weight and is ill-conditioned: a
segment whose opacity correction rounds one ulp of 1 apart (alpha 6e-8
instead of 0) brings its colour whole into a transparent pixel.

Bars:
* float RGBA (the batched frame's float output, which composites the same
  per-block segments as the oracle's float composite): within 2e-3 absolute
  per channel (premultiplied colour and alpha), the north-star tolerance;
* 8-bit ARGB of one block over a viewport: within 1 LSB per channel.  A
  segment's float colour moves by a few 1e-6 (or by at most (1 - cutoff) x
  one sample's alpha where early termination lands one sample apart), so the
  encode rounds at most one step differently;
* the per-block 8-bit paths of several blocks requantise after every block, so
  such 1-LSB steps can add up along a ray: within 2 LSB per channel, and
  fewer than 1% of the channels off at all;
* the fast share kernel against the fast per-block launches, and the worker's
  fused path against its per-block path: bit-exact (the same operations).
"""
import numpy as np
import pytest

from revise_amd import api, datasets
from tests.gpu_util import oracle_blocks_frame, say, to_dev

pytestmark = pytest.mark.gpu

FOV, THR, Q = 30.0, 0.5, 2.0


def _blocks(seed, level, depth, p_refine, nan_holes=False):
    rng = np.random.default_rng(seed)
    field = datasets.Field(seed, n_gauss=3, noise_octaves=1)
    m = 2 ** level
    out = []
    for z in range(m):
        for y in range(m):
            for x in range(m):
                ch = datasets.refined_subtree(depth, lambda l, locs: rng.random(locs.shape[0]) < p_refine)
                d = api.subtree_depth(ch, depth)
                n2i = api.subtree_n2i(ch, d)
                occ = None
                if nan_holes:
                    cen = rng.uniform(0, 1, size=3)
                    rad = float(rng.uniform(0.4, 1.2))
                    def occ(u, L=level, x=x, y=y, z=z, cen=cen, rad=rad):
                        lu = u * 2 ** L - np.array([x, y, z])
                        return ((lu - cen) ** 2).sum(-1) < rad * rad
                out.append(datasets.Block(level, (x, y, z), d, n2i,
                                          datasets.sample(field, level, (x, y, z), d, n2i, occ)))
    return out


def premul8(img):
    """ARGB8 pixels as (alpha, colour * alpha / 255) in 8-bit steps, float."""
    a = np.asarray(img).view(np.uint32).reshape(-1)
    al = ((a >> 24) & 255).astype(np.float64)
    return np.stack([al] + [((a >> s) & 255) * al / 255.0 for s in (0, 8, 16)])


def _lsb_diff(a, b):
    """Per-channel difference of two ARGB8 images, premultiplied, in 8-bit
    steps (rounded up)."""
    return np.ceil(np.abs(premul8(a) - premul8(b)) - 1e-9).astype(np.int32)


def premul(rgba):
    """Float RGBA (non-premultiplied) as (r a, g a, b a, a)."""
    rgba = np.asarray(rgba, dtype=np.float64)
    return np.concatenate([rgba[..., :3] * rgba[..., 3:4], rgba[..., 3:4]], axis=-1)


@pytest.mark.parametrize("cut", [0.996, 0.99])
def test_fast_single_block_within_one_lsb(orc, cut):
    W, H = 200, 150
    blocks = _blocks(41, 0, 5, 0.5, nan_holes=True)
    fmin, fmax = datasets.field_range(blocks)
    tf = datasets.seeded_tf(41, 256)
    b = blocks[0]
    brick = orc.interpolate(b.sparse, b.n2i, b.depth)
    init = np.random.default_rng(2).integers(0, 2 ** 32, size=(H, W), dtype=np.uint64).astype(np.uint32)
    for angle in (20.0, 130.0, 250.0):
        M = datasets.orbit_camera(angle, elevation_deg=30.0)
        live, p = api.block_params(M, W, H, FOV, 0, (0, 0, 0), b.depth, fmin, fmax, THR, Q, cut)
        _, po = orc.block_params(M, W, H, FOV, 0, (0, 0, 0), b.depth, fmin, fmax, THR, Q, cut)
        assert live and p.flags == 0
        d_vp = to_dev(init.view(np.int32))
        api.render_block_argb(p, to_dev(brick), to_dev(tf), d_vp)
        exp = init.copy()
        orc.render_block(po, brick, tf, exp)
        d = _lsb_diff(d_vp.cpu().numpy(), exp)
        say("angle", angle, "cut", cut, "max LSB", int(d.max()), "channels off %.4f%%" % (100 * (d > 0).mean()))
        assert d.max() <= 1
        assert (d > 0).mean() < 0.01


@pytest.mark.parametrize("layout", [api.BRICK_LINEAR, api.BRICK_TILED4])
def test_fast_share_equals_fast_per_block(orc, layout):
    import torch
    W, H = 230, 170
    blocks = _blocks(42, 2, 4, 0.6, nan_holes=True)
    fmin, fmax = datasets.field_range(blocks)
    tf = to_dev(datasets.seeded_tf(42, 256))
    M = datasets.orbit_camera(77.0, elevation_deg=-30.0)
    order = api.back_to_front([0] * 3, [4] * 3, [-5] * 3, [5] * 3, api.eye_from_view(M))
    by = {b.index: b for b in blocks}
    lin, til = [], []
    for idx in order:
        b = by[tuple(int(v) for v in idx)]
        live, p = api.block_params(M, W, H, FOV, 2, b.index, b.depth, fmin, fmax, THR, Q, 0.996)
        if not live:
            continue
        dense = orc.interpolate(b.sparse, b.n2i, b.depth)
        lin.append((p, to_dev(dense)))
        t = torch.empty(api.brick_floats(b.depth, api.BRICK_TILED4), dtype=torch.float32, device="cuda")
        plan = api.interp_plan(b.n2i, b.depth)
        api.InterpBatch([(to_dev(b.sparse), plan, t, b.depth)], api.BRICK_TILED4).run()
        til.append((p, t))
    init = np.random.default_rng(3).integers(0, 2 ** 32, size=(H, W), dtype=np.uint64).astype(np.uint32)
    a = to_dev(init.view(np.int32))
    for p, br in lin:
        api.render_block_argb(p, br, tf, a)
    b_ = to_dev(init.view(np.int32))
    items = til if layout == api.BRICK_TILED4 else lin
    api.ShareRenderer(items, tf, W, H, layout=layout).render(b_)
    assert np.array_equal(a.cpu().numpy(), b_.cpu().numpy())


def test_fast_frame_float_within_2e3_and_8bit_share(orc):
    """One scene, three renderings: the batched frame's float image against
    the oracle's float composite (2e-3), its 8-bit image against the
    oracle's encoding of that composite, and the share kernel (per-block 8-bit
    composite) against the oracle's per-block frame."""
    import torch
    W, H = 256, 192
    blocks = _blocks(43, 2, 4, 0.55, nan_holes=True)
    fmin, fmax = datasets.field_range(blocks)
    tf_np = datasets.seeded_tf(43, 256)
    tf = to_dev(tf_np)
    bricks = [orc.interpolate(b.sparse, b.n2i, b.depth) for b in blocks]
    dev = [{"level": b.level, "index": b.index, "depth": b.depth, "brick": to_dev(br)}
           for b, br in zip(blocks, bricks)]
    for angle, elev in ((35.0, 20.0), (200.0, -40.0)):
        M = datasets.orbit_camera(angle, elevation_deg=elev)
        keys = np.array([(b.level,) + b.index for b in blocks], dtype=np.uint32)
        order = api.mixed_back_to_front(keys, api.eye_from_view(M))
        lst = [(blocks[i].level, blocks[i].index, blocks[i].depth, bricks[i]) for i in order]
        exp = orc.render_float_composite(M, W, H, FOV, lst, fmin, fmax, THR, Q, 0.996, tf_np, 0.9999)
        fp = api.FramePlan(dev, W, H, FOV, fmin, fmax, THR, Q, 0.996, 0.9999, tf)
        fp.set_cameras(M[None])
        argb = torch.empty((H, W), dtype=torch.int32, device="cuda")
        rgba = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
        fp.render(0, argb, rgba, interpolate=False)
        got = rgba.cpu().numpy()
        err = np.abs(premul(got) - premul(exp))
        say("angle", angle, "float max %.2e mean %.2e" % (err.max(), err.mean()))
        assert err.max() <= 2e-3
        enc = (np.minimum(exp, 1.0) * np.float32(255.999)).astype(np.uint32)
        enc = (enc[..., 3] << 24) | (enc[..., 0] << 16) | (enc[..., 1] << 8) | enc[..., 2]
        d8 = _lsb_diff(argb.cpu().numpy(), enc.reshape(H, W))
        assert d8.max() <= 1
        fp.close()
        # the share kernel's per-block 8-bit composite, fast, against the oracle's exact one
        items, oitems = [], []
        for i in order:
            b = blocks[i]
            live, p = api.block_params(M, W, H, FOV, b.level, b.index, b.depth, fmin, fmax, THR, Q, 0.996)
            if live:
                items.append((p, dev[i]["brick"]))
                oitems.append((orc.block_params(M, W, H, FOV, b.level, b.index, b.depth, fmin, fmax, THR, Q,
                                                0.996)[1], bricks[i]))
        vp = torch.zeros((H, W), dtype=torch.int32, device="cuda")
        api.ShareRenderer(items, tf, W, H).render(vp)
        ovp = oracle_blocks_frame(orc, oitems, tf_np, W, H)
        d = _lsb_diff(vp.cpu().numpy(), ovp)
        say("share 8-bit: max LSB", int(d.max()), "channels off %.4f%%" % (100 * (d > 0).mean()))
        assert d.max() <= 2
        assert (d > 0).mean() < 0.01


def test_fast_worker_fused_equals_per_block():
    W, H = 220, 160
    blocks = _blocks(44, 2, 4, 0.5)
    fmin, fmax = datasets.field_range(blocks)
    tf = datasets.seeded_tf(44, 256)
    M = datasets.orbit_camera(300.0)
    parts = {}
    for fused in (True, False):
        w = api.Worker(0)
        for b in blocks:
            w.add_subtree(2, b.index, b.depth, b.n2i, b.sparse)
        w.set_state(M, W, H, FOV, Q, THR, 0.996, fmin, fmax, (0.1, 0.2, 0.3), tf)
        w.set_fused(fused)
        parts[fused] = w.render_scene_part(2, (0, 0, 0), (4, 4, 4))
        w.close()
    assert parts[True][0] == parts[False][0]
    assert np.array_equal(parts[True][1], parts[False][1])


@pytest.mark.parametrize("case", ["deep brick", "close camera", "camera inside"])
def test_fast_share_equals_per_block_edge_cases(orc, case):
    """The share kernel against per-block launches where the rays are
    unusual: a deep brick (cells far below a pixel), a camera close to the
    blocks, a camera inside the volume (rays in every direction).  Same
    samples and operations: equal bit for bit."""
    import torch
    W, H = 200, 150
    if case == "deep brick":
        level, depth, M = 0, 7, datasets.orbit_camera(33.0, elevation_deg=25.0)
    elif case == "close camera":
        level, depth, M = 1, 5, datasets.look_at((3.0, 2.5, 6.5), (0.0, -0.5, 0.0))
    else:
        level, depth, M = 1, 4, datasets.look_at((0.3, -0.4, 0.2), (3.0, 1.0, -2.0))
    blocks = _blocks(45, level, depth, 0.7)
    fmin, fmax = datasets.field_range(blocks)
    tf = to_dev(datasets.seeded_tf(45, 256))
    m = 2 ** level
    order = api.back_to_front([0] * 3, [m] * 3, [-5] * 3, [5] * 3, api.eye_from_view(M))
    by = {b.index: b for b in blocks}
    items = []
    for idx in order:
        b = by[tuple(int(v) for v in idx)]
        live, p = api.block_params(M, W, H, FOV, level, b.index, b.depth, fmin, fmax, THR, Q, 0.996)
        if live:
            items.append((p, to_dev(orc.interpolate(b.sparse, b.n2i, b.depth))))
    assert items
    a = torch.zeros((H, W), dtype=torch.int32, device="cuda")
    for p, br in items:
        api.render_block_argb(p, br, tf, a)
    b_ = torch.zeros((H, W), dtype=torch.int32, device="cuda")
    api.ShareRenderer(items, tf, W, H).render(b_)
    got = a.cpu().numpy()
    say(case, "covered pixels", int(((got.view(np.uint32) >> 24) > 0).sum()))
    assert ((got.view(np.uint32) >> 24) > 0).any()
    assert np.array_equal(got, b_.cpu().numpy())
