"""Field modes besides Argb (ArgbLight, Isosurface, MaxIntensityProjection)
on the GPU against the oracle, bit for bit:

* per block (rv_render_block, the drop-in for renderKernel<fieldMode>), over a
  viewport holding earlier blocks' pixels, with NaN (no-field) regions in the
  bricks;
* a whole share in one launch (rv_render_share) equals the per-block calls;
* the batched frame (rv_frame_set_field_mode) equals the oracle's float
  composite of the same segments, resident and interpolating;
* the max-intensity part blend and colouring (rv_blend_image_part_max_intensity,
  rv_mip_colorize)."""
import numpy as np
import pytest

from revise_amd import api, datasets
from tests.gpu_util import say, to_dev

pytestmark = pytest.mark.gpu

FOV = 30.0
LEVELS = [0.55, 0.3, 0.42, 0.3]
MODES = [(api.FIELD_ARGB_LIGHT, 0.5, 1.0, None), (api.FIELD_ISOSURFACE, 0.45, 0.7, None),
         (api.FIELD_MIP, 0.3, 1.0, None), (api.FIELD_ISOSURFACES, 0.5, 0.6, LEVELS),
         (api.FIELD_DOMAIN_VOXELS, 0.35, 1.0, None)]


def _scene(level, depth, seed):
    field = datasets.Field(seed, n_gauss=3, noise_octaves=1)
    n2i = api.subtree_n2i(datasets.full_subtree(depth), depth)
    blocks = []
    m = 2 ** level
    for z in range(m):
        for y in range(m):
            for x in range(m):
                # a no-field ball in some blocks: NaN after interpolation
                occ = (lambda u, c=np.array([x + 0.3, y + 0.6, z + 0.5]) / m:
                       ((u - c) ** 2).sum(-1) > (0.25 / m) ** 2)
                blocks.append(datasets.Block(level, (x, y, z), depth, n2i,
                                             datasets.sample(field, level, (x, y, z), depth, n2i, occ)))
    return blocks


def _init_viewport(mode, W, H, seed):
    rng = np.random.default_rng(seed)
    if mode == api.FIELD_MIP:
        v = rng.uniform(0.0, 0.6, size=(H, W)).astype(np.float32)
        v[rng.random((H, W)) < 0.5] = np.nan
        return v.view(np.uint32)
    return rng.integers(0, 2 ** 32, size=(H, W), dtype=np.uint64).astype(np.uint32)


@pytest.mark.parametrize("mode,thr,iso,levels", MODES)
def test_render_block_modes_bit_exact(orc, mode, thr, iso, levels):
    import torch
    W, H = 160, 120
    blocks = _scene(1, 4, seed=21)
    bricks = [orc.interpolate(b.sparse, b.n2i, b.depth) for b in blocks]
    assert any(np.isnan(br).any() for br in bricks)
    fmin, fmax = datasets.field_range(blocks)
    tf = datasets.seeded_tf(21, 256)
    d_tf = to_dev(tf)
    M = datasets.orbit_camera(40.0, elevation_deg=15.0)
    init = _init_viewport(mode, W, H, 5)
    d_vp = to_dev(init.view(np.int32))
    exp = init.copy()
    order = api.mixed_back_to_front(np.array([(1,) + b.index for b in blocks], np.uint32), api.eye_from_view(M))
    n = 0
    for i in order:
        b, brick = blocks[int(i)], bricks[int(i)]
        live, p = api.block_params(M, W, H, FOV, b.level, b.index, b.depth, fmin, fmax, thr, 2.0, 0.996, exact=True)
        lo, po = orc.block_params(M, W, H, FOV, b.level, b.index, b.depth, fmin, fmax, thr, 2.0, 0.996)
        assert live == lo
        if not live:
            continue
        say("mode", mode, "block", b.index)
        api.render_block(p, mode, iso, to_dev(brick), d_tf, d_vp, levels=levels)
        orc.render_block_mode(po, mode, iso, brick, tf, exp.reshape(-1), levels=levels)
        n += 1
    got = d_vp.cpu().numpy().view(np.uint32)
    assert n == 8
    assert not np.array_equal(exp, init)  # the blocks drew something
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("mode,thr,iso,levels", MODES)
def test_render_share_modes_equal_per_block(orc, mode, thr, iso, levels):
    import torch
    W, H = 150, 110
    blocks = _scene(2, 3, seed=22)
    bricks = [orc.interpolate(b.sparse, b.n2i, b.depth) for b in blocks]
    fmin, fmax = datasets.field_range(blocks)
    tf = datasets.seeded_tf(22, 256)
    d_tf = to_dev(tf)
    M = datasets.orbit_camera(250.0, elevation_deg=-20.0)
    order = api.mixed_back_to_front(np.array([(2,) + b.index for b in blocks], np.uint32), api.eye_from_view(M))
    items, exp = [], None
    d_bricks = [to_dev(br) for br in bricks]
    vp_seq = torch.empty((H, W), dtype=torch.int32, device="cuda")
    api.clear_viewport_mode(vp_seq, W, H, mode)
    exp = vp_seq.cpu().numpy().view(np.uint32).copy()
    for i in order:
        b = blocks[int(i)]
        live, p = api.block_params(M, W, H, FOV, b.level, b.index, b.depth, fmin, fmax, thr, 2.0, 0.996, exact=True)
        lo, po = orc.block_params(M, W, H, FOV, b.level, b.index, b.depth, fmin, fmax, thr, 2.0, 0.996)
        if live:
            items.append((p, d_bricks[int(i)]))
            api.render_block(p, mode, iso, d_bricks[int(i)], d_tf, vp_seq, levels=levels)
            orc.render_block_mode(po, mode, iso, bricks[int(i)], tf, exp.reshape(-1), levels=levels)
    share = api.ShareRenderer(items, d_tf, W, H)
    vp = torch.empty((H, W), dtype=torch.int32, device="cuda")
    api.clear_viewport_mode(vp, W, H, mode)
    share.render_mode(vp, mode, iso, levels=levels)
    torch.cuda.synchronize()
    got = vp.cpu().numpy().view(np.uint32)
    assert np.array_equal(got, vp_seq.cpu().numpy().view(np.uint32))
    assert np.array_equal(got, exp)
    if mode == api.FIELD_MIP:
        assert (~np.isnan(got.view(np.float32))).any()
    else:
        assert (got >> 24).any()


@pytest.mark.parametrize("mode,thr,iso,levels", MODES)
def test_frame_modes_match_oracle(orc, mode, thr, iso, levels):
    import torch
    W, H = 144, 108
    blocks = _scene(1, 4, seed=23)
    fmin, fmax = datasets.field_range(blocks)
    tf = datasets.seeded_tf(23, 256)
    d_tf = to_dev(tf)
    plan = api.interp_plan(blocks[0].n2i, 4)
    dev = [{"level": b.level, "index": b.index, "depth": b.depth, "sparse": to_dev(b.sparse), "plan": plan,
            "brick": torch.empty(api.brick_nodes(4), dtype=torch.float32, device="cuda")} for b in blocks]
    fp = api.FramePlan(dev, W, H, FOV, fmin, fmax, thr, 2.0, 0.996, 0.9999, d_tf, exact=True)
    fp.set_field_mode(mode, iso, levels=levels)
    cams = np.stack([datasets.orbit_camera(a, elevation_deg=e) for a, e in ((60.0, 10.0), (300.0, -35.0))])
    fp.set_cameras(cams)
    bricks = [orc.interpolate(b.sparse, b.n2i, b.depth) for b in blocks]
    keys = np.array([(1,) + b.index for b in blocks], np.uint32)
    argb = torch.empty((H, W), dtype=torch.int32, device="cuda")
    rgba = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    for c in range(2):
        order = api.mixed_back_to_front(keys, api.eye_from_view(cams[c]))
        lst = [(1, blocks[i].index, 4, bricks[i]) for i in order]
        exp = orc.render_float_composite(cams[c], W, H, FOV, lst, fmin, fmax, thr, 2.0, 0.996, tf, 0.9999,
                                         mode=mode, iso_opacity=iso, levels=levels)
        for interp in (True, False):
            say("frame mode", mode, "cam", c, "interpolate", interp)
            fp.render(c, argb, rgba, interpolate=interp)
            torch.cuda.synchronize()
            assert fp.status() == 0
            got = rgba.cpu().numpy()
            assert np.array_equal(got, exp, equal_nan=True)
            if mode == api.FIELD_MIP:
                bits = argb.cpu().numpy().view(np.uint32)
                assert np.array_equal(bits.view(np.float32), exp[..., 0], equal_nan=True)
                assert (~np.isnan(exp[..., 0])).mean() > 0.02
            else:
                assert (exp[..., 3] > 0).mean() > 0.02
    # the tape and adjoint are Argb only
    fp.set_tape(True)
    fp.render(0, argb, rgba, interpolate=False)
    from revise_amd import RvError
    with pytest.raises(RvError, match="tape"):
        fp.adjoint(0, rgba, rgba, torch.zeros((256, 4), dtype=torch.float32, device="cuda"))
    fp.close()


def test_max_intensity_compose_kernels(orc):
    import torch
    rng = np.random.default_rng(9)
    dw, dh = 40, 30
    dst = rng.uniform(0, 1, size=(dh, dw)).astype(np.float32)
    dst[rng.random((dh, dw)) < 0.4] = np.nan
    src = rng.uniform(0, 1, size=(12, 17)).astype(np.float32)
    src[rng.random(src.shape) < 0.3] = np.nan
    for just_copy in (False, True):
        d_dst = to_dev(dst)
        api.blend_image_part_max_intensity(d_dst, dw, dh, (2, 3), to_dev(src), 17, 12, (7, 9), just_copy)
        exp = dst.copy()
        orc.blend_image_part_max_intensity(exp, src, 5, 6, just_copy)
        assert np.array_equal(d_dst.cpu().numpy(), exp, equal_nan=True)
    ctv = rng.uniform(0, 1, size=(512, 4)).astype(np.float32)
    src2 = rng.uniform(-0.2, 1.2, size=(12, 17)).astype(np.float32)
    src2[rng.random(src2.shape) < 0.2] = np.nan
    frame = rng.integers(0, 2 ** 32, size=(dh, dw), dtype=np.uint64).astype(np.uint32)
    d_frame = to_dev(frame.view(np.int32))
    api.mip_colorize(d_frame, dw, dh, to_dev(src2), 17, 12, (7, 9), to_dev(ctv))
    exp = frame.copy()
    orc.mip_colorize(exp, src2, 7, 9, ctv)
    assert np.array_equal(d_frame.cpu().numpy().view(np.uint32), exp)


def test_isosurfaces_without_levels(orc):
    """No levels: every pixel a block covers becomes 0 (per block and share);
    the frame draws nothing."""
    import torch
    W, H = 96, 72
    blocks = _scene(1, 3, seed=24)
    bricks = [orc.interpolate(b.sparse, b.n2i, b.depth) for b in blocks]
    fmin, fmax = datasets.field_range(blocks)
    tf = datasets.seeded_tf(24, 64)
    d_tf = to_dev(tf)
    M = datasets.orbit_camera(10.0)
    init = _init_viewport(api.FIELD_ARGB, W, H, 3)
    d_vp = to_dev(init.view(np.int32))
    exp = init.copy()
    for b, br in zip(blocks, bricks):
        live, p = api.block_params(M, W, H, FOV, b.level, b.index, b.depth, fmin, fmax, 0.5, 2.0, 0.996, exact=True)
        lo, po = orc.block_params(M, W, H, FOV, b.level, b.index, b.depth, fmin, fmax, 0.5, 2.0, 0.996)
        if live:
            api.render_block(p, api.FIELD_ISOSURFACES, 1.0, to_dev(br), d_tf, d_vp, levels=[])
            orc.render_block_mode(po, api.FIELD_ISOSURFACES, 1.0, br, tf, exp.reshape(-1), levels=[])
    got = d_vp.cpu().numpy().view(np.uint32)
    assert np.array_equal(got, exp)
    assert (got == 0).sum() > (init == 0).sum()


VALUE_MODES = [(api.FIELD_VALUE_ON_ISOSURFACE, 0.45, 0.7, None),
               (api.FIELD_VALUE_ON_ISOSURFACES, 0.5, 0.6, LEVELS)]


def _secondary(blocks, seed):
    """A second field over the same blocks (same n2i), with its own no-field ball."""
    field = datasets.Field(seed, n_gauss=4, noise_octaves=2)
    out = []
    for b in blocks:
        m = 2 ** b.level
        c = np.array([b.index[0] + 0.7, b.index[1] + 0.4, b.index[2] + 0.5]) / m
        occ = lambda u, c=c: ((u - c) ** 2).sum(-1) > (0.2 / m) ** 2
        out.append(datasets.Block(b.level, b.index, b.depth, b.n2i,
                                  datasets.sample(field, b.level, b.index, b.depth, b.n2i, occ)))
    return out


@pytest.mark.parametrize("mode,thr,iso,levels", VALUE_MODES)
def test_value_modes_block_and_share_bit_exact(orc, mode, thr, iso, levels):
    """ValueOnIsosurface(s): per block (rv_render_block_secondary) over a
    viewport holding earlier pixels, and a whole share
    (rv_render_share_secondary), against the oracle, bit for bit."""
    import torch
    W, H = 160, 120
    blocks = _scene(2, 3, seed=24)
    sec = _secondary(blocks, 77)
    bricks = [orc.interpolate(b.sparse, b.n2i, b.depth) for b in blocks]
    bricks2 = [orc.interpolate(b.sparse, b.n2i, b.depth) for b in sec]
    assert any(np.isnan(br).any() for br in bricks2)
    fmin, fmax = datasets.field_range(blocks)
    smin, smax = datasets.field_range(sec)
    tf = datasets.seeded_tf(24, 256)
    d_tf = to_dev(tf)
    M = datasets.orbit_camera(130.0, elevation_deg=20.0)
    init = _init_viewport(mode, W, H, 6)
    d_vp = to_dev(init.view(np.int32))
    exp = init.copy()
    order = api.mixed_back_to_front(np.array([(2,) + b.index for b in blocks], np.uint32), api.eye_from_view(M))
    d_b = [to_dev(br) for br in bricks]
    d_b2 = [to_dev(br) for br in bricks2]
    items, items2 = [], []
    for i in order:
        i = int(i)
        b = blocks[i]
        live, p = api.block_params(M, W, H, FOV, b.level, b.index, b.depth, fmin, fmax, thr, 2.0, 0.996, exact=True)
        lo, po = orc.block_params(M, W, H, FOV, b.level, b.index, b.depth, fmin, fmax, thr, 2.0, 0.996)
        assert live == lo
        if not live:
            continue
        say("value mode", mode, "block", b.index)
        api.render_block(p, mode, iso, d_b[i], d_tf, d_vp, levels=levels, d_brick2=d_b2[i], srange=(smin, smax))
        orc.render_block_mode(po, mode, iso, bricks[i], tf, exp.reshape(-1), levels=levels, brick2=bricks2[i],
                              srange=(smin, smax))
        items.append((p, d_b[i]))
        items2.append(d_b2[i])
    got = d_vp.cpu().numpy().view(np.uint32)
    assert not np.array_equal(exp, init)
    assert np.array_equal(got, exp)
    # the share, from a cleared viewport, equals the per-block calls from one
    share = api.ShareRenderer(items, d_tf, W, H, bricks2=items2)
    vp = torch.zeros((H, W), dtype=torch.int32, device="cuda")
    share.render_mode(vp, mode, iso, levels=levels, srange=(smin, smax))
    vp_seq = torch.zeros((H, W), dtype=torch.int32, device="cuda")
    for (p, br), br2 in zip(items, items2):
        api.render_block(p, mode, iso, br, d_tf, vp_seq, levels=levels, d_brick2=br2, srange=(smin, smax))
    torch.cuda.synchronize()
    assert (vp.cpu().numpy().view(np.uint32) >> 24).any()
    assert torch.equal(vp, vp_seq)


@pytest.mark.parametrize("fp16", [False, True])
@pytest.mark.parametrize("mode,thr,iso,levels", VALUE_MODES)
def test_frame_value_modes_match_oracle(orc, mode, thr, iso, levels, fp16):
    """ValueOnIsosurface(s) in the batched frame: the secondary bricks beside
    the primary ones, interpolated in the same launch (both fields' tiles) or
    resident; the float image equals the oracle's composite of the same
    segments."""
    import torch
    W, H = 144, 108
    blocks = _scene(1, 4, seed=25)
    sec = _secondary(blocks, 26)
    fmin, fmax = datasets.field_range(blocks)
    smin, smax = datasets.field_range(sec)
    tf = datasets.seeded_tf(25, 256)
    d_tf = to_dev(tf)
    plan = api.interp_plan(blocks[0].n2i, 4)
    bt = torch.float16 if fp16 else torch.float32
    dev = [{"level": b.level, "index": b.index, "depth": b.depth, "sparse": to_dev(b.sparse), "plan": plan,
            "brick": torch.empty(api.brick_nodes(4), dtype=bt, device="cuda")} for b in blocks]
    fp = api.FramePlan(dev, W, H, FOV, fmin, fmax, thr, 2.0, 0.996, 0.9999, d_tf, exact=True)
    from revise_amd import RvError
    with pytest.raises(RvError, match="secondary"):
        fp.set_field_mode(mode, iso, levels=levels)
    sparse2 = [to_dev(b.sparse) for b in sec]
    bricks2 = [torch.empty(api.brick_nodes(4), dtype=bt, device="cuda") for _ in sec]
    fp.set_secondary(sparse2, bricks2, (smin, smax))
    fp.set_field_mode(mode, iso, levels=levels)
    cams = np.stack([datasets.orbit_camera(a, elevation_deg=e) for a, e in ((40.0, 15.0), (220.0, -25.0))])
    fp.set_cameras(cams)
    cast = (lambda x: x.astype(np.float16).astype(np.float32)) if fp16 else (lambda x: x)
    bricks = [cast(orc.interpolate(b.sparse, b.n2i, b.depth)) for b in blocks]
    bricks_s = [cast(orc.interpolate(b.sparse, b.n2i, b.depth)) for b in sec]
    keys = np.array([(1,) + b.index for b in blocks], np.uint32)
    argb = torch.empty((H, W), dtype=torch.int32, device="cuda")
    rgba = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    for c in range(2):
        order = api.mixed_back_to_front(keys, api.eye_from_view(cams[c]))
        lst = [(1, blocks[i].index, 4, bricks[i], bricks_s[i]) for i in order]
        exp = orc.render_float_composite(cams[c], W, H, FOV, lst, fmin, fmax, thr, 2.0, 0.996, tf, 0.9999,
                                         mode=mode, iso_opacity=iso, levels=levels, srange=(smin, smax))
        assert (exp[..., 3] > 0).mean() > 0.01  # the surfaces cover part of the image
        for interp in (True, False):
            say("frame value mode", mode, "fp16", fp16, "cam", c, "interpolate", interp)
            fp.render(c, argb, rgba, interpolate=interp)
            torch.cuda.synchronize()
            assert fp.status() == 0
            assert np.array_equal(rgba.cpu().numpy(), exp)
    fp.close()


@pytest.mark.parametrize("mode,thr,iso,levels", MODES + VALUE_MODES + [(api.FIELD_ARGB, 0.5, 1.0, None)])
def test_render_share_tiled_bricks_equal_linear(orc, mode, thr, iso, levels):
    """The share kernel reading 4^3 micro-brick bricks (interpolated in that
    layout) gives the same viewport as with x-fastest bricks, in every mode."""
    import torch
    W, H = 150, 110
    blocks = _scene(2, 3, seed=27)
    sec = _secondary(blocks, 28)
    fmin, fmax = datasets.field_range(blocks)
    smin, smax = datasets.field_range(sec)
    tf = datasets.seeded_tf(27, 256)
    d_tf = to_dev(tf)
    M = datasets.orbit_camera(75.0, elevation_deg=-10.0)
    order = api.mixed_back_to_front(np.array([(2,) + b.index for b in blocks], np.uint32), api.eye_from_view(M))
    plan = api.interp_plan(blocks[0].n2i, 3)
    value = mode in (api.FIELD_VALUE_ON_ISOSURFACE, api.FIELD_VALUE_ON_ISOSURFACES)
    out = {}
    for layout in (api.BRICK_LINEAR, api.BRICK_TILED4):
        nf = api.brick_floats(3, layout)
        br = [torch.empty(nf, dtype=torch.float32, device="cuda") for _ in blocks]
        br2 = [torch.empty(nf, dtype=torch.float32, device="cuda") for _ in blocks]
        jobs = [(to_dev(b.sparse), plan, br[i], 3) for i, b in enumerate(blocks)]
        jobs += [(to_dev(b.sparse), plan, br2[i], 3) for i, b in enumerate(sec)]
        api.InterpBatch(jobs, layout).run()
        items, items2 = [], []
        for i in order:
            b = blocks[int(i)]
            live, p = api.block_params(M, W, H, FOV, b.level, b.index, b.depth, fmin, fmax, thr, 2.0, 0.996, exact=True)
            if live:
                items.append((p, br[int(i)]))
                items2.append(br2[int(i)])
        share = api.ShareRenderer(items, d_tf, W, H, bricks2=items2 if value else None, layout=layout)
        vp = torch.empty((H, W), dtype=torch.int32, device="cuda")
        api.clear_viewport_mode(vp, W, H, mode)
        say("share layout", layout, "mode", mode)
        share.render_mode(vp, mode, iso, levels=levels, srange=(smin, smax))
        torch.cuda.synchronize()
        out[layout] = vp.cpu().numpy().view(np.uint32)
    assert np.array_equal(out[api.BRICK_LINEAR], out[api.BRICK_TILED4])
    if mode == api.FIELD_MIP:
        assert (~np.isnan(out[api.BRICK_LINEAR].view(np.float32))).any()
    else:
        assert (out[api.BRICK_LINEAR] >> 24).any()
