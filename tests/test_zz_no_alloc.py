"""rv_render_share_argb_noalloc (block table in the library's fixed device
array, no allocation) against rv_render_share_argb (caller workspace) through
the C ABI: bit-exact 8-bit viewports on two calls in a row, the second with
fewer blocks and a smaller viewport, then the first again; more blocks than
the array holds is an argument error."""
import numpy as np
import pytest

from revise_amd import BlockParams, RvError, api, datasets
from tests.gpu_util import say, to_dev

pytestmark = pytest.mark.gpu


def _scene(seed, level, W, H, angle):
    rng = np.random.default_rng(seed)
    field = datasets.Field(seed, n_gauss=3, noise_octaves=1)
    m = 2 ** level
    blocks = []
    for z in range(m):
        for y in range(m):
            for x in range(m):
                depth = int(rng.integers(2, 5))
                ch = datasets.refined_subtree(depth, lambda l, locs: rng.random(locs.shape[0]) < 0.7)
                d = api.subtree_depth(ch, depth)
                n2i = api.subtree_n2i(ch, d)
                blocks.append(datasets.Block(level, (x, y, z), d, n2i,
                                             datasets.sample(field, level, (x, y, z), d, n2i)))
    fmin, fmax = datasets.field_range(blocks)
    tf = to_dev(datasets.seeded_tf(seed, 256))
    M = datasets.orbit_camera(angle, elevation_deg=25.0)
    order = api.back_to_front([0] * 3, [m] * 3, [-5] * 3, [5] * 3, api.eye_from_view(M))
    by = {b.index: b for b in blocks}
    params, bricks = [], []
    for idx in order:
        b = by[tuple(int(v) for v in idx)]
        live, p = api.block_params(M, W, H, 30.0, level, b.index, b.depth, fmin, fmax, 0.5, 2.0, 0.996)
        if not live:
            continue
        import torch
        dense = torch.empty(api.brick_nodes(b.depth), dtype=torch.float32, device="cuda")
        api.interp_brick(to_dev(b.sparse), b.n2i, b.depth, dense)
        params.append(p)
        bricks.append(dense)
    init = rng.integers(0, 2 ** 32, size=(H, W), dtype=np.uint64).astype(np.uint32)
    return params, bricks, tf, init


def _both(params, bricks, tf, init, W, H):
    ref = to_dev(init.view(np.int32))
    api.ShareRenderer(list(zip(params, bricks)), tf, W, H).render(ref)
    got = to_dev(init.view(np.int32))
    api.render_share_argb_noalloc(params, bricks, tf, got, W, H)
    return ref.cpu().numpy(), got.cpu().numpy()


def test_noalloc_share_equals_workspace_share_twice():
    big = _scene(31, 2, 210, 160, 55.0)
    small = _scene(32, 1, 96, 72, 200.0)
    for name, (params, bricks, tf, init), (W, H) in (("64 blocks", big, (210, 160)),
                                                     ("8 blocks", small, (96, 72)),
                                                     ("64 blocks again", big, (210, 160))):
        say("noalloc share:", name, len(params), "live blocks", W, "x", H)
        ref, got = _both(params, bricks, tf, init, W, H)
        assert (ref.view(np.uint32) >> 24).any()
        assert np.array_equal(ref, got)


def test_noalloc_share_rejects_too_many_blocks():
    import torch
    n = api.lib().rv_share_workspace_bytes(0)  # library loaded
    assert n > 0
    p = BlockParams()
    params = [p] * 4097
    bricks = [torch.empty(1, device="cuda")] * 4097
    with pytest.raises(RvError, match="RV_SHARE_STATIC_MAX_BLOCKS"):
        api.render_share_argb_noalloc(params, bricks, torch.empty((1, 4), device="cuda"),
                                      torch.empty(1, dtype=torch.int32, device="cuda"), 1, 1)
