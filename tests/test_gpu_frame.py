"""Batched persistent frame (one launch per frame: brick interpolation +
front-to-back compositing of all blocks) against the oracle's float
composite of the same per-block segments (bit-exact float RGBA), with
uniform and mixed block levels, fp16 bricks, and graph replays."""
import numpy as np
import pytest

from revise_amd import api, datasets
from tests.gpu_util import say, to_dev

pytestmark = pytest.mark.gpu

FOV, THR, Q, CUT = 30.0, 0.5, 2.0, 0.996


def _uniform_blocks(level, depth, seed):
    field = datasets.Field(seed, n_gauss=3, noise_octaves=2)
    n2i = api.subtree_n2i(datasets.full_subtree(depth), depth)
    return datasets.uniform_level(level, depth, field, n2i)


def _mixed_blocks(seed, depth=3, n_splits=12):
    rng = np.random.default_rng(seed)
    keys = {(2, x, y, z) for z in range(4) for y in range(4) for x in range(4)}
    for _ in range(n_splits):
        cand = sorted(b for b in keys if b[0] < 4)
        b = cand[int(rng.integers(len(cand)))]
        keys.remove(b)
        L, x, y, z = b
        for k in range(8):
            keys.add((L + 1, 2 * x + (k & 1), 2 * y + ((k >> 1) & 1), 2 * z + ((k >> 2) & 1)))
    field = datasets.Field(seed, n_gauss=4, noise_octaves=1)
    blocks = []
    for (L, x, y, z) in sorted(keys):
        p = float(rng.uniform(0.2, 1.0))
        ch = datasets.refined_subtree(depth, lambda l, locs: rng.random(locs.shape[0]) < p)
        d = api.subtree_depth(ch, depth)
        n2i = api.subtree_n2i(ch, d)
        cen = rng.uniform(0, 1, size=3)
        rad = float(rng.uniform(0.3, 1.5))
        def occ(u, L=L, x=x, y=y, z=z, cen=cen, rad=rad):
            lu = u * 2 ** L - np.array([x, y, z])
            return ((lu - cen) ** 2).sum(-1) < rad * rad
        blocks.append(datasets.Block(L, (x, y, z), d, n2i,
                                     datasets.sample(field, L, (x, y, z), d, n2i, occ)))
    return blocks


def _setup(blocks, W, H, fp16=False, acc_cut=2.0, seed=0, tile=(0, 0), layout=api.BRICK_LINEAR):
    import torch
    fmin, fmax = datasets.field_range(blocks)
    tf = datasets.seeded_tf(seed, 256)
    d_tf = to_dev(tf)
    dev_blocks = []
    plans = {}
    for b in blocks:
        key = (id(b.n2i), b.depth)
        if key not in plans:
            plans[key] = api.interp_plan(b.n2i, b.depth)
        brick = torch.empty(api.brick_floats(b.depth, layout), dtype=torch.float16 if fp16 else torch.float32,
                            device="cuda")
        dev_blocks.append({"level": b.level, "index": b.index, "depth": b.depth, "brick": brick,
                           "sparse": to_dev(b.sparse), "plan": plans[key]})
    fp = api.FramePlan(dev_blocks, W, H, FOV, fmin, fmax, THR, Q, CUT, acc_cut, d_tf, tile, layout, exact=True)
    return fp, tf, fmin, fmax, dev_blocks


def _oracle_frame(orc, blocks, bricks_np, M, W, H, fmin, fmax, tf, acc_cut):
    keys = np.array([(b.level,) + b.index for b in blocks], dtype=np.uint32)
    order = api.mixed_back_to_front(keys, api.eye_from_view(M))
    lst = [(blocks[i].level, blocks[i].index, blocks[i].depth, bricks_np[i]) for i in order]
    return orc.render_float_composite(M, W, H, FOV, lst, fmin, fmax, THR, Q, CUT, tf, acc_cut)


SCHEDULES = [api.SCHEDULE_BINS, api.SCHEDULE_LANES, api.SCHEDULE_HYBRID]


def _set_schedule(fp, schedule):
    """The hybrid mixes both walks in one frame: a low threshold sends the
    crowded bins to the per-lane walk."""
    fp.set_schedule(schedule)
    if schedule == api.SCHEDULE_HYBRID:
        fp.set_hybrid_min(5)


@pytest.mark.parametrize("schedule", SCHEDULES)
@pytest.mark.parametrize("fp16,acc_cut", [(False, 2.0), (False, 0.9999), (True, 0.9999)])
def test_frame_uniform_matches_oracle(orc, fp16, acc_cut, schedule):
    import torch
    blocks = _uniform_blocks(2, 4, seed=5)
    W, H = 200, 150
    fp, tf, fmin, fmax, dev = _setup(blocks, W, H, fp16, acc_cut, seed=5)
    _set_schedule(fp, schedule)
    assert fp.schedule() == schedule
    cams = np.stack([datasets.orbit_camera(a) for a in (15.0, 140.0)])
    fp.set_cameras(cams)
    argb = torch.empty((H, W), dtype=torch.int32, device="cuda")
    rgba = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    for c in range(2):
        say("frame uniform cam", c, "fp16", fp16, "schedule", schedule)
        fp.render(c, argb, rgba, interpolate=True)
        torch.cuda.synchronize()
        assert fp.status() == 0
        bricks = [orc.interpolate(b.sparse, b.n2i, b.depth) for b in blocks]
        if fp16:
            bricks = [x.astype(np.float16).astype(np.float32) for x in bricks]
        exp = _oracle_frame(orc, blocks, bricks, cams[c], W, H, fmin, fmax, tf, acc_cut)
        got = rgba.cpu().numpy()
        assert np.array_equal(got, exp)
        a = argb.cpu().numpy().view(np.uint32)
        q = lambda v: (v * np.float32(255.999)).astype(np.uint32)
        assert np.array_equal(a, (q(exp[..., 3]) << 24) + (q(exp[..., 0]) << 16) + (q(exp[..., 1]) << 8) + q(exp[..., 2]))
        assert (a >> 24).any()


@pytest.mark.parametrize("layout", [api.BRICK_LINEAR, api.BRICK_TILED4])
@pytest.mark.parametrize("schedule", SCHEDULES)
def test_frame_mixed_levels_matches_oracle(orc, schedule, layout):
    """Mixed levels and depths (0-6), both brick layouts, every schedule."""
    import torch
    blocks = _mixed_blocks(seed=8)
    say("mixed blocks", len(blocks), "schedule", schedule, "layout", layout)
    W, H = 180, 140
    fp, tf, fmin, fmax, dev = _setup(blocks, W, H, acc_cut=0.9999, seed=8, layout=layout)
    _set_schedule(fp, schedule)
    cams = np.stack([datasets.orbit_camera(a, elevation_deg=e) for a, e in ((70.0, 25.0), (250.0, -40.0))])
    fp.set_cameras(cams)
    argb = torch.empty((H, W), dtype=torch.int32, device="cuda")
    rgba = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    bricks = [orc.interpolate(b.sparse, b.n2i, b.depth) for b in blocks]
    for c in range(2):
        fp.render(c, argb, rgba, interpolate=True)
        torch.cuda.synchronize()
        assert fp.status() == 0
        exp = _oracle_frame(orc, blocks, bricks, cams[c], W, H, fmin, fmax, tf, 0.9999)
        assert np.array_equal(rgba.cpu().numpy(), exp)
        # resident bricks give the same frame
        fp.render(c, argb, rgba, interpolate=False)
        torch.cuda.synchronize()
        assert np.array_equal(rgba.cpu().numpy(), exp)


@pytest.mark.parametrize("layout,fp16", [(api.BRICK_LINEAR, False), (api.BRICK_TILED4, False),
                                         (api.BRICK_LINEAR, True)])
@pytest.mark.parametrize("schedule", [api.SCHEDULE_BINS, api.SCHEDULE_HYBRID])
def test_frame_split_launches_match_oracle(orc, schedule, layout, fp16):
    """Interpolating frames split in two launches (the bricks, then the
    resident-brick kernel): bricks cleared before every frame are rebuilt,
    the images equal the oracle's and the one-launch frame's, also when
    replayed from a graph."""
    import torch
    blocks = _mixed_blocks(seed=12)
    W, H = 170, 130
    fp, tf, fmin, fmax, dev = _setup(blocks, W, H, fp16=fp16, acc_cut=0.9999, seed=12, layout=layout)
    _set_schedule(fp, schedule)
    assert not fp.split()
    cams = np.stack([datasets.orbit_camera(a, elevation_deg=e) for a, e in ((30.0, 10.0), (200.0, -30.0))])
    fp.set_cameras(cams)
    argb = torch.empty((H, W), dtype=torch.int32, device="cuda")
    rgba = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    bricks = [orc.interpolate(b.sparse, b.n2i, b.depth) for b in blocks]
    if fp16:
        bricks = [x.astype(np.float16).astype(np.float32) for x in bricks]
    one = []
    for c in range(2):
        fp.render(c, argb, rgba, interpolate=True)
        one.append(argb.cpu().numpy().copy())
    fp.set_split(True)
    assert fp.split()
    for c in range(2):
        say("split frame cam", c, "schedule", schedule, "layout", layout, "fp16", fp16)
        for b in dev:
            b["brick"].zero_()
        fp.render(c, argb, rgba, interpolate=True)
        torch.cuda.synchronize()
        assert fp.status() == 0
        exp = _oracle_frame(orc, blocks, bricks, cams[c], W, H, fmin, fmax, tf, 0.9999)
        assert np.array_equal(rgba.cpu().numpy(), exp)
        assert np.array_equal(argb.cpu().numpy(), one[c])
    frames = torch.zeros((2, H, W), dtype=torch.int32, device="cuda")
    fp.graph_build(frames, H * W, interpolate=True)
    for rep in range(2):
        for b in dev:
            b["brick"].zero_()
        frames.zero_()
        fp.graph_launch()
        torch.cuda.synchronize()
        got = frames.cpu().numpy()
        for c in range(2):
            assert np.array_equal(got[c], one[c]), (rep, c)


@pytest.mark.parametrize("fp16", [False, True])
def test_frame_split_small_tiles_match_one_launch(orc, fp16):
    """Split frames of depth-4 bricks, whose tiles all fit the small LDS
    class (256-thread workgroups): the bricks are rebuilt and the image
    equals the oracle's and the one-launch frame's."""
    import torch
    blocks = _uniform_blocks(2, 4, seed=14)
    W, H = 150, 120
    fp, tf, fmin, fmax, dev = _setup(blocks, W, H, fp16=fp16, acc_cut=0.9999, seed=14)
    fp.set_cameras(datasets.orbit_camera(110.0, elevation_deg=20.0)[None])
    argb = torch.empty((H, W), dtype=torch.int32, device="cuda")
    rgba = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    fp.render(0, argb, rgba, interpolate=True)
    one = argb.cpu().numpy().copy()
    fp.set_split(True)
    for b in dev:
        b["brick"].zero_()
    say("split frame, depth-4 bricks, fp16", fp16)
    fp.render(0, argb, rgba, interpolate=True)
    torch.cuda.synchronize()
    assert fp.status() == 0
    bricks = [orc.interpolate(b.sparse, b.n2i, b.depth) for b in blocks]
    if fp16:
        bricks = [x.astype(np.float16).astype(np.float32) for x in bricks]
    exp = _oracle_frame(orc, blocks, bricks, datasets.orbit_camera(110.0, elevation_deg=20.0), W, H, fmin, fmax,
                        tf, 0.9999)
    assert np.array_equal(rgba.cpu().numpy(), exp)
    assert np.array_equal(argb.cpu().numpy(), one)


@pytest.mark.parametrize("schedule", SCHEDULES)
def test_frame_graph_replays_match_eager(orc, schedule):
    import torch
    blocks = _uniform_blocks(2, 4, seed=6)
    W, H = 160, 120
    fp, tf, fmin, fmax, dev = _setup(blocks, W, H, acc_cut=0.9999, seed=6)
    _set_schedule(fp, schedule)
    cams = np.stack([datasets.orbit_camera(a) for a in (0.0, 45.0, 90.0)])
    fp.set_cameras(cams)
    eager = []
    one = torch.empty((H, W), dtype=torch.int32, device="cuda")
    for c in range(3):
        fp.render(c, one, None, interpolate=True)
        eager.append(one.cpu().numpy().copy())
    frames = torch.zeros((3, H, W), dtype=torch.int32, device="cuda")
    fp.graph_build(frames, H * W, interpolate=True)
    for rep in range(3):
        frames.zero_()
        fp.graph_launch()
        torch.cuda.synchronize()
        got = frames.cpu().numpy()
        for c in range(3):
            assert np.array_equal(got[c], eager[c]), (rep, c)
        # eager work between replays
        fp.render(1, one, None, interpolate=True)
        torch.cuda.synchronize()
        assert np.array_equal(one.cpu().numpy(), eager[1])
    assert fp.status() == 0


@pytest.mark.parametrize("exact", [True, False])
def test_frame_grouped_bins_equal_queue(orc, exact):
    """Resident frames from the grouped workgroups (2 x 2 neighbouring bins
    each) equal the queue's, for every schedule and two footprints, in both
    sample sequences (exact: also the oracle's float composite)."""
    import torch
    blocks = _mixed_blocks(31)
    W, H = 200, 136
    fp, tf, fmin, fmax, dev = _setup(blocks, W, H, acc_cut=0.9999, seed=31, layout=api.BRICK_TILED4)
    fp.set_exact(exact)
    M = datasets.orbit_camera(65.0, elevation_deg=20.0)
    fp.set_cameras(M[None])
    argb = torch.empty((H, W), dtype=torch.int32, device="cuda")
    rgba = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    fp.render(0, argb, rgba, interpolate=True)  # the bricks
    torch.cuda.synchronize()
    ref = rgba.clone()
    if exact:
        bricks = [orc.interpolate(b.sparse, b.n2i, b.depth) for b in blocks]
        assert np.array_equal(ref.cpu().numpy(), _oracle_frame(orc, blocks, bricks, M, W, H, fmin, fmax, tf, 0.9999))
    for tile in ((8, 8), (16, 4)):
        for sch in SCHEDULES:
            _set_schedule(fp, sch)
            for grouped in (False, True):
                fp.set_grouped(grouped)
                fp.render(0, argb, rgba, interpolate=False)
                torch.cuda.synchronize()
                assert fp.status() == 0
                say("grouped", grouped, "schedule", sch, "exact", exact)
                assert torch.equal(rgba, ref)
    fp.close()


def test_frame_autotune_picks_valid_tile(tmp_path):
    import torch
    blocks = _uniform_blocks(1, 4, seed=2)
    W, H = 128, 96
    fp, *_ = _setup(blocks, W, H, acc_cut=0.9999, seed=2)
    fp.set_cameras(datasets.orbit_camera(30.0)[None])
    argb = torch.empty((H, W), dtype=torch.int32, device="cuda")
    cache = str(tmp_path / "tune.jsonl")
    w, h = fp.autotune(argb, True, cache)
    assert w * h == 64 and (w, h) == fp.tile()
    sched = fp.schedule()
    assert sched in SCHEDULES
    w2, h2 = fp.autotune(argb, True, cache)  # served from the cache
    assert (w2, h2) == (w, h) and fp.schedule() == sched
    text = open(cache).read()
    assert text.count("\n") == 1 and '"schedule": %d' % sched in text
    assert '"split": %d' % int(fp.split()) in text


@pytest.mark.parametrize("schedule", SCHEDULES)
@pytest.mark.parametrize("tile", [(8, 8), (16, 4), (4, 16), (32, 2)])
def test_frame_small_blocks_every_footprint(orc, tile, schedule):
    """A sparse level-4 share of depth-2 bricks: blocks a few pixels wide, many
    per bin; every wave footprint and both schedules give the oracle's frame."""
    import torch
    rng = np.random.default_rng(11)
    field = datasets.Field(11, n_gauss=3, noise_octaves=1)
    n2i = api.subtree_n2i(datasets.full_subtree(2), 2)
    keys = [(x, y, z) for z in range(16) for y in range(16) for x in range(16) if rng.random() < 0.3]
    blocks = [datasets.Block(4, k, 2, n2i, datasets.sample(field, 4, k, 2, n2i)) for k in keys]
    W, H = 150, 110
    fp, tf, fmin, fmax, dev = _setup(blocks, W, H, acc_cut=0.999, seed=11, tile=tile)
    assert fp.tile() == tile
    _set_schedule(fp, schedule)
    cams = np.stack([datasets.orbit_camera(a, elevation_deg=e) for a, e in ((20.0, 10.0), (200.0, -30.0))])
    fp.set_cameras(cams)
    argb = torch.empty((H, W), dtype=torch.int32, device="cuda")
    rgba = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    bricks = [orc.interpolate(b.sparse, b.n2i, b.depth) for b in blocks]
    for c in range(2):
        say("small blocks", len(blocks), "tile", tile, "schedule", schedule, "cam", c)
        fp.render(c, argb, rgba, interpolate=True)
        torch.cuda.synchronize()
        assert fp.status() == 0
        exp = _oracle_frame(orc, blocks, bricks, cams[c], W, H, fmin, fmax, tf, 0.999)
        assert np.array_equal(rgba.cpu().numpy(), exp)
        assert (exp[..., 3] > 0).mean() > 0.2
        fp.render(c, argb, rgba, interpolate=False)
        torch.cuda.synchronize()
        assert np.array_equal(rgba.cpu().numpy(), exp)
