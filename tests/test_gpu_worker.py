"""The VsWorkerInterface drop-in (rv_worker) against the oracle: a worker
share (renderScenePart: back-to-front block loop, per-block interpolation and
ray-cast, crop) and a multi-box frame with on-device assembly and final
composition (AssembleFunc / FinalComposeFunc restated in the oracle's
blendImagePart), all bit-exact; cancellation and absent blocks."""
import numpy as np
import pytest

from revise_amd import RvError, api, datasets
from tests.gpu_util import say

pytestmark = pytest.mark.gpu

FOV, Q, THR, CUT = 30.0, 2.0, 0.5, 0.996
BG = (0.1, 0.2, 0.3)


def _make(level, depth, seed, W, H, angle):
    rng = np.random.default_rng(seed)
    field = datasets.Field(seed, n_gauss=3, noise_octaves=1)
    m = 2 ** level
    blocks = {}
    for z in range(m):
        for y in range(m):
            for x in range(m):
                p = float(rng.uniform(0.3, 1.0))
                ch = datasets.refined_subtree(depth, lambda l, locs: rng.random(locs.shape[0]) < p)
                d = api.subtree_depth(ch, depth)
                n2i = api.subtree_n2i(ch, d)
                blocks[(x, y, z)] = datasets.Block(level, (x, y, z), d, n2i,
                                                   datasets.sample(field, level, (x, y, z), d, n2i))
    fmin, fmax = datasets.field_range(blocks.values())
    tf = datasets.seeded_tf(seed, 256)
    M = datasets.orbit_camera(angle)
    w = api.Worker(0)
    for b in blocks.values():
        w.add_subtree(level, b.index, b.depth, b.n2i, b.sparse)
    w.set_state(M, W, H, FOV, Q, THR, CUT, fmin, fmax, BG, tf, exact=True)
    return w, blocks, fmin, fmax, tf, M


def _oracle_share(orc, blocks, level, bmin, bmax, M, W, H, fmin, fmax, tf):
    """Oracle worker share with blocks of possibly different depths: the
    oracle's frame routine takes one depth, so render block by block."""
    vp = np.zeros((H, W), np.uint32)
    eye = api.eye_from_view(M)
    cmin, cmax = api.index_box_coords(level, bmin, bmax)
    order = orc.b2f_macro(bmin, bmax, cmin, cmax, eye)
    for idx in order:
        b = blocks[tuple(int(v) for v in idx)]
        live, po = orc.block_params(M, W, H, FOV, level, b.index, b.depth, fmin, fmax, THR, Q, CUT)
        if live:
            orc.render_block(po, orc.interpolate(b.sparse, b.n2i, b.depth), tf, vp)
    r = orc.bound_rect(vp)
    if r is None:
        return None, None
    x, y, w, h = r
    return r, vp[y:y + h, x:x + w]


@pytest.mark.parametrize("fused", [True, False])
def test_worker_scene_part_bit_exact(orc, fused):
    W, H = 220, 160
    w, blocks, fmin, fmax, tf, M = _make(2, 4, 1, W, H, 35.0)
    w.set_fused(fused)
    for bmin, bmax in [((0, 0, 0), (4, 4, 4)), ((1, 0, 2), (3, 4, 4)), ((3, 3, 3), (4, 4, 4))]:
        say("worker part", bmin, bmax)
        got = w.render_scene_part(2, bmin, bmax)
        rect, bits = got
        erect, ebits = _oracle_share(orc, blocks, 2, bmin, bmax, M, W, H, fmin, fmax, tf)
        assert rect == erect
        assert np.array_equal(bits, ebits)
    w.close()


@pytest.mark.parametrize("fused", [True, False])
def test_worker_empty_box_and_cancel(orc, fused):
    W, H = 64, 48
    w, *_ = _make(1, 3, 2, W, H, 10.0)
    w.set_fused(fused)
    rect, bits = w.render_scene_part(1, (1, 1, 1), (1, 2, 2))  # empty index box
    assert rect == (0, 0, 0, 0) and bits.size == 0
    assert w.render_scene_part(1, (0, 0, 0), (2, 2, 2), cancel=True) is None
    with pytest.raises(RvError, match="Block does not exist"):
        w.render_scene_part(2, (0, 0, 0), (1, 1, 1))
    w.close()


def _oracle_frame(orc, blocks, level, K, M, W, H, fmin, fmax, tf):
    eye = api.eye_from_view(M)
    boxes = orc.index_cube_split_sorted(level, K, eye)
    parts = []
    for b in boxes:
        r, bits = _oracle_share(orc, blocks, level, b[:3], b[3:], M, W, H, fmin, fmax, tf)
        parts.append((r, bits) if r is not None else ((0, 0, 0, 0), None))
    def bbox(a, b):
        xs = [a[0][0], a[0][0] + a[0][2], b[0][0], b[0][0] + b[0][2]]
        ys = [a[0][1], a[0][1] + a[0][3], b[0][1], b[0][1] + b[0][3]]
        return min(xs), min(ys), max(xs) - min(xs), max(ys) - min(ys)
    layer = parts
    while len(layer) > 1:
        nxt = []
        for i in range(1, len(layer), 2):
            a, b = layer[i - 1], layer[i]
            x, y, w_, h_ = bbox(a, b)
            img = np.zeros((h_, w_), np.uint32)
            if a[1] is not None:
                orc.blend_image_part(img, a[1], a[0][:2], (x, y), True)
            if b[1] is not None:
                orc.blend_image_part(img, b[1], b[0][:2], (x, y), False)
            nxt.append(((x, y, w_, h_), img))
        if len(layer) % 2:
            nxt.append(layer[-1])
        layer = nxt
    frame = np.full((H, W), orc.background_argb(*BG), np.uint32)
    if layer[0][1] is not None:
        orc.blend_image_part(frame, layer[0][1], layer[0][0][:2], (0, 0), False)
    return frame


@pytest.mark.parametrize("K,fused", [(1, True), (2, True), (3, False), (5, True), (8, False), (8, True)])
def test_worker_compose_frame_bit_exact(orc, K, fused):
    W, H = 200, 150
    w, blocks, fmin, fmax, tf, M = _make(2, 3, 7, W, H, 120.0)
    w.set_fused(fused)
    say("compose frame K", K)
    got = w.compose_frame(2, K)
    exp = _oracle_frame(orc, blocks, 2, K, M, W, H, fmin, fmax, tf)
    assert np.array_equal(got, exp)
    w.close()


def _oracle_share_mode(orc, blocks, level, bmin, bmax, M, W, H, fmin, fmax, tf, mode, lvl, iso, levels=None):
    init = 0x7FC00000 if mode == api.FIELD_MIP else 0
    vp = np.full((H, W), init, np.uint32)
    eye = api.eye_from_view(M)
    cmin, cmax = api.index_box_coords(level, bmin, bmax)
    thr = lvl if mode == api.FIELD_ISOSURFACE else THR
    for idx in orc.b2f_macro(bmin, bmax, cmin, cmax, eye):
        b = blocks[tuple(int(v) for v in idx)]
        live, po = orc.block_params(M, W, H, FOV, level, b.index, b.depth, fmin, fmax, thr, Q, CUT)
        if live:
            orc.render_block_mode(po, mode, iso, orc.interpolate(b.sparse, b.n2i, b.depth), tf, vp.reshape(-1),
                                  levels=levels)
    r = orc.bound_rect(vp)
    if r is None:
        return None, None
    x, y, w, h = r
    return r, vp[y:y + h, x:x + w]


@pytest.mark.parametrize("mode,fused", [(api.FIELD_ARGB_LIGHT, True), (api.FIELD_ISOSURFACE, False),
                                        (api.FIELD_ISOSURFACE, True), (api.FIELD_MIP, True),
                                        (api.FIELD_MIP, False), (api.FIELD_ISOSURFACES, True),
                                        (api.FIELD_ISOSURFACES, False), (api.FIELD_DOMAIN_VOXELS, True)])
def test_worker_field_modes_bit_exact(orc, mode, fused):
    """renderScenePart and the composed frame in the other field modes: the
    isosurface level replaces the threshold; max intensity parts are float
    images (NaN: none) merged by the larger value and coloured with the
    colour table."""
    W, H = 180, 130
    w, blocks, fmin, fmax, tf, M = _make(2, 3, 11, W, H, 70.0)
    lvl, iso = 0.42, 0.8
    levels = [0.6, 0.35, 0.45] if mode == api.FIELD_ISOSURFACES else None
    w.set_state(M, W, H, FOV, Q, THR, CUT, fmin, fmax, BG, tf, field_mode=mode, iso_level_rel=lvl,
                iso_opacity=iso, iso_levels=levels, exact=True)
    w.set_fused(fused)
    say("worker mode", mode, "fused", fused)
    rect, bits = w.render_scene_part(2, (0, 0, 0), (4, 4, 4))
    erect, ebits = _oracle_share_mode(orc, blocks, 2, (0, 0, 0), (4, 4, 4), M, W, H, fmin, fmax, tf, mode,
                                      lvl, iso, levels)
    assert rect == erect
    assert np.array_equal(bits, ebits)
    if mode != api.FIELD_MIP:
        w.close()
        return
    # composed frame of 3 boxes: max-intensity assembly, colouring over the background
    got = w.compose_frame(2, 3)
    eye = api.eye_from_view(M)
    parts = []
    for b in orc.index_cube_split_sorted(2, 3, eye):
        r, pb = _oracle_share_mode(orc, blocks, 2, b[:3], b[3:], M, W, H, fmin, fmax, tf, mode, lvl, iso)
        parts.append((r, pb) if r is not None else ((0, 0, 0, 0), None))
    layer = parts
    while len(layer) > 1:
        nxt = []
        for i in range(1, len(layer), 2):
            a, b = layer[i - 1], layer[i]
            xs = [a[0][0], a[0][0] + a[0][2], b[0][0], b[0][0] + b[0][2]]
            ys = [a[0][1], a[0][1] + a[0][3], b[0][1], b[0][1] + b[0][3]]
            x, y = min(xs), min(ys)
            img = np.full((max(ys) - y, max(xs) - x), np.nan, np.float32)
            for q, copy in ((a, True), (b, False)):
                if q[1] is not None:
                    orc.blend_image_part_max_intensity(img, q[1].view(np.float32), q[0][0] - x, q[0][1] - y, copy)
            nxt.append(((x, y, img.shape[1], img.shape[0]), img.view(np.uint32)))
        if len(layer) % 2:
            nxt.append(layer[-1])
        layer = nxt
    frame = np.full((H, W), orc.background_argb(*BG), np.uint32)
    if layer[0][1] is not None:
        orc.mip_colorize(frame, layer[0][1].view(np.float32), layer[0][0][0], layer[0][0][1], tf)
    assert np.array_equal(got, frame)
    assert (got != orc.background_argb(*BG)).any()
    w.close()


@pytest.mark.parametrize("mode,fused", [(api.FIELD_VALUE_ON_ISOSURFACE, True),
                                        (api.FIELD_VALUE_ON_ISOSURFACE, False),
                                        (api.FIELD_VALUE_ON_ISOSURFACES, True),
                                        (api.FIELD_VALUE_ON_ISOSURFACES, False)])
def test_worker_value_modes_bit_exact(orc, mode, fused):
    """renderScenePart in ValueOnIsosurface(s): each block's secondary brick
    interpolated beside the primary one (per block, or both in the batched
    interpolation), the isosurface(s) of the primary field coloured by the
    secondary's value."""
    W, H = 170, 120
    w, blocks, fmin, fmax, tf, M = _make(2, 3, 12, W, H, 110.0)
    field2 = datasets.Field(99, n_gauss=4, noise_octaves=2)
    sec = {}
    for key, b in blocks.items():
        sec[key] = datasets.sample(field2, 2, b.index, b.depth, b.n2i)
        w.add_subtree_secondary(2, b.index, sec[key])
    lo = min(float(v.min()) for v in sec.values())
    hi = max(float(v.max()) for v in sec.values())
    lvl, iso = 0.47, 0.75
    levels = [0.3, 0.5, 0.68] if mode == api.FIELD_VALUE_ON_ISOSURFACES else None
    w.set_state(M, W, H, FOV, Q, THR, CUT, fmin, fmax, BG, tf, field_mode=mode, iso_level_rel=lvl,
                iso_opacity=iso, iso_levels=levels, secondary_range=(lo, hi), exact=True)
    w.set_fused(fused)
    say("worker value mode", mode, "fused", fused)
    for bmin, bmax in [((0, 0, 0), (4, 4, 4)), ((1, 0, 1), (3, 4, 4))]:
        rect, bits = w.render_scene_part(2, bmin, bmax)
        vp = np.zeros((H, W), np.uint32)
        eye = api.eye_from_view(M)
        cmin, cmax = api.index_box_coords(2, bmin, bmax)
        thr = lvl if mode == api.FIELD_VALUE_ON_ISOSURFACE else THR
        for idx in orc.b2f_macro(bmin, bmax, cmin, cmax, eye):
            key = tuple(int(v) for v in idx)
            b = blocks[key]
            live, po = orc.block_params(M, W, H, FOV, 2, b.index, b.depth, fmin, fmax, thr, Q, CUT)
            if live:
                orc.render_block_mode(po, mode, iso, orc.interpolate(b.sparse, b.n2i, b.depth), tf,
                                      vp.reshape(-1), levels=levels,
                                      brick2=orc.interpolate(sec[key], b.n2i, b.depth), srange=(lo, hi))
        r = orc.bound_rect(vp)
        assert r is not None and rect == r
        x, y, ww, hh = r
        assert np.array_equal(bits, vp[y:y + hh, x:x + ww])
    w.close()


def _level_blocks(field, level, depth, rng=None, p=1.0):
    m = 2 ** level
    out = {}
    for z in range(m):
        for y in range(m):
            for x in range(m):
                ch = (datasets.full_subtree(depth) if rng is None else
                      datasets.refined_subtree(depth, lambda l, locs: rng.random(locs.shape[0]) < p))
                d = api.subtree_depth(ch, depth)
                n2i = api.subtree_n2i(ch, d)
                out[(x, y, z)] = datasets.Block(level, (x, y, z), d, n2i,
                                                datasets.sample(field, level, (x, y, z), d, n2i))
    return out


def test_worker_fused_deeper_subtree_at_next_level(orc):
    """The fused path keeps one brick per (level, index) sized for its
    subtree: block (3,3,3) is a depth-1 subtree at level 2 and a depth-2
    subtree at level 3 (64 then 512 micro-brick floats).  Level 2, then level
    3, then level 2 again render bit-exact against the oracle and the
    per-block path."""
    W, H = 160, 120
    field = datasets.Field(21, n_gauss=3)
    levels = {2: _level_blocks(field, 2, 1), 3: _level_blocks(field, 3, 2)}
    fmin = min(datasets.field_range(b.values())[0] for b in levels.values())
    fmax = max(datasets.field_range(b.values())[1] for b in levels.values())
    tf = datasets.seeded_tf(21, 256)
    M = datasets.orbit_camera(40.0)
    ws = {}
    for fused in (True, False):
        w = api.Worker(0)
        for lv, bl in levels.items():
            for b in bl.values():
                w.add_subtree(lv, b.index, b.depth, b.n2i, b.sparse)
        w.set_state(M, W, H, FOV, Q, THR, CUT, fmin, fmax, BG, tf, exact=True)
        w.set_fused(fused)
        ws[fused] = w
    for lv in (2, 3, 2):
        m = 2 ** lv
        erect, ebits = _oracle_share(orc, levels[lv], lv, (0, 0, 0), (m,) * 3, M, W, H, fmin, fmax, tf)
        for fused, w in ws.items():
            say("level", lv, "fused", fused)
            rect, bits = w.render_scene_part(lv, (0, 0, 0), (m,) * 3)
            assert rect == erect
            assert np.array_equal(bits, ebits)
    for w in ws.values():
        w.close()


def test_worker_replaced_values_rerender(orc):
    """Replacing a block's values (rv_worker_add_subtree again, same depth)
    re-interpolates it in both paths: per-block equals fused equals the
    oracle after the change, and the part differs from before."""
    W, H = 128, 96
    field = datasets.Field(22, n_gauss=3)
    blocks = _level_blocks(field, 1, 3)
    fmin, fmax = datasets.field_range(blocks.values())
    tf = datasets.seeded_tf(22, 256)
    M = datasets.orbit_camera(15.0)
    ws = {}
    for fused in (True, False):
        w = api.Worker(0)
        for b in blocks.values():
            w.add_subtree(1, b.index, b.depth, b.n2i, b.sparse)
        w.set_state(M, W, H, FOV, Q, THR, CUT, fmin, fmax, BG, tf, exact=True)
        w.set_fused(fused)
        ws[fused] = w
    # a one-block box, so the per-block renderer's cached brick is this block
    box = ((1, 1, 1), (2, 2, 2))
    before = {f: w.render_scene_part(1, *box) for f, w in ws.items()}
    assert before[True][0] == before[False][0] and np.array_equal(before[True][1], before[False][1])
    b = blocks[(1, 1, 1)]
    new = datasets.Block(1, b.index, b.depth, b.n2i, (fmax - (b.sparse - fmin)).astype(np.float32))
    blocks[(1, 1, 1)] = new
    for w in ws.values():
        w.add_subtree(1, new.index, new.depth, new.n2i, new.sparse)
    erect, ebits = _oracle_share(orc, blocks, 1, box[0], box[1], M, W, H, fmin, fmax, tf)
    for fused, w in ws.items():
        rect, bits = w.render_scene_part(1, *box)
        assert rect == erect and np.array_equal(bits, ebits), fused
    assert not (before[True][0] == erect and np.array_equal(before[True][1], ebits))
    for w in ws.values():
        w.close()


def test_worker_compose_frame_repeated_uses_cached_batches(orc):
    """Two frames of 5 boxes in a row on one worker (one cached batch per
    box), then a frame after a value change: each equals the oracle."""
    W, H = 180, 140
    w, blocks, fmin, fmax, tf, M = _make(2, 3, 9, W, H, 200.0)
    exp = _oracle_frame(orc, blocks, 2, 5, M, W, H, fmin, fmax, tf)
    assert np.array_equal(w.compose_frame(2, 5), exp)
    assert np.array_equal(w.compose_frame(2, 5), exp)
    b = blocks[(2, 1, 1)]
    blocks[(2, 1, 1)] = datasets.Block(2, b.index, b.depth, b.n2i, np.flip(b.sparse).copy())
    w.add_subtree(2, b.index, b.depth, b.n2i, blocks[(2, 1, 1)].sparse)
    assert np.array_equal(w.compose_frame(2, 5), _oracle_frame(orc, blocks, 2, 5, M, W, H, fmin, fmax, tf))
    w.close()


@pytest.mark.parametrize("exact", [True, False])
def test_worker_level5_share_bins(exact):
    """A level-5 share, 32768 blocks (full depth-2 subtrees), at 640 x 360:
    the fused path's share launch visits each 16 x 16 screen bin's blocks
    only (the bins built with the block table); its part equals the
    per-block path's bit for bit, in both sample sequences."""
    import time
    W, H = 640, 360
    field = datasets.Field(23, n_gauss=3)
    n2i = api.subtree_n2i(datasets.full_subtree(2), 2)
    blocks = datasets.uniform_level(5, 2, field, n2i)
    fmin, fmax = datasets.field_range(blocks)
    tf = datasets.seeded_tf(23, 256)
    M = datasets.orbit_camera(55.0)
    parts = {}
    for fused in (True, False):
        w = api.Worker(0)
        for b in blocks:
            w.add_subtree(5, b.index, b.depth, b.n2i, b.sparse)
        w.set_state(M, W, H, FOV, Q, THR, CUT, fmin, fmax, BG, tf, exact=exact)
        w.set_fused(fused)
        w.render_scene_part(5, (0, 0, 0), (32, 32, 32))
        t0 = time.perf_counter()
        parts[fused] = w.render_scene_part(5, (0, 0, 0), (32, 32, 32))
        say("level-5 share, fused %s: %.1f ms" % (fused, (time.perf_counter() - t0) * 1e3))
        w.close()
    assert parts[True][0] == parts[False][0]
    assert parts[True][0][2] * parts[True][0][3] > W * H // 10
    assert np.array_equal(parts[True][1], parts[False][1])


@pytest.mark.parametrize("exact", [True, False])
def test_share_scan_equals_bins_level5(exact):
    """The same level-5 share (32768 full depth-2 blocks, 640 x 360) through
    the share launch that scans every block's view rectangle per wave
    (rv_render_share_layout) and through the worker's binned launch: the
    cropped parts are equal bit for bit."""
    import torch
    from revise_amd import configs
    W, H = 640, 360
    field = datasets.Field(24, n_gauss=3)
    wl = configs.uniform(torch, field, level=5, depth=2)
    tf_np = datasets.seeded_tf(24, 256)
    M = datasets.orbit_camera(125.0, elevation_deg=30.0)
    fmin, fmax = float(wl.fmin), float(wl.fmax)
    host = wl.sparse.cpu().numpy()
    w = api.Worker(0)
    for i, b in enumerate(wl.blocks):
        w.add_subtree(5, b["index"], 2, wl.n2i["full"], host[i])
    w.set_state(M, W, H, FOV, Q, THR, CUT, fmin, fmax, BG, tf_np, exact=exact)
    w.set_fused(True)
    rect, bits = w.render_scene_part(5, (0, 0, 0), (32, 32, 32))
    w.close()
    tiled = [torch.empty(api.brick_floats(2, api.BRICK_TILED4), dtype=torch.float32, device="cuda")
             for _ in wl.blocks]
    api.InterpBatch([(b["sparse"], b["plan"], t, 2) for b, t in zip(wl.blocks, tiled)], api.BRICK_TILED4).run()
    keys = np.array([(5,) + tuple(b["index"]) for b in wl.blocks], dtype=np.uint32)
    items = []
    for i in api.mixed_back_to_front(keys, api.eye_from_view(M)):
        b = wl.blocks[int(i)]
        live, p = api.block_params(M, W, H, FOV, 5, b["index"], 2, fmin, fmax, THR, Q, CUT, exact=exact)
        if live:
            items.append((p, tiled[int(i)]))
    tf = torch.from_numpy(tf_np).cuda()
    vp = torch.empty((H, W), dtype=torch.int32, device="cuda")
    api.clear_viewport(vp, W, H)
    api.ShareRenderer(items, tf, W, H, layout=api.BRICK_TILED4).render(vp)
    torch.cuda.synchronize()
    say("level-5 share: %d live blocks, part %s" % (len(items), rect))
    assert tuple(api.bound_rect(vp, W, H)) == tuple(rect)
    x, y, pw, ph = rect
    img = vp.cpu().numpy().view(np.uint32)
    assert np.array_equal(img[y:y + ph, x:x + pw].reshape(-1), np.asarray(bits).reshape(-1))


def test_worker_part_into_pinned_host_buffer():
    """A pinned host destination receives the part from the crop kernel
    itself: the same rectangle and pixels as a pageable destination; a
    destination too small for the part is an error and is left untouched."""
    import torch
    W, H = 220, 160
    w, blocks, fmin, fmax, tf, M = _make(2, 4, 3, W, H, 65.0)
    w.set_fused(True)
    for exact in (True, False):
        w.set_state(M, W, H, FOV, Q, THR, CUT, fmin, fmax, BG, tf, exact=exact)
        ref = w.render_scene_part(2, (0, 0, 0), (4, 4, 4))
        out = torch.full((W * H,), -7, dtype=torch.int32, pin_memory=True)
        got = w.render_scene_part(2, (0, 0, 0), (4, 4, 4), out=out)
        assert got[0] == ref[0]
        assert np.array_equal(got[1], ref[1])
        n = ref[0][2] * ref[0][3]
        assert n > 0 and (out.numpy()[n:] == -7).all()
        small = torch.full((n - 1,), -7, dtype=torch.int32, pin_memory=True)
        with pytest.raises(RvError):
            w.render_scene_part(2, (0, 0, 0), (4, 4, 4), out=small)
        assert (small.numpy() == -7).all()
    w.close()
