"""Parity at the benchmark sizes.

* GPU plain drop-in: 512 level-3 blocks, each a full depth-6 subtree (65^3
  fp32), 3 seeded Gaussians, 1920x1080, quality 2, cutoff 0.99, through
  rv_worker_render_scene_part (VsWorker_cu.cpp:225-313) -- fused and per
  block -- against the oracle's per-block frame over the whole viewport
  (every host core) and its crop (getBoundRect, VsWorker_cu.cpp:51-88).
  Exact sequence: bit-exact; fast sequence: fused equals per-block bit for
  bit, and the 8-bit part is within the per-block requantisation bar of
  test_gpu_fast (2 LSB premultiplied, under 1% of the channels off).
* Batched persistent: 2000 blocks of levels 2-6, fill 5-100%, 8 Gaussians +
  3-octave noise, 1920x1080: interpolating frames replayed from a hipGraph
  over 8 cameras (bricks cleared before the replay), two of them against the
  oracle's float composite (exact: the 8-bit frames bit-exact with its
  encoding; fast: float RGBA within 2e-3, premultiplied).
"""
import concurrent.futures as cf
import time

import numpy as np
import pytest

from revise_amd import api, configs, datasets
from tests.gpu_util import host_threads, oracle_blocks_frame, say, to_dev
from tests.test_gpu_fast import _lsb_diff, premul

pytestmark = pytest.mark.gpu

W, H = 1920, 1080
FOV, THR, Q, CUT2 = 30.0, 0.5, 2.0, 0.99
SEED = 2024




@pytest.fixture(scope="module")
def config2(orc):
    t0 = time.time()
    field = datasets.Field(SEED, n_gauss=3)
    n2i = api.subtree_n2i(datasets.full_subtree(6), 6)
    blocks = datasets.uniform_level(3, 6, field, n2i)
    fmin, fmax = datasets.field_range(blocks)
    tf = datasets.seeded_tf(SEED, 256)
    angle = float(np.random.default_rng(SEED).uniform(0.0, 360.0))
    M = datasets.orbit_camera(angle)
    with cf.ThreadPoolExecutor(host_threads()) as ex:
        bricks = list(ex.map(lambda b: orc.interpolate(b.sparse, b.n2i, b.depth), blocks))
    by = {b.index: (b, br) for b, br in zip(blocks, bricks)}
    eye = api.eye_from_view(M)
    order = orc.b2f_macro((0, 0, 0), (8, 8, 8), *api.index_box_coords(3, (0, 0, 0), (8, 8, 8)), eye)
    items = []
    for idx in order:
        b, br = by[tuple(int(v) for v in idx)]
        live, po = orc.block_params(M, W, H, FOV, 3, b.index, 6, fmin, fmax, THR, Q, CUT2)
        if live:
            items.append((po, br))
    vp = oracle_blocks_frame(orc, items, tf, W, H)
    rect = orc.bound_rect(vp)
    x, y, w, h = rect
    say("config 2 oracle frame: %d live blocks, crop %s, %.1fs on %d threads"
        % (len(items), rect, time.time() - t0, host_threads()))
    return dict(blocks=blocks, fmin=fmin, fmax=fmax, tf=tf, M=M, rect=rect, crop=vp[y:y + h, x:x + w])


def _worker(c, fused, exact):
    w = api.Worker(0)
    for b in c["blocks"]:
        w.add_subtree(3, b.index, b.depth, b.n2i, b.sparse)
    w.set_state(c["M"], W, H, FOV, Q, THR, CUT2, c["fmin"], c["fmax"], (0.0, 0.0, 0.0), c["tf"], exact=exact)
    w.set_fused(fused)
    return w


@pytest.mark.parametrize("fused", [True, False])
def test_config2_dropin_exact_bit_exact(config2, fused):
    w = _worker(config2, fused, True)
    say("config 2 worker part, fused", fused)
    rect, bits = w.render_scene_part(3, (0, 0, 0), (8, 8, 8))
    w.close()
    assert rect == config2["rect"]
    assert np.array_equal(bits, config2["crop"])


def test_config2_dropin_fast(config2):
    parts = {}
    for fused in (True, False):
        w = _worker(config2, fused, False)
        parts[fused] = w.render_scene_part(3, (0, 0, 0), (8, 8, 8))
        w.close()
    assert parts[True][0] == parts[False][0]
    assert np.array_equal(parts[True][1], parts[False][1])
    rect, bits = parts[True]
    # the crop may differ by where a nearly transparent edge pixel rounds to 0
    # alpha: compare on the union of both rectangles
    ex, ey, ew, eh = config2["rect"]
    x0, y0 = min(rect[0], ex), min(rect[1], ey)
    x1, y1 = max(rect[0] + rect[2], ex + ew), max(rect[1] + rect[3], ey + eh)
    a = np.zeros((y1 - y0, x1 - x0), np.uint32)
    b = np.zeros_like(a)
    a[rect[1] - y0:rect[1] - y0 + rect[3], rect[0] - x0:rect[0] - x0 + rect[2]] = bits.reshape(rect[3], rect[2])
    b[ey - y0:ey - y0 + eh, ex - x0:ex - x0 + ew] = config2["crop"]
    d = _lsb_diff(a, b)
    say("config 2 fast part: max LSB", int(d.max()), "channels off %.4f%%" % (100 * (d > 0).mean()))
    assert d.max() <= 2
    assert (d > 0).mean() < 0.01


def _sample_counts(p):
    """sampleCount of every pixel of a block's view rectangle (0 where the
    ray misses): setup_ray's fp32 steps in numpy (DenseFieldRenderer_cu.cu:
    62-175), the same operations in the same order."""
    f = np.float32
    vx, vy = np.meshgrid(np.arange(p.view_x, p.view_x + p.view_w, dtype=f),
                         np.arange(p.view_y, p.view_y + p.view_h, dtype=f))
    x = (f(-0.5) + (vx + f(0.5)) / f(p.vp_w)) * f(p.w)
    y = (f(-0.5) + (vy + f(0.5)) / f(p.vp_h)) * f(p.h)
    r = [(f(p.n[a]) + f(p.e1[a]) * x) + f(p.e2[a]) * y for a in range(3)]
    ev, h = [f(p.pos[a]) for a in range(3)], f(p.half)
    te, tx = np.full(x.shape, f(1.0)), np.zeros(x.shape, f)
    with np.errstate(all="ignore"):
        for a in range(3):
            o1, o2 = (1 if a == 0 else 0), (1 if a == 2 else 2)
            for side in range(2):
                face = f(p.center[a]) - h if side == 0 else f(p.center[a]) + h
                tau = (face - ev[a]) / r[a]
                on = ((np.abs(r[a]) > 1e-10) & (np.abs(ev[o1] + r[o1] * tau - f(p.center[o1])) <= h)
                      & (np.abs(ev[o2] + r[o2] * tau - f(p.center[o2])) <= h))
                enter = (r[a] > 0) if side == 0 else ~(r[a] > 0)
                te, tx = np.where(on & enter, tau, te), np.where(on & ~enter, tau, tx)
        te = np.maximum(te, f(1.0))
        dotrr = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2]
        l1 = (np.abs(r[0]) + np.abs(r[1])) + np.abs(r[2])
        sc = (f(p.quality) * (tx - te) * dotrr / (f(p.texel) * l1)).astype(np.int64)
    return np.where(tx > te, sc, 0)


@pytest.mark.parametrize("exact", [False, True])
def test_config2_dropin_overlong_segments(config2, exact):
    """Cameras where rays grazing a block's edge miss its entry face in fp32
    and start at the eye (tauEnter = 1, DenseFieldRenderer_cu.cu:164-165):
    segments of thousands of samples, which the share kernel continues with
    the whole wave; its frame equals the per-block launches' (one lane per
    ray) bit for bit, in both sample sequences."""
    angle = float(np.random.default_rng(SEED).uniform(0.0, 360.0))
    longest = 0
    for da in (15.0, 30.0, 45.0):
        c = dict(config2, M=datasets.orbit_camera(angle + da))
        for b in c["blocks"]:
            live, p = api.block_params(c["M"], W, H, FOV, 3, b.index, 6, c["fmin"], c["fmax"], THR, Q, CUT2)
            if live:
                longest = max(longest, int(_sample_counts(p).max()))
        parts = {}
        for fused in (True, False):
            w = _worker(c, fused, exact)
            parts[fused] = w.render_scene_part(3, (0, 0, 0), (8, 8, 8))
            w.close()
        say("config 2 %s, camera %+.0f deg: fused == per block" % ("exact" if exact else "fast", da))
        assert parts[True][0] == parts[False][0]
        assert np.array_equal(parts[True][1], parts[False][1])
    say("longest segment over the three cameras: %d samples" % longest)
    assert longest > 1000  # (inside a block: 2 sqrt(3) 65 at most)


@pytest.fixture(scope="module")
def config3(orc):
    import torch
    t0 = time.time()
    field = datasets.Field(SEED + 3, n_gauss=8, noise_octaves=3)
    wl = configs.mixed(torch, field, nblocks=2000, depth=6, seed=SEED + 3)
    tf_np = datasets.seeded_tf(SEED, 256)
    rng_a, rng_e = np.random.default_rng(SEED + 3), np.random.default_rng(SEED + 4)
    cams = np.stack([datasets.orbit_camera(a, elevation_deg=e)
                     for a, e in zip(rng_a.uniform(0, 360, 8), rng_e.uniform(-45, 45, 8))])
    sparse = [b["sparse"].cpu().numpy() for b in wl.blocks]

    def interp(i):
        b = wl.blocks[i]
        return orc.interpolate(sparse[i], wl.n2i[b["shape"]], b["depth"])

    with cf.ThreadPoolExecutor(host_threads()) as ex:
        bricks = list(ex.map(interp, range(len(wl.blocks))))
    say("config 3: %d blocks, oracle bricks in %.1fs" % (len(wl.blocks), time.time() - t0))
    return dict(wl=wl, tf=tf_np, cams=cams, bricks=bricks)


def _oracle_composite(orc, c, cam, acc_cut):
    wl = c["wl"]
    keys = np.array([(b["level"],) + tuple(b["index"]) for b in wl.blocks], dtype=np.uint32)
    order = api.mixed_back_to_front(keys, api.eye_from_view(c["cams"][cam]))
    lst = [(wl.blocks[i]["level"], tuple(wl.blocks[i]["index"]), wl.blocks[i]["depth"], c["bricks"][i])
           for i in order]
    return orc.render_float_composite(c["cams"][cam], W, H, FOV, lst, wl.fmin, wl.fmax, THR, Q, CUT2,
                                      c["tf"], acc_cut, nthreads=host_threads())


def _tiled(torch, wl):
    return [dict(b, brick=torch.empty(api.brick_floats(b["depth"], api.BRICK_TILED4), dtype=torch.float32,
                                      device="cuda")) for b in wl.blocks]


def test_config3_graph_frames_exact(orc, config3):
    import torch
    wl = config3["wl"]
    tblocks = _tiled(torch, wl)
    tf = to_dev(config3["tf"])
    fp = api.FramePlan(tblocks, W, H, FOV, wl.fmin, wl.fmax, THR, Q, CUT2, 0.9999, tf, (8, 8),
                       api.BRICK_TILED4, exact=True)
    fp.set_cameras(config3["cams"])
    frames = torch.zeros((8, H, W), dtype=torch.int32, device="cuda")
    fp.graph_build(frames, H * W, True)
    for b in tblocks:  # the graph's frames must interpolate the bricks themselves
        b["brick"].fill_(float("nan"))
    torch.cuda.synchronize()
    fp.graph_launch()
    torch.cuda.synchronize()
    assert fp.status() == 0
    for cam in (0, 5):
        exp = _oracle_composite(orc, config3, cam, 0.9999)
        enc = (np.minimum(exp, 1.0) * np.float32(255.999)).astype(np.uint32)
        enc = (enc[..., 3] << 24) | (enc[..., 0] << 16) | (enc[..., 1] << 8) | enc[..., 2]
        got = frames[cam].cpu().numpy().view(np.uint32)
        say("config 3 camera", cam, "covered pixels", int((enc >> 24).astype(bool).sum()))
        assert (enc >> 24).any()
        assert np.array_equal(got, enc)
    fp.close()


def test_config3_frame_fast_float(orc, config3):
    import torch
    wl = config3["wl"]
    tblocks = _tiled(torch, wl)
    tf = to_dev(config3["tf"])
    fp = api.FramePlan(tblocks, W, H, FOV, wl.fmin, wl.fmax, THR, Q, CUT2, 0.9999, tf, (8, 8), api.BRICK_TILED4)
    fp.set_cameras(config3["cams"][[2]])
    argb = torch.empty((H, W), dtype=torch.int32, device="cuda")
    rgba = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    fp.render(0, argb, rgba, interpolate=True)
    torch.cuda.synchronize()
    assert fp.status() == 0
    exp = _oracle_composite(orc, config3, 2, 0.9999)
    err = np.abs(premul(rgba.cpu().numpy()) - premul(exp))
    say("config 3 fast float: max %.2e mean %.2e" % (err.max(), err.mean()))
    assert err.max() <= 2e-3
    # the walks give the same float frame bit for bit: the bin walk (one
    # block per wave), the per-lane walk (each lane on its own block, levels
    # and depths mixed) and the hybrid, each continuing overlong segments
    # with the whole wave
    first = rgba.clone()
    for sch in (api.SCHEDULE_BINS, api.SCHEDULE_LANES, api.SCHEDULE_HYBRID):
        fp.set_schedule(sch)
        if sch == api.SCHEDULE_HYBRID:
            fp.set_hybrid_min(5)
        fp.render(0, argb, rgba, interpolate=False)
        torch.cuda.synchronize()
        assert fp.status() == 0
        say("config 3 fast float, schedule", sch)
        assert torch.equal(rgba, first)
    fp.close()
