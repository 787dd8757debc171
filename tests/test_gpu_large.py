"""Large single-GPU share: the blocks of one level with the largest centre
values (12%), each a full depth-4 subtree (17^3 sparse nodes, fp16 brick),
field = seeded Gaussians + value noise (configs.large_share).

* reduced size (level 5: 3932 blocks, 384x216, two cameras): the batched
  frame, resident and interpolating, equals the oracle's float composite of
  the fp16 bricks bit for bit;
* colour-table adjoint (level 4, 256x144, 48x48 crop): equals torch autograd
  of the float64 reference of the same forward within 1e-4 relative;
* field-gradient adjoint (level 6, 1920x1080, 16x16 crop around a segment
  of thousands of samples): the same bar for both gradients;
* full size (level 8: about 2.0 M blocks, 3840x2160), through properties
  that do not depend on size: every brick equals its nodes' sparse values
  rounded to fp16 (a full subtree makes the interpolation a scatter), the
  interpolating frame equals the resident one, two wave footprints give the
  same frame, and the colour-table gradient of several views is the sum of
  the per-view gradients (to float-atomic rounding)."""
import numpy as np
import pytest

from revise_amd import api, configs, datasets
from tests.adjoint_ref import SampleTable
from tests.gpu_util import host_threads, say

pytestmark = pytest.mark.gpu

FOV, THR, Q, CUT, ACC = 30.0, 0.5, 2.0, 0.99, 0.9999


def _share(level, seed):
    import torch
    ls = configs.large_share(torch, seed, level=level, depth=4, occupancy=0.12, n_gauss=32, octaves=4)
    api.InterpBatch(ls.brick_jobs()).run()
    torch.cuda.synchronize()
    return ls


def _scatter_check(torch, ls):
    nn = 17
    fi = torch.from_numpy(((ls.n2i[:, 2].astype(np.int64) * nn + ls.n2i[:, 1]) * nn + ls.n2i[:, 0])).cuda()
    N = ls.n2i.shape[0]
    nb = ls.keys.shape[0]
    for c0 in range(0, nb, 1 << 18):
        c1 = min(nb, c0 + (1 << 18))
        assert torch.equal(ls.bricks[c0:c1, :nn ** 3][:, fi], ls.sparse[c0:c1, :N].half())


def test_large_share_reduced_matches_oracle(orc):
    import torch
    ls = _share(5, seed=51)
    nb = ls.keys.shape[0]
    say("level-5 share", nb, "blocks")
    assert nb == round(0.12 * 32 ** 3)
    _scatter_check(torch, ls)
    W, H = 384, 216
    tf_np = datasets.seeded_tf(51, 256)
    tf = torch.from_numpy(tf_np).cuda()
    fp = api.FramePlan(ls.frame_blocks(), W, H, FOV, ls.fmin, ls.fmax, THR, Q, CUT, ACC, tf, exact=True)
    cams = np.stack([datasets.orbit_camera(a, elevation_deg=e) for a, e in ((35.0, 15.0), (220.0, -25.0))])
    fp.set_cameras(cams)
    argb = torch.empty((H, W), dtype=torch.int32, device="cuda")
    rgba = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    bricks = ls.bricks[:, :17 ** 3].float().cpu().numpy()
    keys = np.concatenate([np.full((nb, 1), 5, np.uint32), ls.keys.astype(np.uint32)], 1)
    for c in range(2):
        order = api.mixed_back_to_front(keys, api.eye_from_view(cams[c]))
        lst = [(5, tuple(int(v) for v in ls.keys[i]), 4, bricks[i]) for i in order]
        exp = orc.render_float_composite(cams[c], W, H, FOV, lst, ls.fmin, ls.fmax, THR, Q, CUT, tf_np, ACC)
        assert (exp[..., 3] > 0).mean() > 0.03
        for interp in (False, True):
            say("level-5 frame cam", c, "interpolate", interp)
            fp.render(c, argb, rgba, interpolate=interp)
            torch.cuda.synchronize()
            assert fp.status() == 0
            assert np.array_equal(rgba.cpu().numpy(), exp)
    fp.close()


def test_large_share_exact_overlong_segments_match_oracle(orc):
    """The exact sequence over the same small blocks at the camera whose
    frame holds a segment of 4,000 samples: both walks, which continue such
    segments with the whole wave (march_impl), equal the oracle's float
    composite (one ray at a time) bit for bit."""
    import torch
    ls = _share(6, seed=61)
    W, H = 1920, 1080
    tf_np = datasets.seeded_tf(61, 256)
    tf = torch.from_numpy(tf_np).cuda()
    fp = api.FramePlan(ls.frame_blocks(), W, H, FOV, ls.fmin, ls.fmax, THR, Q, CUT, ACC, tf, (8, 8), exact=True)
    cam = datasets.orbit_camera(35.0, elevation_deg=15.0)
    fp.set_cameras(cam[None])
    argb = torch.empty((H, W), dtype=torch.int32, device="cuda")
    rgba = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    nb = ls.keys.shape[0]
    bricks = ls.bricks[:, :17 ** 3].float().cpu().numpy()
    keys = np.concatenate([np.full((nb, 1), 6, np.uint32), ls.keys.astype(np.uint32)], 1)
    order = api.mixed_back_to_front(keys, api.eye_from_view(cam))
    lst = [(6, tuple(int(v) for v in ls.keys[i]), 4, bricks[i]) for i in order]
    exp = orc.render_float_composite(cam, W, H, FOV, lst, ls.fmin, ls.fmax, THR, Q, CUT, tf_np, ACC,
                                     nthreads=host_threads())
    for sch in (api.SCHEDULE_BINS, api.SCHEDULE_LANES):
        fp.set_schedule(sch)
        fp.render(0, argb, rgba, interpolate=False)
        torch.cuda.synchronize()
        assert fp.status() == 0
        say("level-6 exact frame, schedule", sch)
        assert np.array_equal(rgba.cpu().numpy(), exp)
    fp.close()


def test_large_share_fast_lanes_equal_bins():
    """The fast sequence over small fp16 blocks (level 6: 31,457 blocks,
    1920x1080), where rays grazing a block's edge start at the eye and march
    hundreds to thousands of samples, which the waves continue together
    (march_fast): the per-lane walk, whose lanes are on different blocks,
    gives the float frame of the bin walk (one block per wave) bit for bit.
    (The one-block form equals the one-lane march of per-block launches:
    test_gpu_fullsize.py.)"""
    import torch
    from tests.test_gpu_fullsize import _sample_counts
    ls = _share(6, seed=61)
    W, H = 1920, 1080
    tf = torch.from_numpy(datasets.seeded_tf(61, 256)).cuda()
    fp = api.FramePlan(ls.frame_blocks(), W, H, FOV, ls.fmin, ls.fmax, THR, Q, CUT, ACC, tf, (8, 8))
    cams = np.stack([datasets.orbit_camera(a, elevation_deg=e) for a, e in ((35.0, 15.0), (220.0, -25.0),
                                                                          (300.0, 5.0))])
    fp.set_cameras(cams)
    argb = torch.empty((H, W), dtype=torch.int32, device="cuda")
    rgba = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    longest = 0
    for c in range(3):
        for k in ls.keys[::7]:  # (a sample of the blocks, for the assertion below)
            live, p = api.block_params(cams[c], W, H, FOV, 6, tuple(int(v) for v in k), 4, ls.fmin, ls.fmax,
                                       THR, Q, CUT)
            if live:
                longest = max(longest, int(_sample_counts(p).max()))
        out = {}
        for sch in (api.SCHEDULE_BINS, api.SCHEDULE_LANES):
            fp.set_schedule(sch)
            fp.render(c, argb, rgba, interpolate=False)
            torch.cuda.synchronize()
            assert fp.status() == 0
            out[sch] = rgba.clone()
        say("level-6 fast frame cam", c, "lanes == bins")
        assert (out[api.SCHEDULE_BINS][..., 3] > 0).float().mean() > 0.03
        assert torch.equal(out[api.SCHEDULE_BINS], out[api.SCHEDULE_LANES])
    fp.close()
    say("longest segment among the sampled blocks: %d samples" % longest)
    assert longest > 500  # (inside a 17^3 block: 2 sqrt(3) 17 at most)


def test_large_share_tf_adjoint_matches_float64(orc):
    import torch
    ls = _share(4, seed=44)
    W, H = 256, 144
    crop = (104, 48, 48, 48)
    tf_np = datasets.seeded_tf(44, 256)
    rng = np.random.default_rng(45)
    tf_t = np.clip(tf_np + rng.normal(0.0, 0.05, size=tf_np.shape), 0.0, 1.0).astype(np.float32)
    M = datasets.orbit_camera(120.0, elevation_deg=20.0)
    tf = torch.from_numpy(tf_np).cuda()
    fp = api.FramePlan(ls.frame_blocks(), W, H, FOV, ls.fmin, ls.fmax, THR, Q, CUT, ACC, tf, exact=True)
    fp.set_cameras(M[None])
    argb = torch.empty((H, W), dtype=torch.int32, device="cuda")
    target = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    tf.copy_(torch.from_numpy(tf_t))
    fp.render(0, argb, target, interpolate=False)
    tf.copy_(torch.from_numpy(tf_np))
    fp.set_tape(True)
    rgba = torch.empty_like(target)
    fp.render(0, argb, rgba, interpolate=False)
    x0, y0, cw, ch = crop
    mask = torch.zeros((H, W, 1), dtype=torch.float32, device="cuda")
    for j in range(ch):
        mask[H - 1 - (y0 + j), x0:x0 + cw] = 1.0
    gI = (rgba - target) * mask * mask
    gtf = torch.zeros((256, 4), dtype=torch.float32, device="cuda")
    fp.adjoint(0, rgba, gI, gtf)
    torch.cuda.synchronize()
    g_tf = gtf.double().cpu().numpy()
    # float64 reference over the crop's samples
    blocks = [{"level": 4, "index": tuple(int(v) for v in k), "depth": 4} for k in ls.keys]
    bricks_np = ls.bricks[:, :17 ** 3].float().cpu().numpy()
    table = SampleTable(orc, blocks, bricks_np, M, W, H, FOV, ls.fmin, ls.fmax, THR, Q, CUT, ACC, tf_np, crop)
    say("samples", table.pix.size, "blocks touched", table.touched.size)
    assert table.pix.size > 1000
    tf64 = torch.tensor(tf_np, dtype=torch.float64, device="cuda", requires_grad=True)
    b64 = torch.tensor(bricks_np, dtype=torch.float64, device="cuda")
    img = table.forward(torch, tf64, b64, "cuda")
    rows = [H - 1 - (y0 + j) for j in range(ch)]
    tgt = target[rows][:, x0:x0 + cw].double()
    got = rgba[rows][:, x0:x0 + cw].double()
    assert float((got - img.detach()).abs().max()) < 2e-3
    (0.5 * ((img - tgt) ** 2).sum()).backward()
    r_tf = tf64.grad.cpu().numpy()
    say("tf grad rel err", np.linalg.norm(g_tf - r_tf) / np.linalg.norm(r_tf))
    assert np.linalg.norm(r_tf) > 0
    assert np.linalg.norm(g_tf - r_tf) <= 1e-4 * np.linalg.norm(r_tf)
    fp.close()


def test_large_share_field_adjoint_overlong_segments_matches_float64(orc):
    """The field-gradient adjoint (adjoint_kernel) over a crop around the
    longest segment of the level-6 camera whose rays grazing a block's edge
    start at the eye: segments of thousands of samples, far past the
    one-lane cap (2 * quality * 17 samples), which the wave continues
    together.  The colour-table and the field gradients equal torch autograd
    of the float64 reference of the same forward within 1e-4 relative, and
    blocks the crop's rays do not reach get no gradient."""
    import torch
    from tests.test_gpu_fullsize import _sample_counts
    ls = _share(6, seed=61)
    W, H = 1920, 1080
    cam = datasets.orbit_camera(35.0, elevation_deg=15.0)
    nb = ls.keys.shape[0]
    params = [api.block_params(cam, W, H, FOV, 6, tuple(int(v) for v in k), 4, ls.fmin, ls.fmax, THR, Q, CUT,
                               exact=True) for k in ls.keys]
    longest, at = 0, None
    for live, p in params:
        if live:
            sc = _sample_counts(p)
            if int(sc.max()) > longest:
                longest = int(sc.max())
                iy, ix = np.unravel_index(int(np.argmax(sc)), sc.shape)
                at = (p.view_x + int(ix), p.view_y + int(iy))
    cap = int(2 * Q * 17)
    say("longest segment %d samples at pixel %s (cap %d)" % (longest, at, cap))
    assert longest > 10 * cap
    cw = ch = 16
    x0 = min(max(at[0] - cw // 2, 0), W - cw)
    y0 = min(max(at[1] - ch // 2, 0), H - ch)
    crop = (x0, y0, cw, ch)
    sub = [i for i, (live, p) in enumerate(params)
           if live and p.view_x < x0 + cw and x0 < p.view_x + p.view_w and p.view_y < y0 + ch and y0 < p.view_y + p.view_h]
    say("crop", crop, "blocks over it", len(sub))

    tf_np = datasets.seeded_tf(61, 256)
    rng = np.random.default_rng(62)
    tf_t = np.clip(tf_np + rng.normal(0.0, 0.05, size=tf_np.shape), 0.0, 1.0).astype(np.float32)
    tf = torch.from_numpy(tf_np).cuda()
    fp = api.FramePlan(ls.frame_blocks(), W, H, FOV, ls.fmin, ls.fmax, THR, Q, CUT, ACC, tf, exact=True)
    fp.set_cameras(cam[None])
    argb = torch.empty((H, W), dtype=torch.int32, device="cuda")
    target = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    tf.copy_(torch.from_numpy(tf_t))
    fp.render(0, argb, target, interpolate=False)
    tf.copy_(torch.from_numpy(tf_np))
    fp.set_tape(True)
    rgba = torch.empty_like(target)
    fp.render(0, argb, rgba, interpolate=False)
    mask = torch.zeros((H, W, 1), dtype=torch.float32, device="cuda")
    for j in range(ch):
        mask[H - 1 - (y0 + j), x0:x0 + cw] = 1.0
    gI = (rgba - target) * mask * mask
    gtf = torch.zeros((256, 4), dtype=torch.float32, device="cuda")
    gb = torch.zeros((nb, 17 ** 3), dtype=torch.float32, device="cuda")
    fp.set_grad_bricks([gb[i] for i in range(nb)])
    fp.adjoint(0, rgba, gI, gtf)
    torch.cuda.synchronize()
    assert fp.status() == 0
    g_tf = gtf.double().cpu().numpy()

    # float64 reference over the crop's samples (the blocks over the crop,
    # in the frame's order)
    blocks = [{"level": 6, "index": tuple(int(v) for v in ls.keys[i]), "depth": 4} for i in sub]
    bricks_np = ls.bricks[sub][:, :17 ** 3].float().cpu().numpy()
    table = SampleTable(orc, blocks, bricks_np, cam, W, H, FOV, ls.fmin, ls.fmax, THR, Q, CUT, ACC, tf_np, crop,
                        max_samples=1 << 16)
    _, runs = np.unique(table.pix * len(sub) + table.blk, return_counts=True)
    say("samples", table.pix.size, "blocks touched", table.touched.size, "longest kept segment", runs.max())
    assert runs.max() > 2 * cap  # (a segment that ends late enough for the wave to continue it)
    tf64 = torch.tensor(tf_np, dtype=torch.float64, device="cuda", requires_grad=True)
    clean = np.where(np.isnan(bricks_np), 0.0, bricks_np)
    b64 = torch.tensor(clean, dtype=torch.float64, device="cuda", requires_grad=True)
    img = table.forward(torch, tf64, b64, "cuda")
    rows = [H - 1 - (y0 + j) for j in range(ch)]
    tgt = target[rows][:, x0:x0 + cw].double()
    got = rgba[rows][:, x0:x0 + cw].double()
    assert float((got - img.detach()).abs().max()) < 2e-3
    (0.5 * ((img - tgt) ** 2).sum()).backward()
    r_tf = tf64.grad.cpu().numpy()
    tf_err = np.linalg.norm(g_tf - r_tf) / np.linalg.norm(r_tf)
    r_b = b64.grad
    subt = torch.as_tensor(sub, device="cuda")
    g_sub = gb[subt].double()
    touched = torch.as_tensor(table.touched, device="cuda")
    num = torch.linalg.norm(g_sub[touched] - r_b[touched])
    den = torch.linalg.norm(r_b[touched])
    say("tf grad rel err", tf_err, "field grad rel err", float(num / den))
    assert np.linalg.norm(r_tf) > 0 and tf_err <= 1e-4
    assert float(den) > 0 and float(num) <= 1e-4 * float(den)
    reached = torch.zeros(nb, dtype=torch.bool, device="cuda")
    reached[subt[touched]] = True
    assert float(gb[~reached].abs().max()) == 0.0
    fp.close()


def test_large_share_full_size_properties():
    import torch
    say("level-8 share: generating")
    ls = _share(8, seed=2029)
    nb = ls.keys.shape[0]
    say("level-8 share", nb, "blocks")
    assert nb == round(0.12 * 256 ** 3)
    _scatter_check(torch, ls)
    W, H = 3840, 2160
    tf_np = datasets.seeded_tf(2029, 256)
    tf = torch.from_numpy(tf_np).cuda()
    rng = np.random.default_rng(2029)
    views = np.stack([datasets.orbit_camera(float(a), elevation_deg=float(e))
                      for a, e in zip(rng.uniform(0, 360, 3), rng.uniform(-30, 30, 3))])
    frames = {}
    for tile in ((8, 8), (16, 4)):
        fp = api.FramePlan(ls.frame_blocks(), W, H, FOV, ls.fmin, ls.fmax, THR, Q, CUT, ACC, tf, tile, exact=True)
        fp.set_cameras(views)
        argb = torch.empty((H, W), dtype=torch.int32, device="cuda")
        say("level-8 frame, tile", tile)
        fp.render(0, argb, None, interpolate=False)
        frames[tile] = argb.clone()
        if tile == (8, 8):
            fp.render(0, argb, None, interpolate=True)
            torch.cuda.synchronize()
            assert fp.status() == 0
            assert torch.equal(argb, frames[tile])
            assert float((argb != 0).float().mean()) > 0.02
            # colour-table gradients: several views at once = sum of each
            fp.set_tape(True)
            rgba = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
            gsum = torch.zeros((256, 4), dtype=torch.float32, device="cuda")
            each = []
            for v in range(3):
                fp.render(v, argb, rgba, interpolate=False)
                g = torch.zeros_like(gsum)
                gI = rgba * 0.5 - 0.25
                fp.adjoint(v, rgba, gI, g)
                fp.adjoint(v, rgba, gI, gsum)
                each.append(g)
            torch.cuda.synchronize()
            tot = sum(each)
            scale = float(tot.abs().max())
            assert scale > 0 and bool(torch.isfinite(gsum).all())
            assert float((gsum - tot).abs().max()) <= 1e-4 * scale
        fp.close()
    assert torch.equal(frames[(8, 8)], frames[(16, 4)])
