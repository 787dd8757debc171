"""Adjoint ray-cast (rv_frame_adjoint) against a float64 reference of the same
forward: image loss L = 1/2 sum over a 64x64 crop of (image - target)^2, the
target rendered with the colour table perturbed by seeded N(0, 0.05) noise.

* forward: the kernel's float image on the crop equals the float64 reference
  within 2e-3 (absolute);
* gradients with respect to the colour table and to every brick value: the
  adjoint (through a torch.autograd.Function wrapper) equals torch autograd of
  the float64 reference within 1e-4 relative (Euclidean norm of the
  difference over the norm of the reference gradient); blocks the crop's rays
  never reach get exactly zero;
* finite differences: central differences of the float64 reference along
  unit directions (the gradient itself and seeded random ones) agree with
  <gradient, direction> within 1e-4 |gradient|.

The sample set of each pixel (termination decisions) comes from the oracle's
fp32 march, so every gradient is taken over the same samples."""
import numpy as np
import pytest

from revise_amd import api, configs, datasets
from tests.adjoint_ref import SampleTable
from tests.gpu_util import say, to_dev

pytestmark = pytest.mark.gpu

FOV, THR, Q, CUT, ACC = 30.0, 0.5, 2.0, 0.99, 0.9999


def _frame_fn():
    import torch

    class FrameRender(torch.autograd.Function):
        """Forward: rv_frame_render (resident bricks, float image);
        backward: rv_frame_adjoint.  tf and bricks must be the tensors the
        plan was built on."""

        @staticmethod
        def forward(ctx, tf, bricks, plan, cam):
            rgba = torch.empty((plan.H, plan.W, 4), dtype=torch.float32, device="cuda")
            argb = torch.empty((plan.H, plan.W), dtype=torch.int32, device="cuda")
            plan.render(cam, argb, rgba, interpolate=False)
            ctx.plan, ctx.cam = plan, cam
            ctx.save_for_backward(rgba, bricks)
            return rgba

        @staticmethod
        def backward(ctx, g):
            rgba, bricks = ctx.saved_tensors
            gtf = torch.zeros((ctx.plan.tf_n, 4), dtype=torch.float32, device="cuda")
            gb = torch.zeros_like(bricks)
            ctx.plan.set_grad_bricks([gb[i] for i in range(gb.shape[0])])
            ctx.plan.adjoint(ctx.cam, rgba, g.contiguous(), gtf)
            return gtf, gb, None, None

    return FrameRender


def _scene(level, depth, W, H, seed):
    import torch
    field = datasets.Field(seed, n_gauss=3)
    wl = configs.uniform(torch, field, level=level, depth=depth)
    api.InterpBatch([(b["sparse"], b["plan"], b["brick"], b["depth"]) for b in wl.blocks]).run()
    torch.cuda.synchronize()
    tf_np = datasets.seeded_tf(seed, 256)
    rng = np.random.default_rng(seed + 1)
    tf_target = np.clip(tf_np + rng.normal(0.0, 0.05, size=tf_np.shape), 0.0, 1.0).astype(np.float32)
    M = datasets.orbit_camera(float(rng.uniform(0, 360)), elevation_deg=25.0)
    return wl, tf_np, tf_target, M


def _plan(wl, W, H, tf_dev):
    fp = api.FramePlan(wl.blocks, W, H, FOV, wl.fmin, wl.fmax, THR, Q, CUT, ACC, tf_dev, exact=True)
    fp.tf_n = tf_dev.shape[0]
    return fp


def _crop_rows(img, H, crop):
    """Kernel image rows are bottom-up (row H-1-vy): crop indexed by vy - y0."""
    x0, y0, cw, ch = crop
    rows = [H - 1 - (y0 + j) for j in range(ch)]
    return img[rows][:, x0:x0 + cw]


@pytest.mark.parametrize("level,depth,W,H,crop", [
    (2, 5, 192, 192, (64, 64, 64, 64)),
    (3, 6, 512, 512, (224, 224, 64, 64)),  # the 512-block 65^3 field at 512x512
])
def test_adjoint_matches_float64_reference(orc, level, depth, W, H, crop):
    import torch
    say("adjoint scene level", level, "depth", depth, "viewport", W, H)
    wl, tf_np, tf_target, M = _scene(level, depth, W, H, seed=40 + level)
    nb = len(wl.blocks)
    tf = to_dev(tf_np).requires_grad_(True)
    bricks = wl.bricks  # (nblocks, V) fp32, the plan's bricks
    fp = _plan(wl, W, H, tf.detach())
    fp.set_cameras(M[None])
    fp.set_tape(True)
    # target image: same field, perturbed colour table
    tft = to_dev(tf_target)
    fpt = _plan(wl, W, H, tft)
    fpt.set_cameras(M[None])
    target = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    fpt.render(0, torch.empty((H, W), dtype=torch.int32, device="cuda"), target, interpolate=False)
    torch.cuda.synchronize()

    # adjoint through the autograd wrapper
    x0, y0, cw, ch = crop
    mask = torch.zeros((H, W, 1), dtype=torch.float32, device="cuda")
    for j in range(ch):
        mask[H - 1 - (y0 + j), x0:x0 + cw] = 1.0
    b_leaf = bricks.detach().requires_grad_(True)
    assert b_leaf.data_ptr() == bricks.data_ptr()
    out = _frame_fn().apply(tf, b_leaf, fp, 0)
    loss = 0.5 * (((out - target) * mask) ** 2).sum()
    loss.backward()
    g_tf = tf.grad.detach().double().cpu().numpy()
    g_b = b_leaf.grad.detach()
    say("loss", float(loss.detach()), "|g_tf|", np.linalg.norm(g_tf))
    assert np.isfinite(g_tf).all() and np.linalg.norm(g_tf) > 0

    # float64 reference on the crop
    bricks_np = bricks.cpu().numpy()
    table = SampleTable(orc, wl.blocks, bricks_np, M, W, H, FOV, wl.fmin, wl.fmax, THR, Q, CUT, ACC,
                        tf_np, crop)
    say("samples", table.pix.size, "blocks touched", table.touched.size)
    dev = "cuda"
    tf64 = torch.tensor(tf_np, dtype=torch.float64, device=dev, requires_grad=True)
    clean = np.where(np.isnan(bricks_np), 0.0, bricks_np)
    b64 = torch.tensor(clean, dtype=torch.float64, device=dev, requires_grad=True)
    tgt = _crop_rows(target, H, crop).double()
    img64 = table.forward(torch, tf64, b64, dev)
    # forward parity of the kernel image on the crop
    got = _crop_rows(out.detach(), H, crop).double()
    assert float((got - img64.detach()).abs().max()) < 2e-3
    L64 = 0.5 * ((img64 - tgt) ** 2).sum()
    L64.backward()
    r_tf = tf64.grad.cpu().numpy()
    assert np.linalg.norm(g_tf - r_tf) <= 1e-4 * np.linalg.norm(r_tf)
    r_b = b64.grad
    gb64 = g_b.double()
    touched = torch.as_tensor(table.touched, device=dev)
    num = torch.linalg.norm(gb64[touched] - r_b[touched])
    den = torch.linalg.norm(r_b[touched])
    say("field grad rel err", float(num / den), "tf grad rel err",
        np.linalg.norm(g_tf - r_tf) / np.linalg.norm(r_tf))
    assert float(den) > 0 and float(num) <= 1e-4 * float(den)
    others = torch.ones(nb, dtype=torch.bool, device=dev)
    others[touched] = False
    assert float(gb64[others].abs().max()) == 0.0 if bool(others.any()) else True

    # central differences of the float64 reference
    rng = np.random.default_rng(7)
    def L_at(tf_val, b_val):
        with torch.no_grad():
            return float(0.5 * ((table.forward(torch, tf_val, b_val, dev) - tgt) ** 2).sum())
    tf0 = torch.tensor(tf_np, dtype=torch.float64, device=dev)
    b0 = torch.tensor(clean, dtype=torch.float64, device=dev)
    finite = torch.as_tensor(~np.isnan(bricks_np), device=dev)
    # unit directions: the gradient itself and seeded random ones.  Table
    # alphas that are exactly 0 are left out: raising one makes exactly
    # transparent pixels visible, where the non-premultiplied colour P / A
    # jumps (the image is not differentiable there; the adjoint, like the
    # reference, takes the derivative from below, colour 0).
    eps = 1e-5
    free = np.ones_like(tf_np, dtype=np.float64)
    free[:, 3] = tf_np[:, 3] != 0
    dirs_tf = [g_tf] + [rng.normal(size=tf_np.shape) for _ in range(3)]
    for u in dirs_tf:
        u = u * free
        u = u / np.linalg.norm(u)
        ut = torch.tensor(u, dtype=torch.float64, device=dev)
        fd = (L_at(tf0 + eps * ut, b0) - L_at(tf0 - eps * ut, b0)) / (2 * eps)
        pred = float((g_tf * u).sum())
        say("fd tf", fd, "adjoint", pred)
        assert abs(fd - pred) <= 1e-4 * np.linalg.norm(g_tf * free), (fd, pred)
    gb_t = gb64 * finite
    ngb = float(torch.linalg.norm(gb_t))
    for k in range(4):
        if k == 0:
            ub = gb_t.clone()
        else:
            ub = torch.zeros_like(b0)
            ub[touched] = torch.randn(ub[touched].shape, dtype=torch.float64, device=dev,
                                      generator=torch.Generator(dev).manual_seed(k))
            ub = ub * finite
        ub = ub / torch.linalg.norm(ub)
        fd = (L_at(tf0, b0 + eps * ub) - L_at(tf0, b0 - eps * ub)) / (2 * eps)
        pred = float((gb_t * ub).sum())
        say("fd field", fd, "adjoint", pred)
        assert abs(fd - pred) <= 1e-4 * ngb, (fd, pred)
    fp.close()
    fpt.close()


def test_adjoint_requires_taped_render():
    """The adjoint refuses a camera that was not rendered with the tape on."""
    import torch
    from revise_amd import RvError
    wl, tf_np, tf_target, M = _scene(1, 3, 32, 32, seed=3)
    fp = _plan(wl, 32, 32, to_dev(tf_np))
    fp.set_cameras(np.stack([M, M]))
    rgba = torch.zeros((32, 32, 4), dtype=torch.float32, device="cuda")
    g = torch.zeros_like(rgba)
    gtf = torch.zeros((256, 4), dtype=torch.float32, device="cuda")
    with pytest.raises(RvError, match="tape"):
        fp.adjoint(0, rgba, g, gtf)
    fp.set_tape(True)
    fp.render(0, torch.empty((32, 32), dtype=torch.int32, device="cuda"), rgba, interpolate=False)
    with pytest.raises(RvError, match="tape"):
        fp.adjoint(1, rgba, g, gtf)
    fp.adjoint(0, rgba, g, gtf)
    torch.cuda.synchronize()
    fp.close()


def test_adjoint_accumulates_and_tf_only(orc):
    """Outputs accumulate over calls; without gradient arrays only the colour
    table gradient is produced, and it is the same.  Float atomics add in a
    run-dependent order, so repeated results agree to rounding (compared
    against the largest entry), not bit for bit."""
    import torch
    wl, tf_np, tf_target, M = _scene(2, 4, 96, 96, seed=9)
    tf = to_dev(tf_np)
    fp = _plan(wl, 96, 96, tf)
    fp.set_cameras(M[None])
    fp.set_tape(True)
    rgba = torch.empty((96, 96, 4), dtype=torch.float32, device="cuda")
    fp.render(0, torch.empty((96, 96), dtype=torch.int32, device="cuda"), rgba, interpolate=False)
    g = torch.randn((96, 96, 4), dtype=torch.float32, device="cuda", generator=torch.Generator("cuda").manual_seed(3))
    g1 = torch.zeros((256, 4), dtype=torch.float32, device="cuda")
    grads = [torch.zeros_like(wl.bricks[i]) for i in range(len(wl.blocks))]
    fp.set_grad_bricks(grads)
    fp.adjoint(0, rgba, g, g1)
    g2 = g1.clone()
    fp.adjoint(0, rgba, g, g2)
    torch.cuda.synchronize()
    scale = float(g1.abs().max())
    assert scale > 0
    assert float((g2 - 2 * g1).abs().max()) <= 1e-4 * scale
    assert any(float(x.abs().max()) > 0 for x in grads)
    fp.set_grad_bricks(None)
    g3 = torch.zeros_like(g1)
    fp.adjoint(0, rgba, g, g3)
    torch.cuda.synchronize()
    assert float((g3 - g1).abs().max()) <= 1e-4 * scale
    fp.close()


def test_adjoint_after_lanes_schedule(orc):
    """The per-lane walk records the same adjoint states: its taped float
    image equals the bin walk's bit for bit, and the gradients agree to the
    rounding of the float atomics."""
    import torch
    wl, tf_np, tf_target, M = _scene(2, 4, 96, 96, seed=12)
    tf = to_dev(tf_np)
    fp = _plan(wl, 96, 96, tf)
    fp.set_cameras(M[None])
    fp.set_tape(True)
    g = torch.randn((96, 96, 4), dtype=torch.float32, device="cuda",
                    generator=torch.Generator("cuda").manual_seed(5))
    out = {}
    for sched in (api.SCHEDULE_BINS, api.SCHEDULE_LANES):
        say("adjoint after schedule", sched)
        fp.set_schedule(sched)
        rgba = torch.empty((96, 96, 4), dtype=torch.float32, device="cuda")
        fp.render(0, torch.empty((96, 96), dtype=torch.int32, device="cuda"), rgba, interpolate=False)
        gt = torch.zeros((256, 4), dtype=torch.float32, device="cuda")
        grads = [torch.zeros_like(wl.bricks[i]) for i in range(len(wl.blocks))]
        fp.set_grad_bricks(grads)
        fp.adjoint(0, rgba, g, gt)
        torch.cuda.synchronize()
        out[sched] = (rgba.clone(), gt, torch.stack(grads))
    a, b = out[api.SCHEDULE_BINS], out[api.SCHEDULE_LANES]
    assert torch.equal(a[0], b[0])
    for x, y in ((a[1], b[1]), (a[2], b[2])):
        scale = float(x.abs().max())
        assert scale > 0
        assert float((x - y).abs().max()) <= 1e-4 * scale
    fp.close()


def test_tf_adjoint_of_small_blocks_one_lane_per_record(orc):
    """Blocks a few pixels wide (level 4 at 96x72: the 8 x 8 tiles of their
    view rectangles are mostly empty) take the colour-table-only adjoint with
    one lane per tape record; its table gradient equals the tile kernel's
    (run with field gradients on) to the rounding of the float atomics."""
    import torch
    W, H = 96, 72
    wl, tf_np, tf_target, M = _scene(4, 2, W, H, seed=21)
    tf = to_dev(tf_np)
    fp = _plan(wl, W, H, tf)
    fp.set_cameras(np.stack([M, datasets.orbit_camera(100.0, elevation_deg=-20.0)]))
    fp.set_tape(True)
    for cam in range(2):
        say("tf adjoint small blocks cam", cam)
        rgba = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
        fp.render(cam, torch.empty((H, W), dtype=torch.int32, device="cuda"), rgba, interpolate=False)
        g = torch.randn((H, W, 4), dtype=torch.float32, device="cuda",
                        generator=torch.Generator("cuda").manual_seed(7 + cam))
        fp.set_grad_bricks(None)
        g_rec = torch.zeros((256, 4), dtype=torch.float32, device="cuda")
        fp.adjoint(cam, rgba, g, g_rec)
        grads = [torch.zeros_like(wl.bricks[i]) for i in range(len(wl.blocks))]
        fp.set_grad_bricks(grads)
        g_tile = torch.zeros((256, 4), dtype=torch.float32, device="cuda")
        fp.adjoint(cam, rgba, g, g_tile)
        torch.cuda.synchronize()
        scale = float(g_tile.abs().max())
        assert scale > 0
        assert float((g_rec - g_tile).abs().max()) <= 1e-4 * scale
    fp.close()


def test_adjoint_with_tiled_bricks(orc):
    """A plan over 4^3 micro-brick bricks renders the same float image and
    gives the same gradients (table and field; the field gradient arrays in
    the x-fastest layout) as one over x-fastest bricks."""
    import torch
    W, H = 96, 96
    wl, tf_np, tf_target, M = _scene(2, 4, W, H, seed=14)
    tf = to_dev(tf_np)
    g = torch.randn((H, W, 4), dtype=torch.float32, device="cuda", generator=torch.Generator("cuda").manual_seed(9))
    out = {}
    for layout in (api.BRICK_LINEAR, api.BRICK_TILED4):
        blocks = []
        for b in wl.blocks:
            c = dict(b)
            c["brick"] = torch.empty(api.brick_floats(b["depth"], layout), dtype=torch.float32, device="cuda")
            blocks.append(c)
        fp = api.FramePlan(blocks, W, H, FOV, wl.fmin, wl.fmax, THR, Q, CUT, ACC, tf, layout=layout, exact=True)
        fp.set_cameras(M[None])
        fp.set_tape(True)
        rgba = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
        say("adjoint layout", layout)
        fp.render(0, torch.empty((H, W), dtype=torch.int32, device="cuda"), rgba, interpolate=True)
        gt = torch.zeros((256, 4), dtype=torch.float32, device="cuda")
        grads = [torch.zeros(api.brick_nodes(b["depth"]), dtype=torch.float32, device="cuda") for b in wl.blocks]
        fp.set_grad_bricks(grads)
        fp.adjoint(0, rgba, g, gt)
        torch.cuda.synchronize()
        assert fp.status() == 0
        out[layout] = (rgba.clone(), gt, torch.stack(grads))
        fp.close()
    a, b = out[api.BRICK_LINEAR], out[api.BRICK_TILED4]
    assert torch.equal(a[0], b[0])
    for x, y in ((a[1], b[1]), (a[2], b[2])):
        scale = float(x.abs().max())
        assert scale > 0
        assert float((x - y).abs().max()) <= 1e-4 * scale
