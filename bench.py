"""Benchmark of the sparse->dense interpolation + ray-cast path on one MI355X.

Headline workload (`value`): the GPU plain drop-in frame -- a block tree
refined to a uniform level 3 (512 blocks), each block a fully refined depth-6
subtree (65^3 nodes, fp32), field = 3 seeded Gaussians sampled at the nodes.
One step = one call of the worker's C ABI, rv_worker_render_scene_part
(VsWorker_cu::renderScenePart) over the whole index box, with the camera
moved every step (a new orbit angle, set with rv_worker_set_state): clear the
1920x1080 ARGB viewport, BackToFrontOrder, every block's parameters,
interpolate every block's brick from its sparse node values, ray-cast the
blocks back to front (step 0.5 voxel = quality 2, early termination at alpha
> 0.99) with the per-block 8-bit composite, find the non-transparent
rectangle and download it into a pinned host buffer.  `value` = frames/s.

The JSON line also carries:
* the same frame with one interpolation and one ray-cast launch per block
  (the reference's structure);
* interpolation of all 512 bricks in one launch: Gvoxels/s and the HBM
  roofline of that kernel (`roofline`);
* the render kernel: exact samples per frame S (reference sampling rule, no
  early termination), samples/s, and its VALU and L1 rates against the
  chip's peaks measured by the library's probe kernels;
* the CPU interpolation configuration's brick (level 2, depth 6, refined
  tree) interpolated on the GPU;
* the batched persistent frame on the same data (interpolation + all blocks in
  one launch, front-to-back in registers);
* the 2000-block mixed-level configuration (levels 2-6, fill 5-100 %, 8
  Gaussians + 3-octave noise): batched persistent frames over 64 cameras
  replayed from one hipGraph vs the plain per-block path on the same data;
* forward + adjoint at 512x512, and the level-8 fp16 share at 4K;
* `cpu_baseline`: the CPU oracle on every host core (see cpu_baseline()).
"""
import argparse
import json
import math
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
# the frame autotuner's choices (one JSON object per line), reused by later runs
TUNE_CACHE = os.environ.get("RV_FRAME_TUNE_CACHE",
                            os.path.join(tempfile.gettempdir(), "revise_amd_frame_tune.jsonl"))
SEED = 2024
W, H = 1920, 1080
FOV, THR, Q2, CUT2 = 30.0, 0.5, 2.0, 0.99
ACC_CUT = 0.9999


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def events_time(torch, fn, reps):
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / 1e3 / reps


class PlainFrame:
    """Per-block drop-in path over a workload (one viewport)."""

    def __init__(self, torch, wl, M, quality, cutoff, tf):
        from revise_amd import api
        self.api, self.torch = api, torch
        self.wl, self.tf = wl, tf
        keys = np.array([(b["level"],) + tuple(b["index"]) for b in wl.blocks], dtype=np.uint32)
        order = api.mixed_back_to_front(keys, api.eye_from_view(M))
        self.items = []
        for i in order:
            b = wl.blocks[int(i)]
            live, p = api.block_params(M, W, H, FOV, b["level"], b["index"], b["depth"], wl.fmin, wl.fmax,
                                       THR, quality, cutoff)
            self.items.append((b, live, p))
        self.vp = torch.empty((H, W), dtype=torch.int32, device="cuda")
        # the cropped part is downloaded into a reused pinned host buffer
        self.host = torch.empty(H * W, dtype=torch.int32, pin_memory=True)
        self.batch = None
        self.share = None

    def download(self, r):
        """The cropped part (x, y, w, h) of the viewport, as an h x w host array."""
        x, y, w, h = r
        out = self.host[: w * h].view(h, w)
        out.copy_(self.vp[y:y + h, x:x + w], non_blocking=True)
        self.torch.cuda.synchronize()
        return out

    def fused_step(self, crop=True):
        """Drop-in worker frame in two launches (bit-identical to step()), as
        the worker's fused path runs it: the share's own bricks in 4^3
        micro-bricks."""
        api = self.api
        if self.batch is None:
            self.tiled = [self.torch.empty(api.brick_floats(b["depth"], api.BRICK_TILED4), dtype=self.torch.float32,
                                           device="cuda") for b, _, _ in self.items]
            self.batch = api.InterpBatch([(b["sparse"], b["plan"], t, b["depth"])
                                          for (b, _, _), t in zip(self.items, self.tiled)], api.BRICK_TILED4)
            self.share = api.ShareRenderer([(p, t) for (b, live, p), t in zip(self.items, self.tiled) if live],
                                           self.tf, W, H, layout=api.BRICK_TILED4)
        api.clear_viewport(self.vp, W, H)
        self.batch.run()
        self.share.render(self.vp)
        if crop:
            r = api.bound_rect(self.vp, W, H)
            if r is not None:
                return self.download(r)
        return None

    def step(self, interpolate=True, crop=True):
        api = self.api
        api.clear_viewport(self.vp, W, H)
        for b, live, p in self.items:
            if interpolate:
                api.interp_brick_planned(b["sparse"], b["plan"], b["depth"], b["brick"])
            if live:
                api.render_block_argb(p, b["brick"], self.tf, self.vp)
        if crop:
            r = api.bound_rect(self.vp, W, H)
            if r is not None:
                return self.download(r).clone()
        return None

    def samples(self):
        c = self.torch.zeros(1, dtype=self.torch.int64, device="cuda")
        for b, live, p in self.items:
            if live:
                self.api.block_sample_count(p, c)
        return int(c.item())


class WorkerFrame:
    """The drop-in through the worker's C ABI: the 512 subtrees registered
    with rv_worker_add_subtree (host n2i and values), and per step
    rv_worker_set_state with the step's camera, then
    rv_worker_render_scene_part over the whole index box into a pinned host
    buffer."""

    def __init__(self, torch, wl, tf_np, angle0, fused=True):
        import ctypes as C
        from revise_amd import api, lib, RenderState, check
        self.C, self.api, self.lib, self.check = C, api, lib, check
        self.w = api.Worker(0)
        sparse = wl.sparse.cpu().numpy()
        n2i = wl.n2i["full"]
        for i, b in enumerate(wl.blocks):
            self.w.add_subtree(b["level"], b["index"], b["depth"], n2i, sparse[i])
        self.w.set_fused(fused)
        self.tf = np.ascontiguousarray(tf_np, dtype=np.float32)
        self.view = np.zeros(16, np.float32)
        self.state = RenderState(self.view.ctypes.data, W, H, FOV, Q2, THR, CUT2, wl.fmin, wl.fmax,
                                 (C.c_float * 3)(0.0, 0.0, 0.0), self.tf.ctypes.data, self.tf.shape[0],
                                 api.FIELD_ARGB, 0.5, 1.0, None, 0, 0.0, 1.0, 0)
        self.out = torch.empty(W * H, dtype=torch.int32, pin_memory=True)
        self.out_p = C.cast(self.out.data_ptr(), C.POINTER(C.c_uint32))
        self.rect = (C.c_int32 * 4)()
        self.bmin = (C.c_uint32 * 3)(0, 0, 0)
        self.bmax = (C.c_uint32 * 3)(8, 8, 8)
        self.cancel = C.c_int(0)
        self.angle0 = angle0

    def camera(self, k):
        return datasets_mod().orbit_camera(self.angle0 + 7.5 * k)

    def cameras(self, n):
        """The view matrices of steps 0 .. n - 1, made before the timed loop
        (the application's part, not the renderer's)."""
        self._cams = [np.ascontiguousarray(self.camera(k).reshape(16), dtype=np.float32) for k in range(n)]

    def step(self, k):
        C = self.C
        cams = getattr(self, "_cams", None)
        self.view[:] = cams[k] if cams is not None and k < len(cams) else self.camera(k).reshape(16)
        L = self.lib()
        self.check(L.rv_worker_set_state(self.w.h, C.byref(self.state)))
        self.check(L.rv_worker_render_scene_part(self.w.h, 3, self.bmin, self.bmax, C.byref(self.cancel),
                                                  self.rect, self.out_p, W * H))
        return tuple(self.rect)

    def close(self):
        self.w.close()


def profiled_traffic(key="traffic_bytes_per_launch"):
    """HBM bytes per launch of the batched interpolation, as measured by the
    FETCH_SIZE / WRITE_SIZE counter runs summarised under profiles/ (None if
    that summary is absent)."""
    path = os.path.join(ROOT, "profiles", "interp_batch_512x65cubed_hbm_bytes.md")
    try:
        for line in open(path):
            if line.startswith(key + ":"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def profiled(path, key):
    """A figure from a profile summary under profiles/ ("key: value" line),
    or None."""
    try:
        for line in open(os.path.join(ROOT, "profiles", path)):
            if line.startswith(key + ":"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def tiled_blocks(torch, wl):
    """The workload's blocks with bricks of their own in 4^3 micro-bricks,
    interpolated (the batched frames read this layout)."""
    from revise_amd import api
    blocks = [dict(b, brick=torch.empty(api.brick_floats(b["depth"], api.BRICK_TILED4), dtype=torch.float32,
                                        device="cuda")) for b in wl.blocks]
    api.InterpBatch([(b["sparse"], b["plan"], b["brick"], b["depth"]) for b in blocks], api.BRICK_TILED4).run()
    return blocks


def frame_plan(torch, wl, tf, quality, cutoff, tile=(0, 0), blocks=None):
    """A batched frame plan over the workload; blocks: tiled_blocks(...) for
    bricks in micro-bricks (default: the workload's x-fastest bricks)."""
    from revise_amd import api
    layout = api.BRICK_LINEAR if blocks is None else api.BRICK_TILED4
    return api.FramePlan(blocks if blocks is not None else wl.blocks, W, H, FOV, wl.fmin, wl.fmax, THR, quality,
                         cutoff, ACC_CUT, tf, tile, layout)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(wl, tf_np, cfg1, runs=5):
    """The oracle on the host, every core this process may use:
    * interpolation of the CPU configuration's brick (level-2 root, depth 6,
      65^3 nodes, a partly refined tree, so the midpoint rule does real
      work): Gvoxels/s on 1 core (median of `runs`) and on all cores (each
      core interpolating its own copy, median of `runs` rounds);
    * the per-block ray-cast of config 2's frame (512 blocks, 1920x1080,
      back to front, 8-bit composite) on a stratified sample of the rows --
      16 bands of 8 rows spread evenly over the frame, 128 of 1080 rows --
      on all cores, scaled to the frame: frames/s, median of `runs`."""
    import ctypes as C
    import statistics
    import threading
    from oracle import oracle as orc
    from revise_amd import api
    ncores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    n2i1, d1, sp1 = cfg1["n2i"], cfg1["depth"], cfg1["sparse"]
    V1 = (2 ** d1 + 1) ** 3
    t1 = []
    for _ in range(runs + 1):
        t0 = time.perf_counter()
        orc.interpolate(sp1, n2i1, d1)
        t1.append(time.perf_counter() - t0)
    t_1core = statistics.median(t1[1:])

    def all_cores_round():
        ths = [threading.Thread(target=orc.interpolate, args=(sp1, n2i1, d1)) for _ in range(ncores)]
        t0 = time.perf_counter()
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        return time.perf_counter() - t0

    all_cores_round()
    t_all = statistics.median(all_cores_round() for _ in range(runs))
    # ray-cast: the oracle's per-block 8-bit frame over sampled rows
    bricks = wl.bricks.cpu().numpy()
    angle = float(np.random.default_rng(SEED).uniform(0.0, 360.0))
    M = datasets_mod().orbit_camera(angle)
    keys = np.array([(b["level"],) + tuple(b["index"]) for b in wl.blocks], dtype=np.uint32)
    order = api.mixed_back_to_front(keys, api.eye_from_view(M))
    items = []
    for i in order:
        b = wl.blocks[int(i)]
        live, po = orc.block_params(M, W, H, FOV, b["level"], b["index"], b["depth"], wl.fmin, wl.fmax,
                                    THR, Q2, CUT2)
        if live:
            items.append((po, bricks[int(i)].ctypes.data_as(C.POINTER(C.c_float))))
    bands = [(int(k * H / 16), int(k * H / 16) + 8) for k in range(16)]
    L = orc.lib()
    tp = np.ascontiguousarray(tf_np).ctypes.data_as(C.POINTER(C.c_float))
    vp = np.zeros((H, W), np.uint32)
    vpp = vp.ctypes.data_as(C.POINTER(C.c_uint32))

    def worker(my_bands):
        for r0, r1 in my_bands:
            for po, bp in items:
                if po.view_y < r1 and po.view_y + po.view_h > r0:
                    L.orc_render_block(C.byref(po), bp, tp, tf_np.shape[0], vpp, r0, r1)

    def render_round():
        vp[:] = 0
        ths = [threading.Thread(target=worker, args=(bands[j::ncores],)) for j in range(min(ncores, len(bands)))]
        t0 = time.perf_counter()
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        return time.perf_counter() - t0

    t_rows = statistics.median(render_round() for _ in range(runs))
    rows = sum(r1 - r0 for r0, r1 in bands)
    fps = 1.0 / (t_rows * H / rows)
    return {"value": fps, "unit": "frames/s", "cores": ncores, "kind": "port",
            "sample": "CPU oracle (the C++ restatement of the reference's interpolation and Argb ray-cast), "
                      "%d host threads: the per-block ray-cast of config 2's 512 blocks at 1920x1080 over 16 "
                      "bands of 8 rows spread evenly over the frame (%d of %d rows, a stratified sample), "
                      "scaled to the frame, median of %d runs; interpolation: the CPU configuration's "
                      "level-2 depth-6 brick (%d of %d nodes given)" % (ncores, rows, H, runs, n2i1.shape[0], V1),
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "interp_gvox_s_1core": V1 / t_1core / 1e9,
            "interp_gvox_s_all_cores": ncores * V1 / t_all / 1e9,
            "interp_ms_1core": t_1core * 1e3}


def datasets_mod():
    from revise_amd import datasets
    return datasets


def large_share_bench(torch, args, tf_np, log):
    """Large single-GPU share: level 8 (256^3 block grid), the 12% of blocks
    with the largest centre values, each a full depth-4 subtree (17^3 fp32
    sparse nodes -> 17^3 fp16 brick); field: 32 seeded Gaussians + 4 octaves
    of value noise.  Interpolation of every brick in one launch; batched
    resident-brick frames at 3840x2160; forward + colour-table adjoint over 8
    seeded views (transfer-function fitting)."""
    from revise_amd import api, configs, datasets
    t0 = time.time()
    ls = configs.large_share(torch, SEED + 5, level=8, depth=4, occupancy=0.12, n_gauss=32, octaves=4)
    torch.cuda.synchronize()
    nb = ls.keys.shape[0]
    log("config 5 setup %.1fs: %d blocks, %.3g dense voxels, field [%g, %g]" % (
        time.time() - t0, nb, ls.dense_nodes, ls.fmin, ls.fmax))
    ib = api.InterpBatch(ls.brick_jobs())
    ib.run()
    torch.cuda.synchronize()
    t_i = events_time(torch, ib.run, 5)
    V, N = ls.dense_nodes, ls.sparse_nodes
    B_alg = 2 * V + 10 * N
    # full subtrees: every brick value is its node's sparse value (scatter)
    nn = 2 ** ls.depth + 1
    fi = torch.from_numpy(((ls.n2i[:, 2].astype(np.int64) * nn + ls.n2i[:, 1]) * nn + ls.n2i[:, 0])).cuda()
    Nn = ls.n2i.shape[0]
    for c0 in range(0, nb, 1 << 18):
        c1 = min(nb, c0 + (1 << 18))
        assert torch.equal(ls.bricks[c0:c1, :nn ** 3][:, fi], ls.sparse[c0:c1, :Nn].half()), "interp"
    B_hbm = 2 * V + 4 * N  # the shared plan's coordinates are read from cache, not HBM
    log("config 5 interp %.3f ms, %.1f Gvox/s, %.0f GB/s counting coordinates, %.0f GB/s of values and "
        "bricks (fp16 bricks)" % (t_i * 1e3, V / t_i / 1e9, B_alg / t_i / 1e9, B_hbm / t_i / 1e9))
    W5, H5 = 3840, 2160
    rng = np.random.default_rng(SEED + 5)
    views = np.stack([datasets.orbit_camera(float(a), elevation_deg=float(e))
                      for a, e in zip(rng.uniform(0, 360, 8), rng.uniform(-30, 30, 8))])
    tf_buf = torch.from_numpy(tf_np).cuda()
    t0 = time.time()
    fp = api.FramePlan(ls.frame_blocks(), W5, H5, FOV, ls.fmin, ls.fmax, THR, Q2, CUT2, ACC_CUT, tf_buf)
    fp.set_cameras(views)
    t_setup = time.time() - t0
    argb = torch.empty((H5, W5), dtype=torch.int32, device="cuda")
    tile = fp.autotune(argb, False, TUNE_CACHE)
    fp.render(0, argb, None, False)
    t_f = events_time(torch, lambda: fp.render(0, argb, None, False), 5)
    # the interpolating frame in one launch and split in two (the full
    # autotuner over 2 M blocks at 4K would take minutes; time both)
    t_fi1 = events_time(torch, lambda: fp.render(0, argb, None, True), 3)
    fp.set_split(True)
    fp.render(0, argb, None, True)
    t_fi2 = events_time(torch, lambda: fp.render(0, argb, None, True), 3)
    t_fi = min(t_fi1, t_fi2)
    assert fp.status() == 0
    # the same frame from the interpolating launches and from resident bricks
    a1 = argb.clone()
    fp.render(0, argb, None, False)
    assert torch.equal(a1, argb)
    log("config 5 frame 3840x2160: %.3f ms resident, %.3f ms with interpolation (one launch %.3f, split %.3f; "
        "camera setup %.1fs, tile %s)" % (t_f * 1e3, t_fi * 1e3, t_fi1 * 1e3, t_fi2 * 1e3, t_setup, tile))
    # forward + colour-table adjoint over the 8 views; targets from the
    # perturbed table N(0, 0.05)
    tf_t = np.clip(tf_np + rng.normal(0.0, 0.05, size=tf_np.shape), 0.0, 1.0).astype(np.float32)
    rgba = torch.empty((H5, W5, 4), dtype=torch.float32, device="cuda")
    targets = torch.empty((8, H5, W5, 4), dtype=torch.float32, device="cuda")
    tf_buf.copy_(torch.from_numpy(tf_t))
    for v in range(8):
        fp.render(v, argb, targets[v], False)
    tf_buf.copy_(torch.from_numpy(tf_np))
    gtf = torch.zeros((tf_np.shape[0], 4), dtype=torch.float32, device="cuda")

    def fwd8():
        for v in range(8):
            fp.render(v, argb, rgba, False)

    def fwd_adj8():
        gtf.zero_()
        for v in range(8):
            fp.render(v, argb, rgba, False)
            fp.adjoint(v, rgba, rgba - targets[v], gtf)

    t_fw_fast = events_time(torch, fwd8, 2)
    fp.set_exact(True)  # the sequence the adjoint differentiates
    t_fw = events_time(torch, fwd8, 2)
    fp.set_tape(True)
    fwd_adj8()
    t_fa = events_time(torch, fwd_adj8, 2)
    assert bool(torch.isfinite(gtf).all()) and float(gtf.abs().max()) > 0
    log("config 5 8 views: forward %.3f ms, forward+adjoint (colour table) %.3f ms, ratio %.2f"
        % (t_fw * 1e3, t_fa * 1e3, t_fa / t_fw))
    res = {
        "workload": "level 8 (256^3 block grid), %d blocks (12%%) x 17^3 fp16 bricks = %.3g voxels, "
                    "32 Gaussians + 4-octave noise; 3840x2160; 8 seeded views" % (nb, V),
        "blocks": nb, "dense_voxels": V, "sparse_nodes": N,
        "interp_ms": t_i * 1e3, "interp_gvox_s": V / t_i / 1e9,
        "interp_gb_s_with_coordinates": B_alg / t_i / 1e9,
        "interp_gb_s_values_and_bricks": B_hbm / t_i / 1e9,
        "interp_frac_of_hbm_peak": B_hbm / t_i / 8e12,
        "interp_note": "2 B per fp16 voxel written + 4 B per sparse value read; the 6 B of node "
                       "coordinates per node come from the one shared plan, held in cache",
        "frame_4k_ms": t_f * 1e3, "frame_4k_fps": 1.0 / t_f, "frame_4k_with_interp_ms": t_fi * 1e3,
        "frame_4k_with_interp_one_launch_ms": t_fi1 * 1e3, "frame_4k_with_interp_split_ms": t_fi2 * 1e3,
        "camera_setup_s_8_views": t_setup, "tile": list(tile),
        "fwd_8views_ms": t_fw * 1e3, "fwd_8views_fast_sequence_ms": t_fw_fast * 1e3,
        "fwd_adjoint_8views_ms": t_fa * 1e3, "ratio": t_fa / t_fw}
    fp.close()
    del ib, ls, targets
    torch.cuda.empty_cache()
    return res


def mixed_bench(torch, args, tf, log):
    """Config 3: 2000 blocks of levels 2-6 (fill 5-100%), 1920x1080: the
    batched persistent frames replayed from a hipGraph over args.cams
    cameras, beside the drop-in forms on the same blocks (the fused share
    path -- one interpolation batch and one share launch with the per-block
    8-bit composite -- and one launch pair per block)."""
    from revise_amd import configs, datasets
    t0 = time.time()
    field3 = datasets.Field(SEED + 3, n_gauss=8, noise_octaves=3)
    wl3 = configs.mixed(torch, field3, nblocks=2000, depth=6, seed=SEED + 3)
    cams = np.stack([datasets.orbit_camera(a, elevation_deg=e) for a, e in
                     zip(np.random.default_rng(SEED + 3).uniform(0, 360, args.cams),
                         np.random.default_rng(SEED + 4).uniform(-45, 45, args.cams))])
    torch.cuda.synchronize()
    log("config 3 setup %.1fs: %d blocks, %d sparse nodes" % (time.time() - t0, len(wl3.blocks),
                                                              wl3.sparse_nodes))
    fp3 = frame_plan(torch, wl3, tf, Q2, CUT2, blocks=tiled_blocks(torch, wl3))
    fp3.set_cameras(cams)
    argb3 = torch.empty((H, W), dtype=torch.int32, device="cuda")
    fp3.autotune(argb3, False, TUNE_CACHE)  # grid of the resident-brick kernel (split frames use it)
    tile3 = fp3.autotune(argb3, True, TUNE_CACHE)
    frames = torch.empty((args.cams, H, W), dtype=torch.int32, device="cuda")
    fp3.graph_build(frames, H * W, True)
    fp3.graph_launch()
    torch.cuda.synchronize()
    t_graph = events_time(torch, fp3.graph_launch, 2) / args.cams
    assert fp3.status() == 0
    split3 = fp3.split()
    log("config 3 batched: %.3f ms/frame over %d cameras (graph), tile %s, %s" % (
        t_graph * 1e3, args.cams, tile3, "two launches" if split3 else "one launch"))
    fp3.close()
    del frames
    plain3 = PlainFrame(torch, wl3, cams[0], Q2, CUT2, tf)
    plain3.fused_step(crop=False)
    t_fused3 = events_time(torch, lambda: plain3.fused_step(crop=False), 5)
    log("config 3 drop-in fused (interpolation batch + share launch): %.3f ms/frame" % (t_fused3 * 1e3))
    plain3.step()
    t_plain3 = events_time(torch, lambda: plain3.step(crop=False), 2)
    log("config 3 plain per-block: %.3f ms/frame" % (t_plain3 * 1e3))
    return {"workload": "2000 blocks, levels 2-6, 65^3 fp32 bricks, fill 5-100%%, 8 Gaussians "
                        "+ 3-octave noise, 1920x1080, %d cameras" % args.cams,
            "batched_graph_fps": 1.0 / t_graph, "batched_ms_per_frame": t_graph * 1e3,
            "dropin_fused_share_ms_per_frame": t_fused3 * 1e3,
            "speedup_vs_dropin_fused": t_fused3 / t_graph,
            "plain_per_block_fps": 1.0 / t_plain3, "speedup_vs_plain": t_plain3 / t_graph,
            "tile": list(tile3), "split_launches": split3, "sparse_nodes": wl3.sparse_nodes}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--cpu-runs", type=int, default=5, help="median over this many CPU baseline runs")
    ap.add_argument("--cams", type=int, default=64)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-large", action="store_true", help="skip the level-8 fp16 share")
    ap.add_argument("--only-large", action="store_true", help="run the level-8 fp16 share only")
    ap.add_argument("--only-adjoint", action="store_true",
                    help="512x512 forward with tape x3 and adjoint x3 (profiler runs)")
    ap.add_argument("--only-render", action="store_true",
                    help="run the render kernels only: share kernel x3, batched frame (resident "
                         "bricks, 8x8 bins) x3 at 1920x1080, then x3 at 512x512 (profiler runs)")
    ap.add_argument("--only-dropin", action="store_true",
                    help="the drop-in worker frames only: 5 warm-up and 20 timed steps (profiler runs)")
    ap.add_argument("--only-mixed", action="store_true",
                    help="config 3 only: batched graph frames, the fused drop-in and per-block launches")
    ap.add_argument("--only-interp", action="store_true",
                    help="time a 1 GiB device copy and the batched interpolation only (profiler runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported (the multi-GPU "
              "sort-last exchange is a design, not built)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    from revise_amd import api, configs, datasets, lib
    lib()
    torch.cuda.set_device(0)
    out = {}
    if args.only_large:
        print(json.dumps(large_share_bench(torch, args, datasets.seeded_tf(SEED, 256), log)))
        return
    if args.only_mixed:
        tf_m = torch.from_numpy(datasets.seeded_tf(SEED, 256)).cuda()
        print(json.dumps(mixed_bench(torch, args, tf_m, log)))
        return

    # ---------------- config 2: 512 level-3 bricks, plain drop-in frame
    t0 = time.time()
    field2 = datasets.Field(SEED, n_gauss=3)
    wl = configs.uniform(torch, field2, level=3, depth=6)
    tf_np = datasets.seeded_tf(SEED, 256)
    tf = torch.from_numpy(tf_np).cuda()
    angle = float(np.random.default_rng(SEED).uniform(0.0, 360.0))
    M = datasets.orbit_camera(angle)
    plain = PlainFrame(torch, wl, M, Q2, CUT2, tf)
    torch.cuda.synchronize()
    log("config 2 setup %.1fs" % (time.time() - t0))
    if args.only_interp:
        # plain device copy of known size (1 GiB read + 1 GiB written): the
        # calibration for the HBM byte counters of a profiler run
        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        batch = api.InterpBatch([(b["sparse"], b["plan"], b["brick"], b["depth"]) for b in wl.blocks])
        batch.run()
        log("copy 1 GiB: %.3f ms" % (events_time(torch, lambda: dst.copy_(src), 3) * 1e3))
        log("batched interp: %.3f ms" % (events_time(torch, batch.run, 10) * 1e3))
        own = [b["plan"].clone() for b in wl.blocks]
        batch_own = api.InterpBatch([(b["sparse"], pl, b["brick"], b["depth"]) for b, pl in zip(wl.blocks, own)])
        batch_own.run()
        log("batched interp, private plans: %.3f ms" % (events_time(torch, batch_own.run, 10) * 1e3))
        return
    if args.only_dropin:
        wf = WorkerFrame(torch, wl, tf_np, angle)
        wf.cameras(25)
        for k in range(5):
            wf.step(k)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(20):
            wf.step(5 + k)
        torch.cuda.synchronize()
        log("drop-in worker frame %.3f ms" % ((time.perf_counter() - t0) / 20 * 1e3))
        wf.close()
        return
    if args.only_adjoint:
        plain.step(crop=False)  # interpolates the bricks the frame plan reads
        fpa = api.FramePlan(wl.blocks, 512, 512, FOV, wl.fmin, wl.fmax, THR, Q2, CUT2, ACC_CUT, tf)
        fpa.set_cameras(M[None])
        fpa.set_tape(True)
        argb_a = torch.empty((512, 512), dtype=torch.int32, device="cuda")
        rgba_a = torch.empty((512, 512, 4), dtype=torch.float32, device="cuda")
        gb_a = torch.zeros_like(wl.bricks)
        fpa.set_grad_bricks([gb_a[i] for i in range(gb_a.shape[0])])
        gtf_a = torch.zeros((tf_np.shape[0], 4), dtype=torch.float32, device="cuda")
        for _ in range(3):
            fpa.render(0, argb_a, rgba_a, False)
            fpa.adjoint(0, rgba_a, rgba_a * 0.5, gtf_a)
        torch.cuda.synchronize()
        fpa.close()
        return
    if args.only_render:
        plain.step(crop=False)
        plain.fused_step()
        for _ in range(3):
            plain.share.render(plain.vp)
        # the exact sample sequence of the same share
        from revise_amd import RENDER_EXACT
        exact_items = []
        for (b, live, p), t in zip(plain.items, plain.tiled):
            if live:
                q = api.BlockParams.from_buffer_copy(p)
                q.flags |= RENDER_EXACT
                exact_items.append((q, t))
        share_exact = api.ShareRenderer(exact_items, tf, W, H, layout=api.BRICK_TILED4)
        for _ in range(3):
            share_exact.render(plain.vp)
        fpr = frame_plan(torch, wl, tf, Q2, CUT2, tile=(8, 8))
        fpr.set_cameras(M[None])
        argb_r = torch.empty((H, W), dtype=torch.int32, device="cuda")
        for _ in range(3):
            fpr.render(0, argb_r, None, False)
        fpr.close()
        fps = api.FramePlan(wl.blocks, 512, 512, FOV, wl.fmin, wl.fmax, THR, Q2, CUT2, ACC_CUT, tf, (8, 8))
        fps.set_cameras(M[None])
        argb_s = torch.empty((512, 512), dtype=torch.int32, device="cuda")
        for _ in range(3):
            fps.render(0, argb_s, None, False)
        torch.cuda.synchronize()
        log("share %.3f ms, frame 512x512 %.3f ms" % (events_time(torch, lambda: plain.share.render(plain.vp), 3) * 1e3,
                                                     events_time(torch, lambda: fps.render(0, argb_s, None, False), 3) * 1e3))
        fps.close()
        return
    # the headline: rv_worker_render_scene_part through the C ABI, a new
    # camera every step
    wf = WorkerFrame(torch, wl, tf_np, angle)
    wf.cameras(args.warmup + args.steps)
    log("config 2 worker: 512 subtrees registered")
    for k in range(args.warmup):
        wf.step(k)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        rect = wf.step(args.warmup + k)
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / args.steps
    log("drop-in worker frame through the C ABI, moving camera: %.3f ms (last crop %s)" % (t_step * 1e3, rect))
    wf.close()
    # the reference's structure: one interpolation and one ray-cast launch per block
    wpb = WorkerFrame(torch, wl, tf_np, angle, fused=False)
    wpb.cameras(3)
    wpb.step(0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(2):
        wpb.step(1 + k)
    torch.cuda.synchronize()
    t_step_pb = (time.perf_counter() - t0) / 2
    wpb.close()
    log("worker frame, per-block launches: %.3f ms" % (t_step_pb * 1e3))
    # the same frame composed from Python at one fixed camera: viewport
    # clear, the interpolation batch, the share launch that tests every view
    # rectangle (no screen bins), the bound rectangle and a blocking download
    for _ in range(2):
        plain.fused_step()
    t_fixed = events_time(torch, lambda: plain.fused_step(), 10)
    t_share = events_time(torch, lambda: plain.share.render(plain.vp), 10)
    # the drop-in frame's own interpolation: the same 512 bricks into the
    # micro-brick layout (80 floats per 4^3 cells: 1.43x the x-fastest bytes)
    t_interp_t4 = events_time(torch, plain.batch.run, 20)

    batch = api.InterpBatch([(b["sparse"], b["plan"], b["brick"], b["depth"]) for b in wl.blocks])
    batch.run()
    t_interp = events_time(torch, batch.run, 20)
    # the same bricks with one private copy of the plan per brick: no plan
    # reuse across bricks in the caches (every subtree of a real field has its
    # own node list)
    own_plans = [b["plan"].clone() for b in wl.blocks]
    batch_own = api.InterpBatch([(b["sparse"], pl, b["brick"], b["depth"]) for b, pl in zip(wl.blocks, own_plans)])
    batch_own.run()
    t_interp_own = events_time(torch, batch_own.run, 20)
    del batch_own, own_plans
    V_total, N_total = wl.dense_nodes, wl.sparse_nodes
    B_alg = 4 * V_total + 10 * N_total
    log("batched interp %.3f ms, %.1f Gvox/s, %.0f GB/s" % (t_interp * 1e3, V_total / t_interp / 1e9,
                                                          B_alg / t_interp / 1e9))
    log("batched interp, private plans %.3f ms" % (t_interp_own * 1e3))
    t_interp_plain = events_time(torch, lambda: [api.interp_brick_planned(b["sparse"], b["plan"], 6, b["brick"])
                                                 for b in wl.blocks], 3)
    t_render = events_time(torch, lambda: plain.step(interpolate=False, crop=False), 3)
    S = plain.samples()
    log("per-block render %.3f ms, share launch %.3f ms, S = %.3g samples/frame, %.3g samples/s (share)"
        % (t_render * 1e3, t_share * 1e3, S, S / t_share))

    tblocks = tiled_blocks(torch, wl)
    fp = frame_plan(torch, wl, tf, Q2, CUT2, blocks=tblocks)
    fp.set_cameras(M[None])
    argb = torch.empty((H, W), dtype=torch.int32, device="cuda")
    tune_cache = TUNE_CACHE
    fp.autotune(argb, False, tune_cache)  # grid of the resident-brick kernel
    tile = fp.autotune(argb, True, tune_cache)
    fp.render(0, argb, None, True)
    t_batched = events_time(torch, lambda: fp.render(0, argb, None, True), 5)
    t_batched_res = events_time(torch, lambda: fp.render(0, argb, None, False), 5)
    assert fp.status() == 0
    split2 = fp.split()
    log("batched frame %.3f ms (resident bricks %.3f ms), tile %s, %s" % (
        t_batched * 1e3, t_batched_res * 1e3, tile, "two launches" if split2 else "one launch"))
    fp.close()

    out.update({
        "metric": "sparse→dense Gvoxels/s; 1080p ray-cast frames/s; fwd+adjoint ms/frame",
        "value": 1.0 / t_step,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Gaussian field sampled at block-tree nodes)",
        "config": {"workload": "plain drop-in frame through rv_worker_render_scene_part, a new camera every "
                               "step: 512 level-3 blocks x 65^3 fp32 bricks (full fill), 1920x1080 ARGB, "
                               "quality 2, alpha cutoff 0.99, block order + parameters, interpolate every "
                               "brick, ray-cast back to front with the per-block 8-bit composite (fast sample "
                               "sequence), crop + download into a pinned host buffer",
                   "blocks": len(wl.blocks), "brick_nodes": 65 ** 3, "viewport": [W, H], "seed": SEED,
                   "camera_angle_deg": angle},
        "per_block_launches_frame_ms": t_step_pb * 1e3,
        "python_composed_frame_fixed_camera_ms": t_fixed * 1e3,
        "interp_gvox_s": V_total / t_interp / 1e9,
        "interp_ms_batched": t_interp * 1e3,
        "interp_ms_batched_private_plans": t_interp_own * 1e3,
        "interp_gvox_s_private_plans": V_total / t_interp_own / 1e9,
        "interp_ms_per_block_launches": t_interp_plain * 1e3,
        "interp_ms_batched_micro_bricks": t_interp_t4 * 1e3,
        "interp_frac_micro_bricks": B_alg / t_interp_t4 / HBM_PEAK,
        "render_share_ms": t_share * 1e3,
        "render_per_block_launches_ms": t_render * 1e3,
        "render_samples_per_frame": S,
        "render_gsamples_s": S / t_share / 1e9,
        "batched_frame_ms": t_batched * 1e3,
        "batched_frame_fps": 1.0 / t_batched,
        "batched_frame_resident_ms": t_batched_res * 1e3,
        "batched_tile": list(tile),
        "batched_split_launches": split2,
        "roofline": {"kernel": "interp_batch_kernel (512 bricks, one launch)", "bound": "hbm",
                     "achieved": B_alg / t_interp / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": B_alg / t_interp / HBM_PEAK, "traffic": profiled_traffic(),
                     "alg_bytes_per_launch": B_alg,
                     # the HBM bytes actually moved (the shared plan's coordinates come from cache)
                     "traffic_gb_s": (profiled_traffic() or 0) / t_interp / 1e9,
                     "traffic_frac": (profiled_traffic() or 0) / t_interp / HBM_PEAK,
                     "traffic_private_plans": profiled_traffic("traffic_bytes_per_launch_private_plans"),
                     "note": "achieved = algorithmic bytes (4 B per dense voxel written + 10 B per sparse "
                             "node read) / kernel time; traffic = HBM bytes per launch from rocprofv3 "
                             "FETCH_SIZE/WRITE_SIZE scaled by a 1 GiB copy, traffic_frac its rate "
                             "against the peak; the 512 bricks share one plan, which stays in cache, so "
                             "fewer bytes cross HBM than the algorithmic count (private plans, every byte "
                             "read: interp_ms_batched_private_plans, traffic_private_plans). "
                             "This is the step's HBM-bound kernel, in the interpolator's own (x-fastest) "
                             "layout; the drop-in frame interpolates into micro-bricks "
                             "(interp_ms_batched_micro_bricks). The step's largest kernel is the share "
                             "ray-cast, which is bound by VALU issue and vector-load address processing, "
                             "not HBM: its fractions are in render_kernel."},
        "render_kernel": {"kernel": "share_kernel (all 512 blocks, per-block 8-bit composite, fast sample "
                                    "sequence, micro-brick bricks)",
                          "bound": "valu / l1 tag lookups", "ms": t_share * 1e3,
                          "gsamples_s": S / t_share / 1e9,
                          "brick_bytes_per_frame": 4 * V_total,
                          "brick_gb_s": 4 * V_total / t_share / 1e9},
    })
    del batch

    # ---------------- config 1: the refined tree's level-2 depth-6 brick
    t0 = time.time()
    cfg1 = configs.cpu_interp_brick(SEED)
    n2i1, d1 = cfg1["n2i"], cfg1["depth"]
    sp1 = torch.from_numpy(cfg1["sparse"]).cuda()
    plan1 = api.interp_plan(n2i1, d1)
    dense1 = torch.empty(api.brick_nodes(d1), dtype=torch.float32, device="cuda")
    api.interp_brick_planned(sp1, plan1, d1, dense1)
    t_i1 = events_time(torch, lambda: api.interp_brick_planned(sp1, plan1, d1, dense1), 50)
    V1 = api.brick_nodes(d1)
    log("config 1: %d octree leaves, level-2 root %s, depth %d, %d of %d nodes given; GPU %.1f us, "
        "%.1f Gvox/s (setup %.1fs)" % (len(cfg1["tree"].leaves), cfg1["index"], d1, n2i1.shape[0], V1,
                                      t_i1 * 1e6, V1 / t_i1 / 1e9, time.time() - t0))
    out["cpu_interp_config"] = {
        "workload": "3 Gaussians, octree refined to level 4 where |grad f| > 0.5 max (%d leaf blocks of "
                    "16^3 cells), the level-2 root %s: depth-6 subtree, %d of 65^3 nodes given"
                    % (len(cfg1["tree"].leaves), cfg1["index"], n2i1.shape[0]),
        "gpu_interp_us": t_i1 * 1e6, "gpu_gvox_s": V1 / t_i1 / 1e9}

    # ---------------- peaks measured on this chip, and the render kernel against them
    peaks = api.probe_peaks(torch)
    log("peaks: VALU %.1f T lane-fma/s, LDS %.1f TB/s, L1 gather %.2f G loads/s coalesced, %.2f G scattered, "
        "HBM stream %.0f GB/s" % (peaks["valu_fma_s"] / 1e12, peaks["lds_bytes_s"] / 1e12,
                                  peaks["gather_coalesced_s"] / 1e9, peaks["gather_scattered_s"] / 1e9,
                                  peaks["stream_bytes_s"] / 1e9))
    prof = "share_kernel_512x65cubed_1080p_counters.md"
    valu_insts = profiled(prof, "fast_sq_insts_valu")
    tcp_acc = profiled(prof, "fast_tcp_total_cache_accesses")
    ta_busy = profiled(prof, "fast_ta_busy_frac")
    rk = out["render_kernel"]
    rk["peaks_measured"] = peaks
    if valu_insts:
        # wave instructions per second against the probe's (a lane fma is one
        # 64th of a wave fma instruction); with the samples per frame, the
        # VALU lane-slots per sample
        rk["valu"] = {"achieved": valu_insts / t_share, "peak": peaks["valu_fma_s"] / 64, "unit": "wave insts/s",
                      "frac": valu_insts / t_share / (peaks["valu_fma_s"] / 64),
                      "lane_slots_per_sample": 64.0 * valu_insts / S,
                      "source": "SQ_INSTS_VALU per frame from profiles/" + prof}
    if ta_busy:
        # the texture addresser (one per CU) processes every vector load's
        # addresses and L1 tag lookups: its busy fraction over the kernel
        rk["texture_addresser"] = {"frac": ta_busy, "l1_accesses_per_frame": tcp_acc,
                                   "source": "TA_TA_BUSY_sum / (GRBM_GUI_ACTIVE / 8 XCDs x 256 CUs) from "
                                             "profiles/" + prof}
    if valu_insts and ta_busy:
        rk["frac"] = max(rk["valu"]["frac"], ta_busy)
        rk["bound"] = "valu issue" if rk["valu"]["frac"] >= ta_busy else "texture addresser (L1 tag lookups)"

    # ---------------- config 3: 2000 mixed-level blocks, batched vs plain
    if not args.no_mixed:
        out["mixed2000"] = mixed_bench(torch, args, tf, log)

    # ---------------- config 4: forward + adjoint at 512x512 on the same field
    if not args.no_adjoint:
        W4 = H4 = 512
        rng4 = np.random.default_rng(SEED + 4)
        tf_t = torch.from_numpy(np.clip(tf_np + rng4.normal(0.0, 0.05, size=tf_np.shape), 0.0, 1.0)
                                .astype(np.float32)).cuda()
        # the adjoint differentiates the exact sample sequence (frames with the
        # tape run it), so the forward it is compared with is the exact one
        fp4 = api.FramePlan(tblocks, W4, H4, FOV, wl.fmin, wl.fmax, THR, Q2, CUT2, ACC_CUT, tf, (0, 0),
                            api.BRICK_TILED4, exact=True)
        fp4t = api.FramePlan(tblocks, W4, H4, FOV, wl.fmin, wl.fmax, THR, Q2, CUT2, ACC_CUT, tf_t, (0, 0),
                             api.BRICK_TILED4)
        fp4.set_cameras(M[None])
        fp4t.set_cameras(M[None])
        argb4 = torch.empty((H4, W4), dtype=torch.int32, device="cuda")
        fp4.autotune(argb4, False, TUNE_CACHE)
        rgba4 = torch.empty((H4, W4, 4), dtype=torch.float32, device="cuda")
        target4 = torch.empty_like(rgba4)
        fp4t.render(0, argb4, target4, interpolate=False)
        gb4 = torch.zeros_like(wl.bricks)
        gtf4 = torch.zeros((tf_np.shape[0], 4), dtype=torch.float32, device="cuda")
        fp4.set_grad_bricks([gb4[i] for i in range(gb4.shape[0])])

        def fwd4():
            fp4.render(0, argb4, rgba4, interpolate=False)

        def adj4():
            # dL/dimage of L = 1/2 |image - target|^2, gradients zeroed, adjoint
            gI = rgba4 - target4
            gtf4.zero_()
            gb4.zero_()
            fp4.adjoint(0, rgba4, gI, gtf4)

        t_fwd4_plain = events_time(torch, fwd4, 10)
        fp4.set_exact(False)
        fwd4()
        t_fwd4_fast = events_time(torch, fwd4, 10)
        fp4.set_exact(True)
        fp4.set_tape(True)  # the forward of a training step records the adjoint's segment states
        fwd4()
        adj4()
        t_fwd4 = events_time(torch, fwd4, 10)
        t_fa4 = events_time(torch, lambda: (fwd4(), adj4()), 5)
        t_adj4 = events_time(torch, adj4, 5)
        fp4.set_grad_bricks(None)

        def adj4_tf():
            gI = rgba4 - target4
            gtf4.zero_()
            fp4.adjoint(0, rgba4, gI, gtf4)

        t_fa4_tf = events_time(torch, lambda: (fwd4(), adj4_tf()), 5)
        log("config 4: forward %.3f ms (with tape %.3f), forward+adjoint %.3f ms (ratio %.2f), adjoint "
            "%.3f ms, colour-table-only forward+adjoint %.3f ms" % (t_fwd4_plain * 1e3, t_fwd4 * 1e3,
                                                                    t_fa4 * 1e3, t_fa4 / t_fwd4_plain,
                                                                    t_adj4 * 1e3, t_fa4_tf * 1e3))
        out["adjoint512"] = {
            "workload": "512 level-3 blocks x 65^3 fp32 (resident), 512x512, L = 1/2 |image - target|^2 "
                        "with the target rendered from the colour table + seeded N(0,0.05); gradients "
                        "w.r.t. the 256x4 table and all 140.6M field values (zeroed every step)",
            "forward_ms": t_fwd4_plain * 1e3, "forward_fast_sequence_ms": t_fwd4_fast * 1e3,
            "forward_with_tape_ms": t_fwd4 * 1e3,
            "fwd_adjoint_ms": t_fa4 * 1e3, "adjoint_ms": t_adj4 * 1e3,
            "ratio_fwd_adjoint_over_fwd": t_fa4 / t_fwd4_plain,
            "tf_only_fwd_adjoint_ms": t_fa4_tf * 1e3}
        fp4.close()
        fp4t.close()
        del gb4

    if not args.no_large:
        out["large_share"] = large_share_bench(torch, args, tf_np, log)

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl, tf_np, cfg1, args.cpu_runs)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
