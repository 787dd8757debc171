"""numpy/torch-facing wrappers of the C ABI (host functions take numpy
arrays, device functions take torch CUDA tensors)."""
import ctypes as C

import numpy as np

from . import (RENDER_EXACT, BlockParams, BrickJob, DatasetInfo, FieldInfo, FrameBlock, FrameDesc,
               FrameOutputHeader, RenderState, RvError, SubtreeInfo, check, lib, ptr, stream_ptr)


def _np(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


# ---- host ---------------------------------------------------------------

def subtree_depth(children, max_depth):
    ch = _np(children, np.uint32)
    out = C.c_uint32()
    check(lib().rv_subtree_depth(_p(ch, C.c_uint32), ch.shape[0], max_depth, C.byref(out)))
    return out.value


def subtree_n2i(children, depth):
    ch = _np(children, np.uint32)
    cap = (2 ** depth + 1) ** 3
    out = np.zeros((cap, 3), dtype=np.uint16)
    cnt = C.c_uint64()
    check(lib().rv_subtree_n2i(_p(ch, C.c_uint32), ch.shape[0], depth, _p(out, C.c_uint16), cap,
                               C.byref(cnt)))
    return out[: cnt.value].copy()


def back_to_front(imin, imax, cmin, cmax, eye):
    imin, imax = _np(imin, np.uint32), _np(imax, np.uint32)
    cmin, cmax, eye = _np(cmin, np.float32), _np(cmax, np.float32), _np(eye, np.float32)
    cap = int(np.prod(imax.astype(np.int64) - imin.astype(np.int64)))
    out = np.zeros((max(cap, 1), 3), dtype=np.uint32)
    cnt = C.c_uint64()
    check(lib().rv_back_to_front(_p(imin, C.c_uint32), _p(imax, C.c_uint32), _p(cmin, C.c_float),
                                 _p(cmax, C.c_float), _p(eye, C.c_float), _p(out, C.c_uint32), cap,
                                 C.byref(cnt)))
    return out[: cnt.value].copy()


def eye_from_view(view16):
    v = _np(view16, np.float32).reshape(16)
    e = np.zeros(3, dtype=np.float32)
    check(lib().rv_eye_from_view(_p(v, C.c_float), _p(e, C.c_float)))
    return e


def index_box_coords(level, bmin, bmax):
    bmin, bmax = _np(bmin, np.uint32), _np(bmax, np.uint32)
    cmin, cmax = np.zeros(3, np.float32), np.zeros(3, np.float32)
    check(lib().rv_index_box_coords(level, _p(bmin, C.c_uint32), _p(bmax, C.c_uint32),
                                    _p(cmin, C.c_float), _p(cmax, C.c_float)))
    return cmin, cmax


def index_cube_split_sorted(level, count, eye):
    eye = _np(eye, np.float32)
    cap = count + 8
    out = np.zeros((cap, 6), dtype=np.uint32)
    cnt = C.c_uint64()
    check(lib().rv_index_cube_split_sorted(level, count, _p(eye, C.c_float), _p(out, C.c_uint32), cap,
                                           C.byref(cnt)))
    return out[: cnt.value].copy()


def mixed_back_to_front(blocks, eye):
    b = _np(blocks, np.uint32).reshape(-1, 4)
    eye = _np(eye, np.float32)
    out = np.zeros(max(b.shape[0], 1), dtype=np.uint32)
    check(lib().rv_mixed_back_to_front(_p(b, C.c_uint32), b.shape[0], _p(eye, C.c_float),
                                       _p(out, C.c_uint32)))
    return out[: b.shape[0]].copy()


def block_params(view16, vp_w, vp_h, fov, level, index, tex_depth, fmin, fmax, thr_rel, quality,
                 alpha_cutoff, exact=False):
    """rv_block_params_setup; exact: flag the exact Argb sample sequence
    (RV_RENDER_EXACT), else the default fast one."""
    v = _np(view16, np.float32).reshape(16)
    idx = _np(index, np.uint32)
    p = BlockParams()
    live = C.c_int()
    check(lib().rv_block_params_setup(_p(v, C.c_float), vp_w, vp_h, fov, level, _p(idx, C.c_uint32),
                                      tex_depth, fmin, fmax, thr_rel, quality, alpha_cutoff,
                                      C.byref(p), C.byref(live)))
    if exact:
        p.flags |= RENDER_EXACT
    return live.value, p


def color_transfer_vector(xs, colors, count):
    xs, colors = _np(xs, np.float32), _np(colors, np.float32)
    out = np.zeros((count, 4), dtype=np.float32)
    check(lib().rv_color_transfer_vector(_p(xs, C.c_float), _p(colors, C.c_float), xs.size, count,
                                         _p(out, C.c_float)))
    return out


def background_argb(r, g, b):
    return int(lib().rv_background_argb(r, g, b))


# ---- device (torch CUDA tensors) ------------------------------------------

def brick_nodes(depth):
    return int(lib().rv_brick_nodes(depth))


def node_ordinals(d_n2i, depth, d_out, stream=None):
    check(lib().rv_node_ordinals(ptr(d_n2i), d_n2i.shape[0], depth, ptr(d_out), stream_ptr(stream)))


def interp_plan_bytes(n2i, depth):
    n2i = _np(n2i, np.uint16)
    return int(lib().rv_interp_plan_bytes(_p(n2i, C.c_uint16), n2i.shape[0], depth))


def interp_plan_encode(n2i, depth):
    """Host-encoded gather plan (numpy uint8)."""
    n2i = _np(n2i, np.uint16)
    nb = interp_plan_bytes(n2i, depth)
    if nb == 0:
        check(1)
    out = np.zeros(nb, dtype=np.uint8)
    check(lib().rv_interp_plan_encode(_p(n2i, C.c_uint16), n2i.shape[0], depth,
                                      out.ctypes.data_as(C.c_void_p), nb))
    return out


def interp_plan(n2i, depth):
    """Gather plan of a subtree as a torch uint8 device tensor."""
    import torch
    return torch.from_numpy(interp_plan_encode(n2i, depth)).cuda()


def interp_brick_planned(d_sparse, d_plan, depth, d_dense, stream=None):
    check(lib().rv_interp_brick_planned(ptr(d_sparse), ptr(d_plan), depth, ptr(d_dense),
                                        stream_ptr(stream)))


def interp_brick(d_sparse, n2i, depth, d_dense, stream=None):
    n2i = _np(n2i, np.uint16)
    check(lib().rv_interp_brick(ptr(d_sparse), _p(n2i, C.c_uint16), n2i.shape[0], depth, ptr(d_dense),
                                stream_ptr(stream)))


# numpy layout of rv_brick_job (for batches of millions of bricks)
BRICK_JOB_DTYPE = np.dtype({"names": ["sparse", "plan", "dense", "dense16", "depth", "layout"],
                            "formats": ["<u8", "<u8", "<u8", "<u8", "<u4", "<u4"],
                            "offsets": [0, 8, 16, 24, 32, 36], "itemsize": 40})
assert BRICK_JOB_DTYPE.itemsize == C.sizeof(BrickJob)


BRICK_LINEAR, BRICK_TILED4 = 0, 1


def brick_floats(depth, layout=BRICK_LINEAR):
    """Floats of a brick of this depth in a layout (BRICK_TILED4: 4^3 micro-bricks)."""
    return int(lib().rv_brick_floats(depth, layout))


class InterpBatch:
    """Batched interpolation handle; jobs: list of (sparse, plan, dense, depth)
    torch tensors (dense may be float32 or float16), or a numpy array of
    BRICK_JOB_DTYPE holding device addresses (the caller keeps the memory
    alive); layout: BRICK_LINEAR or BRICK_TILED4 (fp32) for the list form."""

    def __init__(self, jobs, layout=BRICK_LINEAR):
        import torch
        self._keep = []
        if isinstance(jobs, np.ndarray):
            assert jobs.dtype == BRICK_JOB_DTYPE and jobs.flags.c_contiguous
            self._keep.append(jobs)
            h = C.c_void_p()
            check(lib().rv_interp_batch_create(C.cast(jobs.ctypes.data, C.POINTER(BrickJob)), len(jobs),
                                               C.byref(h)))
            self.h = h
            return
        arr = (BrickJob * max(len(jobs), 1))()
        for i, (sp, pl, de, depth) in enumerate(jobs):
            arr[i].sparse = sp.data_ptr()
            arr[i].plan = pl.data_ptr()
            if de.dtype == torch.float16:
                arr[i].dense16 = de.data_ptr()
            else:
                arr[i].dense = de.data_ptr()
            arr[i].depth = depth
            arr[i].layout = layout
            self._keep.append((sp, pl, de))
        h = C.c_void_p()
        check(lib().rv_interp_batch_create(arr, len(jobs), C.byref(h)))
        self.h = h

    def run(self, stream=None):
        check(lib().rv_interp_batch_run(self.h, stream_ptr(stream)))

    def info(self):
        t, n = C.c_uint64(), C.c_uint64()
        check(lib().rv_interp_batch_info(self.h, C.byref(t), C.byref(n)))
        return t.value, n.value

    def close(self):
        if self.h:
            lib().rv_interp_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def clear_viewport(d_vp, w, h, stream=None):
    check(lib().rv_clear_viewport(ptr(d_vp), w, h, stream_ptr(stream)))


def render_block_argb(params, d_brick, d_tf, d_vp, stream=None):
    check(lib().rv_render_block_argb(C.byref(params), ptr(d_brick), ptr(d_tf), d_tf.shape[0],
                                     ptr(d_vp), stream_ptr(stream)))


def subtree_n2i_device(d_children, nnodes, depth, d_out, stream=None):
    """n2i built on the device into d_out (cap x 3 uint16 tensor); returns
    the node count."""
    cnt = C.c_uint64()
    check(lib().rv_subtree_n2i_device(ptr(d_children), nnodes, depth, ptr(d_out), d_out.shape[0],
                                      C.byref(cnt), stream_ptr(stream)))
    return cnt.value


# s3vs::FieldMode (FieldMode.hpp:26-51)
FIELD_ISOSURFACE, FIELD_ISOSURFACES, FIELD_MIP, FIELD_ARGB, FIELD_ARGB_LIGHT, FIELD_DOMAIN_VOXELS = 0, 1, 2, 3, 4, 5
FIELD_VALUE_ON_ISOSURFACE, FIELD_VALUE_ON_ISOSURFACES = 6, 7
SCHEDULE_BINS, SCHEDULE_LANES, SCHEDULE_HYBRID = 0, 1, 2


def _levels(levels):
    lv = _np(levels if levels is not None else [], np.float32).reshape(-1)
    return lv, (_p(lv, C.c_float) if lv.size else None), lv.size


def render_block(params, mode, iso_opacity, d_brick, d_tf, d_vp, levels=None, stream=None, d_brick2=None,
                 srange=None):
    """rv_render_block; with d_brick2 (and srange = the secondary field's
    range) rv_render_block_secondary, for the ValueOnIsosurface(s) modes."""
    lv, lp, ln = _levels(levels)
    if d_brick2 is not None:
        check(lib().rv_render_block_secondary(C.byref(params), mode, iso_opacity, lp, ln, ptr(d_brick),
                                              ptr(d_brick2), srange[0], srange[1], ptr(d_tf), d_tf.shape[0],
                                              ptr(d_vp), stream_ptr(stream)))
        return
    check(lib().rv_render_block(C.byref(params), mode, iso_opacity, lp, ln, ptr(d_brick), ptr(d_tf),
                                d_tf.shape[0], ptr(d_vp), stream_ptr(stream)))


def clear_viewport_mode(d_vp, w, h, mode, stream=None):
    check(lib().rv_clear_viewport_mode(ptr(d_vp), w, h, mode, stream_ptr(stream)))


def blend_image_part_max_intensity(d_dst, dw, dh, dst_off, d_src, sw, sh, src_xy, just_copy=False,
                                   stream=None):
    check(lib().rv_blend_image_part_max_intensity(ptr(d_dst), dw, dh, dst_off[0], dst_off[1], ptr(d_src), sw,
                                                  sh, src_xy[0], src_xy[1], int(just_copy),
                                                  stream_ptr(stream)))


def mip_colorize(d_dst, dw, dh, d_src, sw, sh, src_xy, d_ctv, stream=None):
    check(lib().rv_mip_colorize(ptr(d_dst), dw, dh, ptr(d_src), sw, sh, src_xy[0], src_xy[1], ptr(d_ctv),
                                d_ctv.shape[0], stream_ptr(stream)))


def block_sample_count(params, d_count, stream=None):
    check(lib().rv_block_sample_count(C.byref(params), ptr(d_count), stream_ptr(stream)))


class ShareRenderer:
    """rv_render_share_argb over a fixed list of (params, brick) in
    back-to-front order."""

    def __init__(self, items, d_tf, W, H, bricks2=None, layout=BRICK_LINEAR):
        import torch
        live = [(p, b) for p, b in items]
        self.n = len(live)
        self.params = (BlockParams * max(self.n, 1))(*[p for p, _ in live])
        self.bricks = (C.c_void_p * max(self.n, 1))(*[b.data_ptr() for _, b in live])
        self.bricks2 = None
        if bricks2 is not None:
            assert len(bricks2) == self.n
            self.bricks2 = (C.c_void_p * max(self.n, 1))(*[b.data_ptr() for b in bricks2])
        self._keep = [b for _, b in live] + [d_tf] + list(bricks2 or [])
        self.layout = layout
        self.tf, self.W, self.H = d_tf, W, H
        nbytes = int(lib().rv_share_workspace_bytes(self.n))
        self.ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")

    def render_mode(self, d_vp, mode, iso_opacity=1.0, levels=None, stream=None, srange=None):
        lv, lp, ln = _levels(levels)
        if self.layout != BRICK_LINEAR:
            sr = srange or (0.0, 1.0)
            check(lib().rv_render_share_layout(self.params, self.bricks, self.bricks2, self.n, self.layout, mode,
                                               iso_opacity, lp, ln, sr[0], sr[1], ptr(self.tf), self.tf.shape[0],
                                               ptr(d_vp), self.W, self.H, ptr(self.ws), stream_ptr(stream)))
            return
        if self.bricks2 is not None:
            check(lib().rv_render_share_secondary(self.params, self.bricks, self.bricks2, self.n, mode,
                                                  iso_opacity, lp, ln, srange[0], srange[1], ptr(self.tf),
                                                  self.tf.shape[0], ptr(d_vp), self.W, self.H, ptr(self.ws),
                                                  stream_ptr(stream)))
            return
        check(lib().rv_render_share(self.params, self.bricks, self.n, mode, iso_opacity, lp, ln, ptr(self.tf),
                                    self.tf.shape[0], ptr(d_vp), self.W, self.H, ptr(self.ws),
                                    stream_ptr(stream)))

    def render(self, d_vp, stream=None):
        if self.layout != BRICK_LINEAR:
            self.render_mode(d_vp, FIELD_ARGB, stream=stream)
            return
        check(lib().rv_render_share_argb(self.params, self.bricks, self.n, ptr(self.tf), self.tf.shape[0],
                                         ptr(d_vp), self.W, self.H, ptr(self.ws), stream_ptr(stream)))


def render_share_argb_noalloc(params, bricks, d_tf, d_vp, W, H, stream=None):
    """rv_render_share_argb_noalloc over (params list, device brick tensors),
    back to front: the block table goes to the library's fixed array."""
    n = len(params)
    pa = (BlockParams * max(n, 1))(*params)
    ba = (C.c_void_p * max(n, 1))(*[b.data_ptr() for b in bricks])
    check(lib().rv_render_share_argb_noalloc(pa, ba, n, ptr(d_tf), d_tf.shape[0], ptr(d_vp), W, H,
                                             stream_ptr(stream)))


def bound_rect(d_pix, w, h, stream=None):
    r = (C.c_int32 * 4)()
    ne = C.c_int()
    check(lib().rv_bound_rect(ptr(d_pix), w, h, r, C.byref(ne), stream_ptr(stream)))
    return tuple(r) if ne.value else None


def blend_image_part(d_dst, dw, dh, dst_off, d_src, sw, sh, src_xy, just_copy=False, stream=None):
    check(lib().rv_blend_image_part(ptr(d_dst), dw, dh, dst_off[0], dst_off[1], ptr(d_src), sw, sh,
                                    src_xy[0], src_xy[1], int(just_copy), stream_ptr(stream)))


def fill_u32(d, n, value, stream=None):
    check(lib().rv_fill_u32(ptr(d), n, value, stream_ptr(stream)))


# numpy layout of rv_frame_block (for shares of millions of blocks)
FRAME_BLOCK_DTYPE = np.dtype({"names": ["level", "index", "depth", "sparse", "plan", "brick", "brick16"],
                              "formats": ["<u4", ("<u4", 3), "<u4", "<u8", "<u8", "<u8", "<u8"],
                              "offsets": [0, 4, 16, 24, 32, 40, 48], "itemsize": 56})
assert FRAME_BLOCK_DTYPE.itemsize == C.sizeof(FrameBlock)


class FramePlan:
    """Batched persistent frame (rv_frame_plan).  blocks: list of dicts with
    level, index, depth, brick (torch fp32/fp16 tensor), optional sparse and
    plan tensors; or a numpy array of FRAME_BLOCK_DTYPE holding device
    addresses (the caller keeps the memory alive)."""

    def __init__(self, blocks, vp_w, vp_h, fov, fmin, fmax, thr_rel, quality, alpha_cutoff,
                 acc_cutoff, d_tf, tile=(0, 0), layout=BRICK_LINEAR, exact=False):
        import torch
        self._keep = [d_tf]
        if isinstance(blocks, np.ndarray):
            assert blocks.dtype == FRAME_BLOCK_DTYPE and blocks.flags.c_contiguous
            arr = C.cast(blocks.ctypes.data, C.POINTER(FrameBlock))
            self._keep.append(blocks)
        else:
            arr = (FrameBlock * len(blocks))()
            for i, b in enumerate(blocks):
                arr[i].level = b["level"]
                for k in range(3):
                    arr[i].index[k] = b["index"][k]
                arr[i].depth = b["depth"]
                if b.get("sparse") is not None:
                    arr[i].sparse = b["sparse"].data_ptr()
                    arr[i].plan = b["plan"].data_ptr()
                if b["brick"].dtype == torch.float16:
                    arr[i].brick16 = b["brick"].data_ptr()
                else:
                    arr[i].brick = b["brick"].data_ptr()
                self._keep.append(b)
        d = FrameDesc(vp_w, vp_h, fov, fmin, fmax, thr_rel, quality, alpha_cutoff, acc_cutoff,
                      d_tf.data_ptr(), d_tf.shape[0], tile[0], tile[1], layout, 1 if exact else 0)
        h = C.c_void_p()
        check(lib().rv_frame_plan_create(arr, len(blocks), C.byref(d), C.byref(h)))
        self.h = h
        self.W, self.H = vp_w, vp_h

    def set_cameras(self, views):
        v = _np(views, np.float32).reshape(-1, 16)
        check(lib().rv_frame_plan_set_cameras(self.h, _p(v, C.c_float), v.shape[0]))

    def render(self, cam, d_argb, d_rgba=None, interpolate=True, stream=None):
        check(lib().rv_frame_render(self.h, cam, ptr(d_argb), ptr(d_rgba) if d_rgba is not None else None,
                                    int(interpolate), stream_ptr(stream)))

    def graph_build(self, d_argb, frame_stride, interpolate=True, stream=None):
        check(lib().rv_frame_graph_build(self.h, ptr(d_argb), frame_stride, int(interpolate),
                                         stream_ptr(stream)))

    def graph_launch(self, stream=None):
        check(lib().rv_frame_graph_launch(self.h, stream_ptr(stream)))

    def status(self):
        v = C.c_uint32()
        check(lib().rv_frame_status(self.h, C.byref(v)))
        return v.value

    def tile(self):
        w, h = C.c_uint32(), C.c_uint32()
        check(lib().rv_frame_get_tile(self.h, C.byref(w), C.byref(h)))
        return w.value, h.value

    def set_schedule(self, schedule):
        """SCHEDULE_BINS (the wave marches each block of its bin) or
        SCHEDULE_LANES (each lane walks its own pixel's blocks; Argb)."""
        check(lib().rv_frame_set_schedule(self.h, schedule))

    def set_hybrid_min(self, blocks):
        """SCHEDULE_HYBRID: per-lane walk for bins listing at least `blocks` blocks."""
        check(lib().rv_frame_set_hybrid_min(self.h, blocks))

    def set_grouped(self, grouped):
        """Resident-brick frames: one workgroup per 2 x 2 neighbouring bins
        (True) or the persistent grid with its bin queue (False, the
        default); the autotuner picks the faster."""
        check(lib().rv_frame_set_grouped(self.h, int(bool(grouped))))

    def schedule(self):
        v = C.c_uint32()
        check(lib().rv_frame_get_schedule(self.h, C.byref(v)))
        return v.value

    def set_exact(self, exact):
        check(lib().rv_frame_set_exact(self.h, 1 if exact else 0))

    def set_split(self, split):
        """Interpolating frames in two launches (bricks, then the resident
        kernel) instead of one; the same images."""
        check(lib().rv_frame_set_split(self.h, 1 if split else 0))

    def split(self):
        v = C.c_int()
        check(lib().rv_frame_get_split(self.h, C.byref(v)))
        return bool(v.value)

    def set_field_mode(self, mode, iso_opacity=1.0, levels=None):
        lv, lp, ln = _levels(levels)
        check(lib().rv_frame_set_field_mode(self.h, mode, iso_opacity, lp, ln))

    def set_secondary(self, sparse2, bricks2, srange):
        """The secondary field of the ValueOnIsosurface(s) modes: per block a
        brick (torch tensor, the primary bricks' type) and sparse values
        (torch tensors, or None for resident frames only); bricks2 None
        removes it."""
        if bricks2 is None:
            check(lib().rv_frame_set_secondary(self.h, None, None, 0.0, 1.0))
            self._sec = None
            return
        b2 = (C.c_void_p * len(bricks2))(*[b.data_ptr() for b in bricks2])
        s2 = None
        if sparse2 is not None:
            s2 = (C.c_void_p * len(sparse2))(*[v.data_ptr() for v in sparse2])
        check(lib().rv_frame_set_secondary(self.h, s2, b2, srange[0], srange[1]))
        self._sec = (list(bricks2), list(sparse2 or []))

    def set_tape(self, enable=True):
        """Record the adjoint's per-segment states in later float-image renders."""
        check(lib().rv_frame_set_tape(self.h, int(enable)))

    def set_grad_bricks(self, grads):
        """Per-block fp32 gradient arrays (torch tensors, brick layout), or None."""
        if grads is None:
            check(lib().rv_frame_set_grad_bricks(self.h, None))
            self._grads = None
            return
        arr = (C.c_void_p * len(grads))(*[g.data_ptr() for g in grads])
        check(lib().rv_frame_set_grad_bricks(self.h, arr))
        self._grads = list(grads)

    def adjoint(self, cam, d_rgba, d_grad_rgba, d_grad_tf, stream=None):
        check(lib().rv_frame_adjoint(self.h, cam, ptr(d_rgba), ptr(d_grad_rgba), ptr(d_grad_tf),
                                     stream_ptr(stream)))

    def autotune(self, d_argb, interpolate=True, cache_path=None, stream=None):
        w, h = C.c_uint32(), C.c_uint32()
        check(lib().rv_frame_autotune(self.h, ptr(d_argb), int(interpolate),
                                      cache_path.encode() if cache_path else None, stream_ptr(stream),
                                      C.byref(w), C.byref(h)))
        return w.value, h.value

    def close(self):
        if self.h:
            lib().rv_frame_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Worker:
    """rv_worker: VsWorkerInterface drop-in (field service + render state +
    renderScenePart + on-device share compositing)."""

    def __init__(self, thread_id=0):
        h = C.c_void_p()
        check(lib().rv_worker_create(thread_id, C.byref(h)))
        self.h = h
        self.W = self.H = 0

    def add_subtree(self, level, index, depth, n2i, sparse):
        idx = _np(index, np.uint32)
        n2i = _np(n2i, np.uint16)
        sp = _np(sparse, np.float32)
        check(lib().rv_worker_add_subtree(self.h, level, _p(idx, C.c_uint32), depth, _p(n2i, C.c_uint16),
                                          n2i.shape[0], _p(sp, C.c_float)))

    def add_subtree_secondary(self, level, index, sparse):
        """Secondary field values of a registered subtree (ValueOnIsosurface(s))."""
        idx = _np(index, np.uint32)
        sp = _np(sparse, np.float32)
        check(lib().rv_worker_add_subtree_secondary(self.h, level, _p(idx, C.c_uint32), _p(sp, C.c_float)))

    def set_state(self, view16, vp_w, vp_h, fov, quality, thr_rel, alpha_cutoff, fmin, fmax, background,
                  tf, field_mode=FIELD_ARGB, iso_level_rel=0.5, iso_opacity=1.0, iso_levels=None,
                  secondary_range=(0.0, 1.0), exact=False):
        self._view = _np(view16, np.float32).reshape(16)
        self._tf = _np(tf, np.float32)
        self._levels = _np(iso_levels if iso_levels is not None else [], np.float32).reshape(-1)
        st = RenderState(self._view.ctypes.data, vp_w, vp_h, fov, quality, thr_rel, alpha_cutoff, fmin,
                         fmax, (C.c_float * 3)(*background), self._tf.ctypes.data, self._tf.shape[0],
                         field_mode, iso_level_rel, iso_opacity,
                         self._levels.ctypes.data if self._levels.size else None, self._levels.size,
                         secondary_range[0], secondary_range[1], 1 if exact else 0)
        check(lib().rv_worker_set_state(self.h, C.byref(st)))
        self.W, self.H = vp_w, vp_h

    def render_scene_part(self, level, box_min, box_max, cancel=False, out=None):
        """rv_worker_render_scene_part into a new numpy buffer, or into `out`
        (a contiguous 32-bit torch tensor or numpy array, e.g. pinned host
        memory, which the library writes from the GPU directly); returns
        ((x, y, w, h), the h x w part) or None when cancelled."""
        bmin, bmax = _np(box_min, np.uint32), _np(box_max, np.uint32)
        if out is None:
            cap = self.W * self.H
            bits = np.zeros(max(cap, 1), dtype=np.uint32)
            ptr_out = _p(bits, C.c_uint32)
        else:
            cap = out.numel() if hasattr(out, "numel") else out.size
            bits = out.numpy().view(np.uint32) if hasattr(out, "numpy") else out.view(np.uint32)
            ptr_out = C.cast(out.data_ptr() if hasattr(out, "data_ptr") else out.ctypes.data, C.POINTER(C.c_uint32))
        rect = (C.c_int32 * 4)()
        flag = C.c_int(1 if cancel else 0)
        rc = lib().rv_worker_render_scene_part(self.h, level, _p(bmin, C.c_uint32), _p(bmax, C.c_uint32),
                                               C.byref(flag), rect, ptr_out, cap)
        if rc == 4:
            return None
        check(rc)
        x, y, w, h = tuple(rect)
        return (x, y, w, h), bits[: w * h].reshape(h, w).copy()

    def set_fused(self, fused):
        check(lib().rv_worker_set_fused(self.h, int(fused)))

    def render_progressive(self, level_min, level_max, split_count, out=None, frame_number=0,
                           measure_time=False, cancel=False):
        """rv_worker_render_progressive into the frame output block `out`
        (a uint8 array of FRAME_OUTPUT_HEADER_BYTES + 4*W*H bytes, made if
        None).  Returns (frames written, next frame number, out)."""
        if out is None:
            out = np.zeros(FRAME_OUTPUT_HEADER_BYTES + 4 * self.W * self.H, dtype=np.uint8)
        fn = C.c_uint64(frame_number)
        nw = C.c_uint32()
        flag = C.c_int(1 if cancel else 0)
        check(lib().rv_worker_render_progressive(self.h, level_min, level_max, split_count, C.byref(flag),
                                                 out.ctypes.data, out.nbytes, C.byref(fn), int(measure_time),
                                                 C.byref(nw)))
        return nw.value, fn.value, out

    def compose_frame(self, level, split_count):
        out = np.zeros((self.H, self.W), dtype=np.uint32)
        flag = C.c_int(0)
        check(lib().rv_worker_compose_frame(self.h, level, split_count, _p(out, C.c_uint32), C.byref(flag)))
        return out

    def close(self):
        if self.h:
            lib().rv_worker_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- frame output block -------------------------------------------------

FRAME_OUTPUT_HEADER_BYTES = 24


def read_frame_output(buf, vp_w, vp_h):
    """(header dict, vp_h x vp_w ARGB pixels) of a frame output block."""
    h = FrameOutputHeader()
    pix = np.zeros((vp_h, vp_w), dtype=np.uint32)
    check(lib().rv_frame_output_read(buf.ctypes.data, buf.nbytes, vp_w, vp_h, C.byref(h),
                                     _p(pix, C.c_uint32)))
    return dict(frame_number=h.frame_number, duration_ms=h.duration_ms, level=h.level), pix


# ---- dataset files ------------------------------------------------------

def dataset_file_names(base, field=None, frame=0, has_time_steps=False):
    """(meta, field file, field info) paths of a dataset: <base>.s3dmm-meta,
    <base>.s3dmm-field#<field> and <base>.s3dmm-fields; with time steps the
    frame's base is <dir>/<stem>/<stem>_tNNNNNN<ext> and the info file
    <dir>/<stem>/<stem><ext>.s3dmm-fields (filename_util.hpp:36-48,
    BlockTreeFieldService.hpp:125-131, 177-191)."""
    import os
    if has_time_steps:
        d, name = os.path.split(base)
        stem, ext = os.path.splitext(name)
        fbase = os.path.join(d, stem, "%s_t%06d%s" % (stem, frame, ext))
        info = os.path.join(d, stem, stem + ext + ".s3dmm-fields")
    else:
        fbase, info = base, base + ".s3dmm-fields"
    return (fbase + ".s3dmm-meta", None if field is None else fbase + ".s3dmm-field#" + field, info)


def read_fields_info(path):
    """(frame count, [(name, min, max, animated)]) of a .s3dmm-fields file."""
    steps, count = C.c_uint32(), C.c_uint32()
    rc = lib().rv_fields_info_read(path.encode(), C.byref(steps), None, 0, C.byref(count))
    if count.value == 0:
        check(rc)
    arr = (FieldInfo * max(count.value, 1))()
    check(lib().rv_fields_info_read(path.encode(), C.byref(steps), arr, count.value, C.byref(count)))
    return steps.value, [(f.name.decode(), f.min, f.max, bool(f.animated)) for f in arr[:count.value]]


class Dataset:
    """rv_dataset: a .s3dmm-meta file (block tree + subtree roots)."""

    def __init__(self, meta_path):
        h = C.c_void_p()
        check(lib().rv_dataset_open(meta_path.encode(), C.byref(h)))
        self.h = h
        info = DatasetInfo()
        check(lib().rv_dataset_get_info(self.h, C.byref(info)))
        self.info = {f: (tuple(getattr(info, f)) if f == "bc_min" else getattr(info, f))
                     for f, _ in DatasetInfo._fields_}

    def __len__(self):
        return self.info["subtree_count"]

    def subtree(self, i):
        s = SubtreeInfo()
        check(lib().rv_dataset_subtree(self.h, i, C.byref(s)))
        return dict(level=s.level, location=tuple(s.location), node=s.node, depth=s.depth,
                    values_pos=s.values_pos)

    def find_subtree(self, level, loc):
        i = C.c_uint64()
        l3 = _np(loc, np.uint32)
        check(lib().rv_dataset_find_subtree(self.h, level, _p(l3, C.c_uint32), C.byref(i)))
        return i.value

    def closest_subtree(self, level, block):
        i = C.c_uint64()
        b3 = _np(block, np.uint32)
        check(lib().rv_dataset_closest_subtree(self.h, level, _p(b3, C.c_uint32), C.byref(i)))
        return i.value

    def children(self, i):
        n = C.c_uint64()
        lib().rv_dataset_subtree_children(self.h, i, None, 0, C.byref(n))
        if n.value == 0:
            check(lib().rv_dataset_subtree_children(self.h, i, None, 0, C.byref(n)))
        out = np.zeros((n.value, 8), dtype=np.uint32)
        check(lib().rv_dataset_subtree_children(self.h, i, _p(out, C.c_uint32), n.value, C.byref(n)))
        return out

    def n2i(self, i):
        return subtree_n2i(self.children(i), self.subtree(i)["depth"])

    def read_field(self, field_path, i, n):
        rng = np.zeros(2, dtype=np.float32)
        vals = np.zeros(n, dtype=np.float32)
        check(lib().rv_dataset_read_field(self.h, field_path.encode(), i, _p(rng, C.c_float),
                                          _p(vals, C.c_float), n))
        return rng, vals

    def load_worker(self, worker, field_path, level_min, level_max, secondary=False):
        """Register the subtrees of levels level_min..level_max with the
        worker, values from field_path; secondary: as the secondary field of
        subtrees already registered."""
        n = C.c_uint64()
        f = lib().rv_dataset_load_worker_secondary if secondary else lib().rv_dataset_load_worker
        check(f(self.h, field_path.encode(), worker.h, level_min, level_max, C.byref(n)))
        return n.value

    def close(self):
        if self.h:
            lib().rv_dataset_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SharedMem:
    """rv_shm: a System V segment with the reference's SharedMemory layout
    (size, control word, user bytes at offset 12)."""

    USER_OFFSET = 12

    def __init__(self, nbytes=None, shmid=None):
        h = C.c_void_p()
        if shmid is None:
            check(lib().rv_shm_create(nbytes, C.byref(h)))
        else:
            check(lib().rv_shm_attach(shmid, C.byref(h)))
        self.h = h

    @property
    def shmid(self):
        return lib().rv_shm_id(self.h)

    @property
    def size(self):
        return lib().rv_shm_size(self.h)

    def read(self, start, count):
        out = np.zeros(count, dtype=np.uint8)
        check(lib().rv_shm_read(self.h, out.ctypes.data, start, count))
        return out

    def write(self, start, data):
        d = np.ascontiguousarray(np.frombuffer(bytes(data), dtype=np.uint8))
        check(lib().rv_shm_write(self.h, d.ctypes.data, start, d.size))

    def write_frame(self, frame_size, frame_number, duration_ms=0, level=0, pixels=None):
        hdr = FrameOutputHeader(frame_number, duration_ms, level, 0)
        px = None if pixels is None else np.ascontiguousarray(pixels, dtype=np.uint32)
        check(lib().rv_frame_output_write_shm(self.h, frame_size, C.byref(hdr),
                                              None if px is None else px.ctypes.data))

    def read_frame(self, frame_size, pixels=True):
        """(frame_number, duration_ms, level, pixels or None)."""
        hdr = FrameOutputHeader()
        px = np.zeros(frame_size // 4, dtype=np.uint32) if pixels else None
        check(lib().rv_frame_output_read_shm(self.h, frame_size, C.byref(hdr),
                                             None if px is None else px.ctypes.data))
        return hdr.frame_number, hdr.duration_ms, hdr.level, px

    def close(self):
        if self.h:
            lib().rv_shm_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Controller:
    """rv_controller: VsController + VsRenderer for one GPU; frames go to a
    shared-memory frame output."""

    def __init__(self, thread_id=0, split_count=1, measure_time=False):
        h = C.c_void_p()
        check(lib().rv_controller_create(thread_id, split_count, int(measure_time), C.byref(h)))
        self.h = h

    def set_problem(self, path, time_frame=0, primary=None, secondary=None):
        check(lib().rv_controller_set_problem(self.h, path.encode(), time_frame,
                                              primary.encode() if primary else None,
                                              secondary.encode() if secondary else None))

    def fields(self):
        n = C.c_uint32()
        check(lib().rv_controller_field_count(self.h, C.byref(n)))
        out = []
        for i in range(n.value):
            f = FieldInfo()
            check(lib().rv_controller_field_info(self.h, i, C.byref(f)))
            out.append((f.name.decode(), f.min, f.max, bool(f.animated)))
        return out

    def time_frame_count(self):
        n = C.c_uint32()
        check(lib().rv_controller_time_frame_count(self.h, C.byref(n)))
        return n.value

    def set_input(self, vp_w, vp_h, fov, quality, thr_rel, alpha_cutoff, background, tf, view16=None,
                  field_mode=FIELD_ARGB, iso_level_rel=0.5, iso_opacity=1.0, iso_levels=None, exact=False):
        view = None if view16 is None else _np(view16, np.float32).reshape(16)
        tfa = _np(tf, np.float32)
        lv = _np(iso_levels if iso_levels is not None else [], np.float32).reshape(-1)
        st = RenderState(None if view is None else view.ctypes.data, vp_w, vp_h, fov, quality, thr_rel,
                         alpha_cutoff, 0.0, 1.0, (C.c_float * 3)(*background), tfa.ctypes.data, tfa.shape[0],
                         field_mode, iso_level_rel, iso_opacity, lv.ctypes.data if lv.size else None, lv.size,
                         0.0, 1.0, 1 if exact else 0)
        check(lib().rv_controller_set_input(self.h, C.byref(st)))

    def set_camera(self, view16):
        v = _np(view16, np.float32).reshape(16)
        check(lib().rv_controller_set_camera(self.h, _p(v, C.c_float)))

    def camera(self):
        v = np.zeros(16, dtype=np.float32)
        check(lib().rv_controller_camera(self.h, _p(v, C.c_float)))
        return v.reshape(4, 4)

    def reset_camera(self):
        check(lib().rv_controller_reset_camera(self.h))

    def set_render_level(self, level):
        check(lib().rv_controller_set_render_level(self.h, level))

    def start(self):
        check(lib().rv_controller_start(self.h))

    def frame_output(self):
        """(shmid, frame_size, width, height)."""
        sid, fs, w, h = C.c_int(), C.c_uint64(), C.c_uint32(), C.c_uint32()
        check(lib().rv_controller_frame_output(self.h, C.byref(sid), C.byref(fs), C.byref(w), C.byref(h)))
        return sid.value, fs.value, w.value, h.value

    def wait_idle(self, timeout_s=60.0):
        """(idle, frames written)."""
        idle, n = C.c_int(), C.c_uint64()
        check(lib().rv_controller_wait_idle(self.h, int(timeout_s * 1000), C.byref(idle), C.byref(n)))
        return bool(idle.value), n.value

    def status(self):
        buf = C.create_string_buffer(512)
        rc = lib().rv_controller_status(self.h, buf, 512)
        return rc, buf.value.decode()

    def close(self):
        if self.h:
            lib().rv_controller_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = [n for n in dir() if not n.startswith("_")] + ["RvError"]


def probe_peaks(torch, cus=None, reps=5):
    """Measured peaks of the chip (rv_probe_*), each the median of `reps`
    timed launches after a warm-up, on torch's current stream:
    valu_fma_s (non-packed fp32 lane fmas/s), lds_bytes_s (ds_read_b128),
    gather_coalesced_s and gather_scattered_s (wave dword-load instructions/s
    from L1, 4 and 64 sectors per instruction), stream_bytes_s (HBM bytes
    read + written per second, 1 GiB copy)."""
    import statistics
    if cus is None:
        cus = torch.cuda.get_device_properties(0).multi_processor_count
    blocks = cus * 16
    sink = torch.zeros(blocks * 256, dtype=torch.float32, device="cuda")
    src = torch.ones(4096, dtype=torch.float32, device="cuda")
    big_a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    big_b = torch.empty_like(big_a)
    s = stream_ptr()
    out = {}
    cnt = C.c_uint64()

    def timed(launch):
        launch()
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            launch()
            e1.record()
            e1.synchronize()
            ts.append(e0.elapsed_time(e1) / 1e3)
        return statistics.median(ts)

    t = timed(lambda: check(lib().rv_probe_valu(ptr(sink), blocks, 4096, C.byref(cnt), s)))
    out["valu_fma_s"] = cnt.value / t
    t = timed(lambda: check(lib().rv_probe_lds(ptr(sink), blocks, 4096, C.byref(cnt), s)))
    out["lds_bytes_s"] = cnt.value / t
    for name, sc in (("gather_coalesced_s", 0), ("gather_scattered_s", 1)):
        t = timed(lambda: check(lib().rv_probe_gather(ptr(src), ptr(sink), blocks, 2048, sc, C.byref(cnt), s)))
        out[name] = cnt.value / t
    nbytes = big_a.numel() * 4
    t = timed(lambda: check(lib().rv_probe_stream(ptr(big_a), ptr(big_b), nbytes, cus * 8, s)))
    out["stream_bytes_s"] = 2 * nbytes / t
    out["cus"] = cus
    del big_a, big_b
    return out
