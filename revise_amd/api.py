"""numpy/torch-facing wrappers of the C ABI (host functions take numpy
arrays, device functions take torch CUDA tensors)."""
import ctypes as C

import numpy as np

from . import BlockParams, BrickJob, RvError, check, lib, ptr, stream_ptr


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
                 alpha_cutoff):
    v = _np(view16, np.float32).reshape(16)
    idx = _np(index, np.uint32)
    p = BlockParams()
    live = C.c_int()
    check(lib().rv_block_params_setup(_p(v, C.c_float), vp_w, vp_h, fov, level, _p(idx, C.c_uint32),
                                      tex_depth, fmin, fmax, thr_rel, quality, alpha_cutoff,
                                      C.byref(p), C.byref(live)))
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


def interp_plan_build(d_n2i, depth, d_plan, d_bad=None, stream=None):
    check(lib().rv_interp_plan_build(ptr(d_n2i), d_n2i.shape[0], depth, ptr(d_plan),
                                     ptr(d_bad) if d_bad is not None else None, stream_ptr(stream)))


def interp_brick_planned(d_sparse, d_plan, depth, d_dense, stream=None):
    check(lib().rv_interp_brick_planned(ptr(d_sparse), ptr(d_plan), depth, ptr(d_dense),
                                        stream_ptr(stream)))


def interp_brick(d_sparse, d_n2i, depth, d_dense, d_ws=None, stream=None):
    check(lib().rv_interp_brick(ptr(d_sparse), ptr(d_n2i), d_n2i.shape[0], depth, ptr(d_dense),
                                ptr(d_ws) if d_ws is not None else None, stream_ptr(stream)))


class InterpBatch:
    """Batched interpolation handle; jobs: list of (sparse, plan, dense, depth)
    torch tensors (dense may be float32 or float16)."""

    def __init__(self, jobs):
        import torch
        arr = (BrickJob * max(len(jobs), 1))()
        self._keep = []
        for i, (sp, pl, de, depth) in enumerate(jobs):
            arr[i].sparse = sp.data_ptr()
            arr[i].plan = pl.data_ptr()
            if de.dtype == torch.float16:
                arr[i].dense16 = de.data_ptr()
            else:
                arr[i].dense = de.data_ptr()
            arr[i].depth = depth
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


__all__ = [n for n in dir() if not n.startswith("_")] + ["RvError"]
