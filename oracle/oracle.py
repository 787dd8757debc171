"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and the CPU-baseline
leg of bench.py import this module, as the checker / timed CPU baseline.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


class BlockParams(C.Structure):
    _fields_ = [
        ("vp_w", C.c_uint32), ("vp_h", C.c_uint32),
        ("view_x", C.c_uint32), ("view_y", C.c_uint32), ("view_w", C.c_uint32), ("view_h", C.c_uint32),
        ("center", C.c_float * 3), ("half", C.c_float), ("level", C.c_uint32),
        ("tex_depth", C.c_uint32), ("fmin", C.c_float), ("fmax", C.c_float),
        ("pos", C.c_float * 3), ("e1", C.c_float * 3), ("e2", C.c_float * 3), ("n", C.c_float * 3),
        ("w", C.c_float), ("h", C.c_float), ("pixel_size", C.c_float),
        ("threshold", C.c_float), ("quality", C.c_float), ("alpha_cutoff", C.c_float),
    ]


class FrameDesc(C.Structure):
    _fields_ = [
        ("view16", C.POINTER(C.c_float)), ("vp_w", C.c_uint32), ("vp_h", C.c_uint32),
        ("fov_y_deg", C.c_float), ("level", C.c_uint32),
        ("box_min", C.c_uint32 * 3), ("box_max", C.c_uint32 * 3), ("tex_depth", C.c_uint32),
        ("fmin", C.c_float), ("fmax", C.c_float), ("threshold_rel", C.c_float),
        ("quality", C.c_float), ("alpha_cutoff", C.c_float),
        ("tf", C.POINTER(C.c_float)), ("tf_n", C.c_uint32),
    ]


class BlockRef(C.Structure):
    _fields_ = [("level", C.c_uint32), ("index", C.c_uint32 * 3), ("tex_depth", C.c_uint32),
                ("brick", C.POINTER(C.c_float)), ("brick2", C.POINTER(C.c_float))]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        u32p = C.POINTER(C.c_uint32)
        u16p = C.POINTER(C.c_uint16)
        f32p = C.POINTER(C.c_float)
        L.orc_subtree_depth.restype = C.c_uint32
        L.orc_subtree_depth.argtypes = [u32p, C.c_uint32, C.c_uint32]
        L.orc_subtree_n2i.restype = C.c_int64
        L.orc_subtree_n2i.argtypes = [u32p, C.c_uint32, C.c_uint32, u16p, C.c_uint64]
        L.orc_ordinals.argtypes = [u16p, C.c_uint64, C.c_uint32, u32p]
        L.orc_interpolate.restype = C.c_int
        L.orc_interpolate.argtypes = [f32p, u16p, C.c_uint64, C.c_uint32, f32p]
        L.orc_uniform_split.argtypes = [C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_float, u32p]
        L.orc_nonuniform_split.argtypes = [C.c_uint32, f32p, C.c_uint32, C.c_float, C.c_float,
                                           C.c_float, u32p]
        L.orc_b2f_macro.restype = C.c_int64
        L.orc_b2f_macro.argtypes = [C.c_uint32, u32p, u32p, f32p, f32p, f32p, u32p, C.c_uint64]
        L.orc_index_cube_split_sorted.restype = C.c_int64
        L.orc_index_cube_split_sorted.argtypes = [C.c_uint32, C.c_uint32, f32p, u32p, C.c_uint64]
        L.orc_index_cube_split_all.restype = C.c_int64
        L.orc_index_cube_split_all.argtypes = [C.c_uint32, C.c_uint32, C.c_int, u32p, C.c_uint64]
        L.orc_block_params_setup.restype = C.c_int
        L.orc_block_params_setup.argtypes = [f32p, C.c_uint32, C.c_uint32, C.c_float, C.c_uint32, u32p,
                                             C.c_uint32, C.c_float, C.c_float, C.c_float, C.c_float,
                                             C.c_float, C.POINTER(BlockParams)]
        L.orc_render_block.argtypes = [C.POINTER(BlockParams), f32p, f32p, C.c_uint32, u32p,
                                       C.c_uint32, C.c_uint32]
        L.orc_ray_segment.restype = C.c_int
        L.orc_ray_segment.argtypes = [C.POINTER(BlockParams), f32p, f32p, C.c_uint32, C.c_uint32,
                                      C.c_uint32, f32p]
        L.orc_block_derived.argtypes = [C.POINTER(BlockParams), f32p]
        L.orc_ray_samples.restype = C.c_int
        L.orc_ray_samples.argtypes = [C.POINTER(BlockParams), f32p, f32p, C.c_uint32, C.c_uint32,
                                      C.c_uint32, C.c_uint32, u32p, f32p, C.POINTER(C.c_float)]
        L.orc_powf.restype = C.c_float
        L.orc_powf.argtypes = [C.c_float, C.c_float]
        L.orc_bound_rect.restype = C.c_int
        L.orc_bound_rect.argtypes = [u32p, C.c_uint32, C.c_uint32, C.POINTER(C.c_int32)]
        L.orc_blend_image_part.argtypes = [u32p, C.c_int32, C.c_int32, u32p, C.c_int32, C.c_int32,
                                           C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int]
        L.orc_background_argb.restype = C.c_uint32
        L.orc_background_argb.argtypes = [C.c_float, C.c_float, C.c_float]
        L.orc_color_transfer_vector.argtypes = [f32p, f32p, C.c_uint32, C.c_uint32, f32p]
        L.orc_render_share.argtypes = [C.POINTER(FrameDesc), C.POINTER(f32p), u32p, C.c_uint32]
        L.orc_render_float_composite.argtypes = [f32p, C.c_uint32, C.c_uint32, C.c_float,
                                                 C.POINTER(BlockRef), C.c_uint32, C.c_float, C.c_float,
                                                 C.c_float, C.c_float, C.c_float, C.c_float, f32p,
                                                 C.c_uint32, f32p, C.c_uint32]
        L.orc_render_block_mode.argtypes = [C.POINTER(BlockParams), C.c_uint32, C.c_float, f32p, C.c_uint32,
                                            f32p, f32p, C.c_uint32, u32p, C.c_uint32, C.c_uint32]
        L.orc_render_block_mode2.argtypes = [C.POINTER(BlockParams), C.c_uint32, C.c_float, f32p,
                                             C.c_uint32, f32p, f32p, C.c_float, C.c_float, f32p, C.c_uint32,
                                             u32p, C.c_uint32, C.c_uint32]
        L.orc_blend_image_part_max_intensity.argtypes = [f32p, C.c_int32, f32p, C.c_int32, C.c_int32,
                                                         C.c_int32, C.c_int32, C.c_int]
        L.orc_mip_colorize.argtypes = [u32p, C.c_int32, f32p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                       f32p, C.c_uint32]
        L.orc_render_float_composite_mode.argtypes = [f32p, C.c_uint32, C.c_uint32, C.c_float,
                                                      C.POINTER(BlockRef), C.c_uint32, C.c_float,
                                                      C.c_float, C.c_float, C.c_float, C.c_float,
                                                      C.c_float, f32p, C.c_uint32, C.c_uint32, C.c_float,
                                                      f32p, C.c_uint32, f32p, C.c_uint32]
        L.orc_render_float_composite_mode2.argtypes = [f32p, C.c_uint32, C.c_uint32, C.c_float,
                                                       C.POINTER(BlockRef), C.c_uint32, C.c_float,
                                                       C.c_float, C.c_float, C.c_float, C.c_float,
                                                       C.c_float, f32p, C.c_uint32, C.c_uint32, C.c_float,
                                                       f32p, C.c_uint32, C.c_float, C.c_float, f32p,
                                                       C.c_uint32]
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def subtree_depth(children, max_depth):
    ch = np.ascontiguousarray(children, dtype=np.uint32)
    return int(lib().orc_subtree_depth(_p(ch, C.c_uint32), ch.shape[0], max_depth))


def subtree_n2i(children, depth):
    ch = np.ascontiguousarray(children, dtype=np.uint32)
    cap = (2 ** depth + 1) ** 3
    out = np.zeros((cap, 3), dtype=np.uint16)
    n = lib().orc_subtree_n2i(_p(ch, C.c_uint32), ch.shape[0], depth, _p(out, C.c_uint16), cap)
    if n < 0:
        raise ValueError("oracle n2i failed")
    return out[:n].copy()


def ordinals(n2i, depth):
    n2i = np.ascontiguousarray(n2i, dtype=np.uint16)
    out = np.zeros(n2i.shape[0], dtype=np.uint32)
    lib().orc_ordinals(_p(n2i, C.c_uint16), n2i.shape[0], depth, _p(out, C.c_uint32))
    return out


def interpolate(sparse, n2i, depth):
    sparse = np.ascontiguousarray(sparse, dtype=np.float32)
    n2i = np.ascontiguousarray(n2i, dtype=np.uint16)
    npe = 2 ** depth + 1
    dense = np.empty(npe ** 3, dtype=np.float32)
    rc = lib().orc_interpolate(_p(sparse, C.c_float), _p(n2i, C.c_uint16), n2i.shape[0], depth,
                               _p(dense, C.c_float))
    if rc != 0:
        raise ValueError("oracle interpolate failed: %d" % rc)
    return dense


def _split(out5):
    b = None if out5[0] == out5[1] else [int(out5[0]), int(out5[1])]
    at = None if out5[2] == 0xFFFFFFFF else int(out5[2])
    a = None if out5[3] == out5[4] else [int(out5[3]), int(out5[4])]
    return b, at, a


def uniform_split(i0, i1, c0, c1, x):
    out = np.zeros(5, dtype=np.uint32)
    lib().orc_uniform_split(i0, i1, c0, c1, x, _p(out, C.c_uint32))
    return _split(out)


def nonuniform_split(start, split, c0, c1, x):
    sp = np.ascontiguousarray(split, dtype=np.float32)
    out = np.zeros(5, dtype=np.uint32)
    lib().orc_nonuniform_split(start, _p(sp, C.c_float), sp.size, c0, c1, x, _p(out, C.c_uint32))
    return _split(out)


def b2f_macro(imin, imax, cmin, cmax, eye):
    dim = len(imin)
    a = lambda v, t: np.ascontiguousarray(v, dtype=t)
    imin, imax = a(imin, np.uint32), a(imax, np.uint32)
    cmin, cmax, eye = a(cmin, np.float32), a(cmax, np.float32), a(eye, np.float32)
    cap = int(np.prod(imax.astype(np.int64) - imin.astype(np.int64)))
    out = np.zeros((max(cap, 1), dim), dtype=np.uint32)
    n = lib().orc_b2f_macro(dim, _p(imin, C.c_uint32), _p(imax, C.c_uint32), _p(cmin, C.c_float),
                            _p(cmax, C.c_float), _p(eye, C.c_float), _p(out, C.c_uint32), cap)
    if n < 0:
        raise ValueError("b2f overflow")
    return out[:n].copy()


def index_cube_split_sorted(level, count, eye):
    eye = np.ascontiguousarray(eye, dtype=np.float32)
    cap = count + 8
    out = np.zeros((cap, 6), dtype=np.uint32)
    n = lib().orc_index_cube_split_sorted(level, count, _p(eye, C.c_float), _p(out, C.c_uint32), cap)
    return out[:n].copy()


def index_cube_split_all(level, count, add_trailing_empty=True):
    cap = max(count, 1) + (1 << (3 * level)) + 8
    out = np.zeros((cap, 6), dtype=np.uint32)
    n = lib().orc_index_cube_split_all(level, count, int(add_trailing_empty), _p(out, C.c_uint32), cap)
    return out[:n].copy()


def block_params(view16, vp_w, vp_h, fov, level, index, tex_depth, fmin, fmax, thr_rel, quality,
                 alpha_cutoff):
    v = np.ascontiguousarray(view16, dtype=np.float32).reshape(16)
    idx = np.ascontiguousarray(index, dtype=np.uint32)
    p = BlockParams()
    live = lib().orc_block_params_setup(_p(v, C.c_float), vp_w, vp_h, fov, level, _p(idx, C.c_uint32),
                                        tex_depth, fmin, fmax, thr_rel, quality, alpha_cutoff,
                                        C.byref(p))
    return live, p


def render_block(params, brick, tf, viewport):
    brick = np.ascontiguousarray(brick, dtype=np.float32)
    tf = np.ascontiguousarray(tf, dtype=np.float32)
    assert viewport.dtype == np.uint32 and viewport.flags.c_contiguous
    lib().orc_render_block(C.byref(params), _p(brick, C.c_float), _p(tf, C.c_float), tf.shape[0],
                           _p(viewport, C.c_uint32), 0, params.vp_h)


def render_block_mode(params, mode, iso_opacity, brick, tf, viewport, levels=None, brick2=None,
                      srange=(0.0, 1.0)):
    """Field modes: 0 Isosurface, 1 Isosurfaces (levels relative), 2
    MaxIntensityProjection (float bits), 3 Argb, 4 ArgbLight, 5 DomainVoxels,
    6 ValueOnIsosurface and 7 ValueOnIsosurfaces (brick2: the secondary
    field's brick, srange its range)."""
    brick = np.ascontiguousarray(brick, dtype=np.float32)
    tf = np.ascontiguousarray(tf, dtype=np.float32)
    lv = np.ascontiguousarray(levels if levels is not None else [0.0], dtype=np.float32)
    nlv = 0 if levels is None else len(levels)
    assert viewport.dtype == np.uint32 and viewport.flags.c_contiguous
    if mode in (6, 7):
        b2 = np.ascontiguousarray(brick2, dtype=np.float32)
        assert b2.shape == brick.shape
        lib().orc_render_block_mode2(C.byref(params), mode, iso_opacity, _p(lv, C.c_float), nlv,
                                     _p(brick, C.c_float), _p(b2, C.c_float), srange[0], srange[1],
                                     _p(tf, C.c_float), tf.shape[0], _p(viewport, C.c_uint32), 0,
                                     params.vp_h)
        return
    lib().orc_render_block_mode(C.byref(params), mode, iso_opacity, _p(lv, C.c_float), nlv,
                                _p(brick, C.c_float), _p(tf, C.c_float), tf.shape[0],
                                _p(viewport, C.c_uint32), 0, params.vp_h)


def blend_image_part_max_intensity(dst, src, dx, dy, just_copy=False):
    """dst, src: 2-D float32 arrays (rows of the image); in place on dst."""
    assert dst.dtype == np.float32 and src.dtype == np.float32 and dst.flags.c_contiguous
    src = np.ascontiguousarray(src)
    lib().orc_blend_image_part_max_intensity(_p(dst, C.c_float), dst.shape[1], _p(src, C.c_float),
                                             src.shape[1], src.shape[0], dx, dy, int(just_copy))


def mip_colorize(dst, src, dx, dy, ctv):
    """dst: 2-D uint32 frame, src: 2-D float32 part at (dx, dy), ctv: n x 4."""
    assert dst.dtype == np.uint32 and dst.flags.c_contiguous
    src = np.ascontiguousarray(src, dtype=np.float32)
    ctv = np.ascontiguousarray(ctv, dtype=np.float32)
    lib().orc_mip_colorize(_p(dst, C.c_uint32), dst.shape[1], _p(src, C.c_float), src.shape[1],
                           src.shape[0], dx, dy, _p(ctv, C.c_float), ctv.shape[0])


def ray_segment(params, brick, tf, vx, vy):
    out = np.zeros(4, dtype=np.float32)
    hit = lib().orc_ray_segment(C.byref(params), _p(brick, C.c_float), _p(tf, C.c_float), tf.shape[0],
                                vx, vy, _p(out, C.c_float))
    return hit, out


def block_derived(params):
    """{texel, tcf, i_layer_count, opacity_factor} of a block (float32)."""
    out = np.zeros(4, dtype=np.float32)
    lib().orc_block_derived(C.byref(params), _p(out, C.c_float))
    return out


def ray_samples(params, brick, tf, vx, vy, max_samples=8192):
    """Kept samples of one pixel's segment of one block: (corners (k, 8)
    uint32, weights (k, 3) float32, segment alpha)."""
    corners = np.zeros((max_samples, 8), dtype=np.uint32)
    weights = np.zeros((max_samples, 3), dtype=np.float32)
    a = C.c_float()
    k = lib().orc_ray_samples(C.byref(params), _p(brick, C.c_float), _p(tf, C.c_float), tf.shape[0], vx, vy,
                              max_samples, _p(corners, C.c_uint32), _p(weights, C.c_float), C.byref(a))
    if k < 0:
        raise ValueError("more than %d samples" % max_samples)
    return corners[:k], weights[:k], np.float32(a.value)


def powf(b, e):
    return float(lib().orc_powf(b, e))


def bound_rect(pix):
    pix = np.ascontiguousarray(pix, dtype=np.uint32)
    r = (C.c_int32 * 4)()
    ok = lib().orc_bound_rect(_p(pix, C.c_uint32), pix.shape[1], pix.shape[0], r)
    return tuple(r) if ok else None


def blend_image_part(dst, src, src_xy, dst_off=(0, 0), just_copy=False):
    assert dst.dtype == np.uint32 and dst.flags.c_contiguous
    src = np.ascontiguousarray(src, dtype=np.uint32)
    lib().orc_blend_image_part(_p(dst, C.c_uint32), dst.shape[1], dst.shape[0], _p(src, C.c_uint32),
                               src.shape[1], src.shape[0], src_xy[0], src_xy[1], dst_off[0],
                               dst_off[1], int(just_copy))


def background_argb(r, g, b):
    return int(lib().orc_background_argb(r, g, b))


def color_transfer_vector(xs, colors, count):
    xs = np.ascontiguousarray(xs, dtype=np.float32)
    colors = np.ascontiguousarray(colors, dtype=np.float32)
    out = np.zeros((count, 4), dtype=np.float32)
    lib().orc_color_transfer_vector(_p(xs, C.c_float), _p(colors, C.c_float), xs.size, count,
                                    _p(out, C.c_float))
    return out


def render_share(view16, vp_w, vp_h, fov, level, box_min, box_max, tex_depth, fmin, fmax, thr_rel,
                 quality, alpha_cutoff, tf, bricks, nthreads=None):
    """bricks: list of float32 arrays in box index order (x fastest)."""
    v = np.ascontiguousarray(view16, dtype=np.float32).reshape(16)
    tf = np.ascontiguousarray(tf, dtype=np.float32)
    d = FrameDesc()
    d.view16 = _p(v, C.c_float)
    d.vp_w, d.vp_h, d.fov_y_deg, d.level = vp_w, vp_h, fov, level
    for a in range(3):
        d.box_min[a], d.box_max[a] = box_min[a], box_max[a]
    d.tex_depth, d.fmin, d.fmax = tex_depth, fmin, fmax
    d.threshold_rel, d.quality, d.alpha_cutoff = thr_rel, quality, alpha_cutoff
    d.tf, d.tf_n = _p(tf, C.c_float), tf.shape[0]
    bricks = [np.ascontiguousarray(b, dtype=np.float32) for b in bricks]
    ptrs = (C.POINTER(C.c_float) * len(bricks))(*[_p(b, C.c_float) for b in bricks])
    vp = np.zeros((vp_h, vp_w), dtype=np.uint32)
    lib().orc_render_share(C.byref(d), ptrs, _p(vp, C.c_uint32), nthreads or os.cpu_count())
    return vp


def render_float_composite(view16, vp_w, vp_h, fov, blocks_b2f, fmin, fmax, thr_rel, quality,
                           alpha_cutoff, tf, acc_cutoff=2.0, nthreads=None, mode=3, iso_opacity=1.0,
                           levels=None, srange=(0.0, 1.0)):
    """blocks_b2f: list of (level, (ix,iy,iz), tex_depth, brick[, brick2]) in back-to-front
    order; brick2 (the secondary field) for the ValueOnIsosurface(s) modes."""
    v = np.ascontiguousarray(view16, dtype=np.float32).reshape(16)
    tf = np.ascontiguousarray(tf, dtype=np.float32)
    keep = []
    lv = np.ascontiguousarray(levels if levels is not None else [0.0], dtype=np.float32)
    nlv = 0 if levels is None else len(levels)
    arr = (BlockRef * len(blocks_b2f))()
    for i, blk in enumerate(blocks_b2f):
        lvl, idx, td, brick = blk[:4]
        b = np.ascontiguousarray(brick, dtype=np.float32)
        keep.append(b)
        arr[i].level = lvl
        for a in range(3):
            arr[i].index[a] = idx[a]
        arr[i].tex_depth = td
        arr[i].brick = _p(b, C.c_float)
        if len(blk) > 4 and blk[4] is not None:
            b2 = np.ascontiguousarray(blk[4], dtype=np.float32)
            assert b2.shape == b.shape
            keep.append(b2)
            arr[i].brick2 = _p(b2, C.c_float)
        elif mode in (6, 7):
            raise ValueError("the ValueOnIsosurface(s) modes need each block's secondary brick")
    out = np.zeros((vp_h, vp_w, 4), dtype=np.float32)
    lib().orc_render_float_composite_mode2(_p(v, C.c_float), vp_w, vp_h, fov, arr, len(blocks_b2f), fmin,
                                           fmax, thr_rel, quality, alpha_cutoff, acc_cutoff,
                                           _p(tf, C.c_float), tf.shape[0], mode, iso_opacity,
                                           _p(lv, C.c_float), nlv, srange[0], srange[1], _p(out, C.c_float),
                                           nthreads or os.cpu_count())
    return out
