"""revise_amd: MI355X (gfx950) sparse->dense interpolation and volume
ray-casting path of ReVisE, behind a C ABI (include/revise_hip.h).

This module is a thin ctypes binding of librevise_hip.so for Python callers
(tests, bench).  It fails loudly when the library is missing: there is no
CPU fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librevise_hip.so")

RV_OK = 0
RENDER_EXACT = 1  # rv_block_params.flags: the exact sample sequence


class RvError(RuntimeError):
    pass


class BlockParams(C.Structure):
    _fields_ = [
        ("vp_w", C.c_uint32), ("vp_h", C.c_uint32),
        ("view_x", C.c_uint32), ("view_y", C.c_uint32), ("view_w", C.c_uint32), ("view_h", C.c_uint32),
        ("center", C.c_float * 3), ("half", C.c_float), ("level", C.c_uint32),
        ("tex_depth", C.c_uint32), ("fmin", C.c_float), ("fmax", C.c_float),
        ("pos", C.c_float * 3), ("e1", C.c_float * 3), ("e2", C.c_float * 3), ("n", C.c_float * 3),
        ("w", C.c_float), ("h", C.c_float), ("pixel_size", C.c_float),
        ("threshold", C.c_float), ("quality", C.c_float), ("alpha_cutoff", C.c_float),
        ("nodes", C.c_uint32), ("texel", C.c_float), ("tex_origin", C.c_float * 3), ("tcf", C.c_float),
        ("i_layer_count", C.c_float), ("opacity_factor", C.c_float), ("flags", C.c_uint32),
    ]


class FrameBlock(C.Structure):
    _fields_ = [("level", C.c_uint32), ("index", C.c_uint32 * 3), ("depth", C.c_uint32),
                ("sparse", C.c_void_p), ("plan", C.c_void_p), ("brick", C.c_void_p),
                ("brick16", C.c_void_p)]


class FrameDesc(C.Structure):
    _fields_ = [("vp_w", C.c_uint32), ("vp_h", C.c_uint32), ("fov_y_deg", C.c_float),
                ("fmin", C.c_float), ("fmax", C.c_float), ("threshold_rel", C.c_float),
                ("quality", C.c_float), ("alpha_cutoff", C.c_float), ("acc_cutoff", C.c_float),
                ("d_tf", C.c_void_p), ("tf_n", C.c_uint32), ("tile_w", C.c_uint32),
                ("tile_h", C.c_uint32), ("layout", C.c_uint32), ("exact", C.c_uint32)]


class RenderState(C.Structure):
    _fields_ = [("view16", C.c_void_p), ("vp_w", C.c_uint32), ("vp_h", C.c_uint32),
                ("fov_y_deg", C.c_float), ("quality", C.c_float), ("threshold_rel", C.c_float),
                ("alpha_cutoff", C.c_float), ("fmin", C.c_float), ("fmax", C.c_float),
                ("background", C.c_float * 3), ("tf", C.c_void_p), ("tf_n", C.c_uint32),
                ("field_mode", C.c_uint32), ("iso_level_rel", C.c_float), ("iso_opacity", C.c_float),
                ("iso_levels_rel", C.c_void_p), ("iso_level_count", C.c_uint32),
                ("smin", C.c_float), ("smax", C.c_float), ("exact", C.c_uint32)]


class BrickJob(C.Structure):
    _fields_ = [("sparse", C.c_void_p), ("plan", C.c_void_p), ("dense", C.c_void_p),
                ("dense16", C.c_void_p), ("depth", C.c_uint32), ("layout", C.c_uint32)]


class DatasetInfo(C.Structure):
    _fields_ = [("format_version", C.c_uint32), ("max_subtree_depth", C.c_uint32),
                ("max_full_level", C.c_uint32), ("level_count", C.c_uint32),
                ("bc_min", C.c_float * 3), ("bc_size", C.c_float), ("node_count", C.c_uint64),
                ("subtree_count", C.c_uint64), ("field_file_size", C.c_uint64)]


class SubtreeInfo(C.Structure):
    _fields_ = [("level", C.c_uint32), ("location", C.c_uint32 * 3), ("node", C.c_uint32),
                ("depth", C.c_uint32), ("values_pos", C.c_uint64)]


class FrameOutputHeader(C.Structure):
    _fields_ = [("frame_number", C.c_uint64), ("duration_ms", C.c_uint64), ("level", C.c_uint32),
                ("pad", C.c_uint32)]


class FieldInfo(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("min", C.c_float), ("max", C.c_float),
                ("animated", C.c_int32)]


_u32p = C.POINTER(C.c_uint32)
_u16p = C.POINTER(C.c_uint16)
_f32p = C.POINTER(C.c_float)
_vp = C.c_void_p

# name: (restype, argtypes)
_SIGS = {
    "rv_last_error": (C.c_char_p, []),
    "rv_version": (C.c_int, []),
    "rv_subtree_depth": (C.c_int, [_u32p, C.c_uint32, C.c_uint32, _u32p]),
    "rv_subtree_n2i": (C.c_int, [_u32p, C.c_uint32, C.c_uint32, _u16p, C.c_uint64,
                                 C.POINTER(C.c_uint64)]),
    "rv_node_ordinals": (C.c_int, [_vp, C.c_uint64, C.c_uint32, _vp, _vp]),
    "rv_brick_nodes": (C.c_uint64, [C.c_uint32]),
    "rv_interp_plan_bytes": (C.c_uint64, [_u16p, C.c_uint64, C.c_uint32]),
    "rv_interp_plan_encode": (C.c_int, [_u16p, C.c_uint64, C.c_uint32, _vp, C.c_uint64]),
    "rv_interp_plan_build": (C.c_int, [_u16p, C.c_uint64, C.c_uint32, C.POINTER(C.c_void_p),
                                       C.POINTER(C.c_uint64)]),
    "rv_interp_plan_free": (C.c_int, [_vp]),
    "rv_interp_brick_planned": (C.c_int, [_vp, _vp, C.c_uint32, _vp, _vp]),
    "rv_interp_brick": (C.c_int, [_vp, _u16p, C.c_uint64, C.c_uint32, _vp, _vp]),
    "rv_interp_batch_create": (C.c_int, [C.POINTER(BrickJob), C.c_uint32, C.POINTER(C.c_void_p)]),
    "rv_interp_batch_run": (C.c_int, [_vp, _vp]),
    "rv_interp_batch_info": (C.c_int, [_vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "rv_interp_batch_destroy": (None, [_vp]),
    "rv_block_params_setup": (C.c_int, [_f32p, C.c_uint32, C.c_uint32, C.c_float, C.c_uint32, _u32p,
                                        C.c_uint32, C.c_float, C.c_float, C.c_float, C.c_float,
                                        C.c_float, C.POINTER(BlockParams), C.POINTER(C.c_int)]),
    "rv_clear_viewport": (C.c_int, [_vp, C.c_uint32, C.c_uint32, _vp]),
    "rv_render_block_argb": (C.c_int, [C.POINTER(BlockParams), _vp, _vp, C.c_uint32, _vp, _vp]),
    "rv_block_sample_count": (C.c_int, [C.POINTER(BlockParams), _vp, _vp]),
    "rv_subtree_n2i_device": (C.c_int, [_vp, C.c_uint32, C.c_uint32, _vp, C.c_uint64,
                                        C.POINTER(C.c_uint64), _vp]),
    "rv_render_block": (C.c_int, [C.POINTER(BlockParams), C.c_uint32, C.c_float, _f32p, C.c_uint32, _vp,
                                  _vp, C.c_uint32, _vp, _vp]),
    "rv_clear_viewport_mode": (C.c_int, [_vp, C.c_uint32, C.c_uint32, C.c_uint32, _vp]),
    "rv_blend_image_part_max_intensity": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _vp,
                                                    C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int,
                                                    _vp]),
    "rv_mip_colorize": (C.c_int, [_vp, C.c_int32, C.c_int32, _vp, C.c_int32, C.c_int32, C.c_int32,
                                  C.c_int32, _vp, C.c_uint32, _vp]),
    "rv_render_share": (C.c_int, [C.POINTER(BlockParams), C.POINTER(C.c_void_p), C.c_uint32, C.c_uint32,
                                  C.c_float, _f32p, C.c_uint32, _vp, C.c_uint32, _vp, C.c_uint32,
                                  C.c_uint32, _vp, _vp]),
    "rv_share_workspace_bytes": (C.c_uint64, [C.c_uint32]),
    "rv_probe_valu": (C.c_int, [_vp, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64), _vp]),
    "rv_probe_lds": (C.c_int, [_vp, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64), _vp]),
    "rv_probe_gather": (C.c_int, [_vp, _vp, C.c_uint32, C.c_uint32, C.c_int, C.POINTER(C.c_uint64), _vp]),
    "rv_probe_stream": (C.c_int, [_vp, _vp, C.c_uint64, C.c_uint32, _vp]),
    "rv_render_share_argb_noalloc": (C.c_int, [C.POINTER(BlockParams), C.POINTER(C.c_void_p), C.c_uint32, _vp,
                                               C.c_uint32, _vp, C.c_uint32, C.c_uint32, _vp]),
    "rv_brick_floats": (C.c_uint64, [C.c_uint32, C.c_uint32]),
    "rv_render_share_layout": (C.c_int, [C.POINTER(BlockParams), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                         C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, _f32p, C.c_uint32,
                                         C.c_float, C.c_float, _vp, C.c_uint32, _vp, C.c_uint32, C.c_uint32,
                                         _vp, _vp]),
    "rv_render_block_secondary": (C.c_int, [C.POINTER(BlockParams), C.c_uint32, C.c_float, _f32p,
                                            C.c_uint32, _vp, _vp, C.c_float, C.c_float, _vp, C.c_uint32,
                                            _vp, _vp]),
    "rv_render_share_secondary": (C.c_int, [C.POINTER(BlockParams), C.POINTER(C.c_void_p),
                                            C.POINTER(C.c_void_p), C.c_uint32, C.c_uint32, C.c_float,
                                            _f32p, C.c_uint32, C.c_float, C.c_float, _vp, C.c_uint32, _vp,
                                            C.c_uint32, C.c_uint32, _vp, _vp]),
    "rv_render_share_argb": (C.c_int, [C.POINTER(BlockParams), C.POINTER(C.c_void_p), C.c_uint32, _vp,
                                       C.c_uint32, _vp, C.c_uint32, C.c_uint32, _vp, _vp]),
    "rv_color_transfer_vector": (C.c_int, [_f32p, _f32p, C.c_uint32, C.c_uint32, _f32p]),
    "rv_back_to_front": (C.c_int, [_u32p, _u32p, _f32p, _f32p, _f32p, _u32p, C.c_uint64,
                                   C.POINTER(C.c_uint64)]),
    "rv_eye_from_view": (C.c_int, [_f32p, _f32p]),
    "rv_index_box_coords": (C.c_int, [C.c_uint32, _u32p, _u32p, _f32p, _f32p]),
    "rv_index_cube_split_sorted": (C.c_int, [C.c_uint32, C.c_uint32, _f32p, _u32p, C.c_uint64,
                                             C.POINTER(C.c_uint64)]),
    "rv_mixed_back_to_front": (C.c_int, [_u32p, C.c_uint64, _f32p, _u32p]),
    "rv_bound_rect": (C.c_int, [_vp, C.c_uint32, C.c_uint32, C.POINTER(C.c_int32), C.POINTER(C.c_int),
                                _vp]),
    "rv_blend_image_part": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _vp, C.c_int32,
                                      C.c_int32, C.c_int32, C.c_int32, C.c_int, _vp]),
    "rv_fill_u32": (C.c_int, [_vp, C.c_uint64, C.c_uint32, _vp]),
    "rv_background_argb": (C.c_uint32, [C.c_float, C.c_float, C.c_float]),
    "rv_frame_plan_create": (C.c_int, [C.POINTER(FrameBlock), C.c_uint32, C.POINTER(FrameDesc),
                                       C.POINTER(C.c_void_p)]),
    "rv_frame_plan_set_cameras": (C.c_int, [_vp, _f32p, C.c_uint32]),
    "rv_frame_render": (C.c_int, [_vp, C.c_uint32, _vp, _vp, C.c_int, _vp]),
    "rv_frame_graph_build": (C.c_int, [_vp, _vp, C.c_uint64, C.c_int, _vp]),
    "rv_frame_graph_launch": (C.c_int, [_vp, _vp]),
    "rv_frame_status": (C.c_int, [_vp, _u32p]),
    "rv_frame_get_tile": (C.c_int, [_vp, _u32p, _u32p]),
    "rv_frame_set_schedule": (C.c_int, [_vp, C.c_uint32]),
    "rv_frame_get_schedule": (C.c_int, [_vp, _u32p]),
    "rv_frame_set_hybrid_min": (C.c_int, [_vp, C.c_uint32]),
    "rv_frame_set_grouped": (C.c_int, [_vp, C.c_int]),
    "rv_frame_set_exact": (C.c_int, [_vp, C.c_int]),
    "rv_frame_set_split": (C.c_int, [_vp, C.c_int]),
    "rv_frame_get_split": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "rv_frame_autotune": (C.c_int, [_vp, _vp, C.c_int, C.c_char_p, _vp, _u32p, _u32p]),
    "rv_frame_set_tape": (C.c_int, [_vp, C.c_int]),
    "rv_frame_set_field_mode": (C.c_int, [_vp, C.c_uint32, C.c_float, _f32p, C.c_uint32]),
    "rv_frame_set_secondary": (C.c_int, [_vp, _vp, _vp, C.c_float, C.c_float]),
    "rv_frame_set_grad_bricks": (C.c_int, [_vp, _vp]),
    "rv_frame_adjoint": (C.c_int, [_vp, C.c_uint32, _vp, _vp, _vp, _vp]),
    "rv_frame_plan_destroy": (None, [_vp]),
    "rv_worker_create": (C.c_int, [C.c_uint32, C.POINTER(C.c_void_p)]),
    "rv_worker_add_subtree": (C.c_int, [_vp, C.c_uint32, _u32p, C.c_uint32, _u16p, C.c_uint64, _f32p]),
    "rv_worker_add_subtree_secondary": (C.c_int, [_vp, C.c_uint32, _u32p, _f32p]),
    "rv_worker_set_state": (C.c_int, [_vp, C.POINTER(RenderState)]),
    "rv_worker_render_scene_part": (C.c_int, [_vp, C.c_uint32, _u32p, _u32p, C.POINTER(C.c_int),
                                              C.POINTER(C.c_int32), _u32p, C.c_uint64]),
    "rv_worker_compose_frame": (C.c_int, [_vp, C.c_uint32, C.c_uint32, _u32p, C.POINTER(C.c_int)]),
    "rv_worker_set_fused": (C.c_int, [_vp, C.c_int]),
    "rv_worker_render_progressive": (C.c_int, [_vp, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_int),
                                               _vp, C.c_uint64, C.POINTER(C.c_uint64), C.c_int, _u32p]),
    "rv_frame_output_read": (C.c_int, [_vp, C.c_uint64, C.c_uint32, C.c_uint32,
                                       C.POINTER(FrameOutputHeader), _u32p]),
    "rv_worker_destroy": (None, [_vp]),
    "rv_dataset_open": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "rv_dataset_close": (None, [_vp]),
    "rv_dataset_get_info": (C.c_int, [_vp, C.POINTER(DatasetInfo)]),
    "rv_dataset_subtree": (C.c_int, [_vp, C.c_uint64, C.POINTER(SubtreeInfo)]),
    "rv_dataset_find_subtree": (C.c_int, [_vp, C.c_uint32, _u32p, C.POINTER(C.c_uint64)]),
    "rv_dataset_closest_subtree": (C.c_int, [_vp, C.c_uint32, _u32p, C.POINTER(C.c_uint64)]),
    "rv_dataset_subtree_children": (C.c_int, [_vp, C.c_uint64, _u32p, C.c_uint64,
                                              C.POINTER(C.c_uint64)]),
    "rv_dataset_read_field": (C.c_int, [_vp, C.c_char_p, C.c_uint64, _f32p, _f32p, C.c_uint64]),
    "rv_dataset_load_worker": (C.c_int, [_vp, C.c_char_p, _vp, C.c_uint32, C.c_uint32,
                                         C.POINTER(C.c_uint64)]),
    "rv_dataset_load_worker_secondary": (C.c_int, [_vp, C.c_char_p, _vp, C.c_uint32, C.c_uint32,
                                                   C.POINTER(C.c_uint64)]),
    "rv_fields_info_read": (C.c_int, [C.c_char_p, _u32p, C.POINTER(FieldInfo), C.c_uint32, _u32p]),
    "rv_shm_create": (C.c_int, [C.c_uint64, C.POINTER(C.c_void_p)]),
    "rv_shm_attach": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "rv_shm_id": (C.c_int, [_vp]),
    "rv_shm_size": (C.c_uint64, [_vp]),
    "rv_shm_read": (C.c_int, [_vp, _vp, C.c_uint64, C.c_uint64]),
    "rv_shm_write": (C.c_int, [_vp, _vp, C.c_uint64, C.c_uint64]),
    "rv_shm_close": (None, [_vp]),
    "rv_frame_output_write_shm": (C.c_int, [_vp, C.c_uint64, C.POINTER(FrameOutputHeader), _vp]),
    "rv_frame_output_read_shm": (C.c_int, [_vp, C.c_uint64, C.POINTER(FrameOutputHeader), _vp]),
    "rv_controller_create": (C.c_int, [C.c_uint32, C.c_uint32, C.c_int, C.POINTER(C.c_void_p)]),
    "rv_controller_set_problem": (C.c_int, [_vp, C.c_char_p, C.c_uint32, C.c_char_p, C.c_char_p]),
    "rv_controller_field_count": (C.c_int, [_vp, _u32p]),
    "rv_controller_field_info": (C.c_int, [_vp, C.c_uint32, C.POINTER(FieldInfo)]),
    "rv_controller_time_frame_count": (C.c_int, [_vp, _u32p]),
    "rv_controller_set_input": (C.c_int, [_vp, C.POINTER(RenderState)]),
    "rv_controller_set_camera": (C.c_int, [_vp, _f32p]),
    "rv_controller_camera": (C.c_int, [_vp, _f32p]),
    "rv_controller_reset_camera": (C.c_int, [_vp]),
    "rv_controller_set_render_level": (C.c_int, [_vp, C.c_int32]),
    "rv_controller_start": (C.c_int, [_vp]),
    "rv_controller_frame_output": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_uint64), _u32p, _u32p]),
    "rv_controller_wait_idle": (C.c_int, [_vp, C.c_uint32, C.POINTER(C.c_int), C.POINTER(C.c_uint64)]),
    "rv_controller_status": (C.c_int, [_vp, C.c_char_p, C.c_uint64]),
    "rv_controller_destroy": (None, [_vp]),
}

_lib = None


def lib():
    """Load librevise_hip.so (built by revise_amd.build); raise if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RvError("librevise_hip.so is not built: run `python -m revise_amd.build` "
                          "(there is no CPU fallback)")
        # torch first: its HIP runtime (soname libamdhip64.so.7) then
        # satisfies the library's dependency, so the process holds one HIP
        # runtime.  Loaded the other way round, torch loads a second runtime of
        # its own, and the library's runtime finds no device.
        import torch  # noqa: F401
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            if not hasattr(L, name):
                continue
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


def check(rc):
    if rc != RV_OK:
        msg = lib().rv_last_error()
        raise RvError("revise_hip error %d: %s" % (rc, msg.decode() if msg else ""))
    return rc


def stream_ptr(stream=None):
    """hipStream_t of a torch stream (None -> torch's current stream)."""
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def ptr(t):
    return C.c_void_p(t.data_ptr())
