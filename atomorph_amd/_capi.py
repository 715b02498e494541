"""ctypes binding of the C ABI in include/atomorph_hip.h.

Loads the in-tree libatomorph_hip.so built for gfx950.  There is no CPU
fallback: if the library is missing, loading raises, so nothing can silently
run a substitute path.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
HIP_LIB = os.path.join(LIB_DIR, "libatomorph_hip.so")
HOST_LIB = os.path.join(LIB_DIR, "libatomorph.so")

AMH_OK = 0

vp = C.c_void_p
i32, i64, u32, u64, f32, f64, sz = (C.c_int32, C.c_int64, C.c_uint32, C.c_uint64,
                                    C.c_float, C.c_double, C.c_size_t)


class BlobStat(C.Structure):
    _fields_ = [("count", i64), ("sum_x", i64), ("sum_y", i64), ("sum_c", i64 * 4),
                ("x1", i32), ("y1", i32), ("x2", i32), ("y2", i32), ("first", i64)]


class RefineStats(C.Structure):
    _fields_ = [("attempts", u64), ("accepted", u64), ("gain", i64), ("reserved", u64)]


class TuneEntry(C.Structure):
    _fields_ = [("tile", i32), ("rounds", i32), ("passes", i32), ("reserved", i32), ("ms", f64),
                ("gain_at_budget", f64), ("gain", i64)]


class ChainBuild(C.Structure):
    _fields_ = [("nkeys", i32), ("nchains", i32), ("wtot", i64), ("frame_width", i32), ("pad_", i32),
                ("chain_off_dev", vp), ("chain_w_dev", vp), ("chain_blob_dev", vp),
                ("empty_xy_dev", vp), ("pixel_list_dev", C.POINTER(vp)),
                ("blob_first_dev", C.POINTER(vp)), ("seed", u64)]


class RenderArgs(C.Structure):
    _fields_ = [("width", i32), ("height", i32), ("nkeys", i32), ("nblobs", i32),
                ("bbox_x1", i32), ("bbox_y1", i32), ("bbox_x2", i32), ("bbox_y2", i32),
                ("motion", i32), ("fading", i32), ("show_blobs", i32), ("density", i32),
                ("feather", i32), ("keep_background", i32), ("blend_blobs", i32), ("nchains", i32),
                ("wtot", i64), ("keys_dev", vp), ("pts_dev", vp), ("chain_off_dev", vp), ("chain_w_dev", vp),
                ("rank_dev", vp), ("blob_rgba_dev", vp), ("lag_perm_dev", vp), ("slope_perm_dev", vp)]


class FluidFrame(C.Structure):
    _fields_ = [("r", C.POINTER(RenderArgs)), ("rank_dev", vp), ("surface_dev", vp), ("blob_col_dev", vp),
                ("sparseness", f64), ("fluidsteps", i32), ("pad_", i32), ("atom_keys_dev", vp)]


# name -> (restype, argtypes); every entry point declared in include/*.h
PROTOTYPES = {
    "amh_version": (C.c_char_p, []),
    "amh_last_error": (C.c_char_p, []),
    "amh_device_info": (C.c_int, [C.POINTER(i32), C.POINTER(u64), C.c_char_p, sz]),
    "amh_malloc": (C.c_int, [C.POINTER(vp), sz]),
    "amh_free": (C.c_int, [vp]),
    "amh_memcpy_h2d": (C.c_int, [vp, vp, sz, vp]),
    "amh_memcpy_d2h": (C.c_int, [vp, vp, sz, vp]),
    "amh_memcpy_d2d": (C.c_int, [vp, vp, sz, vp]),
    "amh_memset": (C.c_int, [vp, C.c_int, sz, vp]),
    "amh_stream_create": (C.c_int, [C.POINTER(vp)]),
    "amh_stream_destroy": (C.c_int, [vp]),
    "amh_stream_sync": (C.c_int, [vp]),
    "amh_timer_begin": (C.c_int, [C.POINTER(vp), vp]),
    "amh_timer_end": (C.c_int, [vp, vp, C.POINTER(f32)]),
    "amh_timer_free": (C.c_int, [vp]),
    "amh_profile_enable": (C.c_int, [C.c_int]),
    "amh_profile_read": (C.c_int, [C.c_char_p, C.POINTER(u64), C.POINTER(f64)]),
    "amh_profile_reset": (C.c_int, []),
    "amh_ingest_frame": (C.c_int, [vp, i32, i32, i32, vp, vp, vp]),
    "amh_segment": (C.c_int, [vp, i32, i32, vp, vp]),
    "amh_blob_workspace_bytes": (sz, [i32, i32]),
    "amh_blob_count": (C.c_int, [vp, i32, i32, vp, C.POINTER(i32), vp]),
    "amh_blob_stats": (C.c_int, [vp, vp, i32, i32, vp, i32, vp, vp, vp]),
    "amh_group_workspace_bytes": (sz, [i32, i32]),
    "amh_group_pixels": (C.c_int, [vp, i32, i32, vp, sz, vp, vp, vp]),
    "amh_build_chains": (C.c_int, [C.POINTER(ChainBuild), vp, vp]),
    "amh_chain_cost": (C.c_int, [vp, i32, i64, vp, vp, i32, vp, vp]),
    "amh_refine_chain": (C.c_int, [vp, i32, i64, i64, i32, i32, u64, u64, u64, vp, vp]),
    "amh_refine_tiles": (C.c_int, [C.POINTER(i32), vp, vp, i32]),
    "amh_refine_items": (C.c_int, [vp, i32, i32, vp, i32, C.POINTER(i32)]),
    "amh_refine_batched": (C.c_int, [vp, i32, i64, vp, vp, vp, i32, i32, u64, i32, u64, vp, vp, vp]),
    "amh_refine_knn": (C.c_int, [vp, i32, i64, vp, vp, vp, i32, i32, u64, i32, u64, vp, vp, vp]),
    "amh_refine_graph_create": (C.c_int, [C.POINTER(vp), vp, i32, i64, vp, vp, vp, i32, i32, i32, i32, u64, vp,
                                          vp, vp]),
    "amh_graph_launch": (C.c_int, [vp, vp]),
    "amh_graph_destroy": (C.c_int, [vp]),
    "amh_fluid_ranks": (C.c_int, [vp, i32, i64, vp, vp, i32, u64, vp, vp]),
    "amh_fluid_atom_keys": (C.c_int, [vp, i32, i64, vp, vp]),
    "amh_refine_autotune": (C.c_int, [vp, vp, i32, i64, vp, vp, vp, i32, f64, u64, i32, C.POINTER(i32),
                                      C.POINTER(i32), vp, i32, C.POINTER(i32), vp]),
    "amh_rgb_to_hsp": (u32, [u32]),
    "amh_hsp_to_rgb": (u32, [u32]),
    "amh_rgb2hsp": (None, [f64, f64, f64, C.POINTER(f64), C.POINTER(f64), C.POINTER(f64)]),
    "amh_hsp2rgb": (None, [f64, f64, f64, C.POINTER(f64), C.POINTER(f64), C.POINTER(f64)]),
    "amh_render_workspace_bytes": (sz, [i32, i32, i64]),
    "amh_render_workspace_init": (C.c_int, [vp, sz, i32, i32, i64, vp]),
    "amh_render_frame": (C.c_int, [C.POINTER(RenderArgs), f64, vp, sz, vp, vp]),
    "amh_render_noalloc_max_bytes": (sz, []),
    "amh_render_frame_noalloc": (C.c_int, [C.POINTER(RenderArgs), f64, vp, vp]),
    "amh_render_graph_create": (C.c_int, [C.POINTER(vp), C.POINTER(RenderArgs), vp, sz, vp]),
    "amh_render_graph_frame": (C.c_int, [vp, f64, vp, vp]),
    "amh_render_graph_destroy": (C.c_int, [vp]),
    "amh_fluid_create": (C.c_int, [C.POINTER(vp), i32, i32, i32, i32, i32, i32, vp]),
    "amh_fluid_destroy": (C.c_int, [vp]),
    "amh_fluid_reset": (C.c_int, [vp, f64, vp]),
    "amh_fluid_set_particles": (C.c_int, [vp, vp, vp]),
    "amh_fluid_get_particles": (C.c_int, [vp, vp, vp]),
    "amh_fluid_step": (C.c_int, [vp, i32, f64, u64, vp]),
    "amh_fluid_render": (C.c_int, [vp, C.POINTER(FluidFrame), f64, vp, vp]),
    "amh_fluid_draw": (C.c_int, [vp, C.POINTER(FluidFrame), f64, vp, vp]),
    "amh_fluid_use_graphs": (C.c_int, [vp, i32]),
    "amh_mlp_create": (C.c_int, [C.POINTER(vp), i32, vp]),
    "amh_mlp_destroy": (C.c_int, [vp]),
    "amh_mlp_set_params": (C.c_int, [vp, vp, vp]),
    "amh_mlp_get_params": (C.c_int, [vp, vp, vp]),
    "amh_mlp_forward": (C.c_int, [vp, vp, i64, vp, vp]),
    "amh_mlp_gradients": (C.c_int, [vp, vp, vp, i32, vp, vp, vp]),
    "amh_mlp_adam_step": (C.c_int, [vp, f32, vp]),
    "amh_mlp_gemm_test": (C.c_int, [vp, vp, vp, i32, i32, i32, vp]),
    "amh_warm_workspace_bytes": (sz, [i64, i32, i32]),
    "amh_learn_features": (C.c_int, [C.POINTER(RenderArgs), i32, vp, vp, vp, sz, vp]),
    "amh_warm_start": (C.c_int, [vp, C.POINTER(RenderArgs), vp, vp, sz, vp]),
}

MLP_PARAMS = 50946


class AmhError(RuntimeError):
    pass


_lib = None


def load(path=HIP_LIB):
    """Load libatomorph_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    # One HIP runtime per process: torch ships its own libamdhip64.so.7 (same
    # soname as ROCm's).  Whichever is loaded first serves every library, and
    # torch only works with its own, so let torch load first when present.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(path):
        raise AmhError(f"{path} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(path)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc, what=""):
    if rc != AMH_OK:
        msg = load().amh_last_error().decode(errors="replace")
        raise AmhError(f"{what} failed ({rc}): {msg}")
    return rc
