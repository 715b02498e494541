"""ctypes wrapper of the CPU oracle (liboracle.so) and, when it has been built
in this container, of the reference harness (_ref/libamref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker / CPU baseline.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libamref.so")

vp, i32, i64, u32, u64, f64, sz = (C.c_void_p, C.c_int32, C.c_int64, C.c_uint32, C.c_uint64,
                                   C.c_double, C.c_size_t)


class RenderCfg(C.Structure):
    _fields_ = [("width", i32), ("height", i32), ("nkeys", i32),
                ("bbox_x1", i32), ("bbox_y1", i32), ("bbox_x2", i32), ("bbox_y2", i32),
                ("motion", i32), ("fading", i32), ("show_blobs", i32), ("density", i32),
                ("feather", i32), ("keep_background", i32), ("blend_blobs", i32), ("pad_", i32),
                ("lag_perm", vp), ("slope_perm", vp)]


_ORACLE = {
    "or_point_distance": (u64, [u64, u64]),
    "or_chain_cost": (u64, [vp, sz, sz, C.POINTER(u64)]),
    "or_morph_asynch": (u64, [vp, sz, sz, sz, C.c_uint, C.POINTER(f64)]),
    "or_morph_visit": (u64, [vp, sz, sz, sz, C.POINTER(u64), C.POINTER(f64)]),
    "or_refine_rounds": (u64, [vp, sz, sz, sz, u64, u64, u64, C.POINTER(i64), C.POINTER(u64)]),
    "or_refine_pass": (u64, [vp, vp, vp, i32, i32, i32, u64, i32, u64, C.POINTER(i64), C.POINTER(u64)]),
    "or_refine_knn_pass": (u64, [vp, vp, vp, i32, i32, i32, u64, i32, u64, C.POINTER(i64), C.POINTER(u64)]),
    "or_perm_index": (u32, [u32, u32, u64]),
    "or_hash64": (u64, [u64]),
    "or_rgb_to_hsp": (u32, [u32]),
    "or_hsp_to_rgb": (u32, [u32]),
    "or_ingest": (None, [vp, sz, C.c_int, vp, vp]),
    "or_perlin_perm": (None, [C.c_uint, vp]),
    "or_perlin_noise3": (f64, [vp, f64, f64, f64]),
    "or_perlin_octave2": (f64, [vp, f64, f64, C.c_int]),
    "or_spline_eval": (None, [vp, vp, C.c_int, f64, vp, vp]),
    "or_segment": (C.c_int, [vp, C.c_int, C.c_int, vp]),
    "or_render_atoms": (None, [C.POINTER(RenderCfg), vp, vp, vp, vp, i32, vp, i32, vp, f64, vp]),
    "or_get_time": (f64, [sz, sz, C.c_int, sz]),
    "or_create_color": (u32, [f64, f64, f64, f64]),
    "or_distinct_color": (u32, [sz]),
    "or_fluid_step": (None, [vp, i32, i32, i32, i32, f64, C.c_uint]),
    "or_lap_min_cost": (i64, [vp, i32, vp]),
    "or_draw_fluid": (None, [vp, i32, f64, i32, i32, f64, vp, i32, i32, vp, vp]),
    "or_match": (i32, [i32, vp, vp, vp, vp, vp, vp, u64, u64, u64, C.c_int, vp, vp, vp, vp, C.POINTER(u64)]),
    "or_match_energy": (f64, [i32, i32, vp, vp, vp, vp, vp]),
    "or_match_brute": (f64, [i32, i32, vp, vp, vp, vp, vp]),
    "or_mlp_forward": (None, [vp, vp, i32, C.c_int, vp]),
    "or_mlp_grad": (f64, [vp, vp, vp, i32, C.c_int, vp]),
}

MLP_PARAMS = 50946

_REF = {
    "ref_rgb_to_hsp": (u32, [u32]),
    "ref_hsp_to_rgb": (u32, [u32]),
    "ref_perlin_perm": (None, [C.c_uint, vp]),
    "ref_perlin_octave2": (f64, [C.c_uint, f64, f64, C.c_int]),
    "ref_perlin_noise3": (f64, [C.c_uint, f64, f64, f64]),
    "ref_spline_eval": (None, [vp, vp, C.c_int, f64, vp, vp]),
    "ref_point_distance": (u64, [u64, u64]),
    "ref_morph_asynch": (f64, [vp, sz, sz, sz, C.c_uint]),
    "ref_fluid_step": (None, [vp, C.c_int, C.c_int, C.c_int, C.c_int, f64, C.c_uint]),
    "ref_morph_new": (vp, []),
    "ref_morph_free": (None, [vp]),
    "ref_morph_set_option": (None, [vp, C.c_char_p, f64]),
    "ref_morph_add_rgba": (None, [vp, sz, vp, C.c_uint, C.c_uint]),
    "ref_morph_set_resolution": (None, [vp, C.c_uint, C.c_uint]),
    "ref_morph_run": (C.c_uint, [vp, sz]),
    "ref_morph_run_until": (C.c_uint, [vp, C.c_uint, sz]),
    "ref_morph_next_state": (None, [vp]),
    "ref_morph_state": (C.c_uint, [vp]),
    "ref_morph_energy": (f64, [vp]),
    "ref_morph_frame_count": (sz, [vp]),
    "ref_morph_blob_count": (sz, [vp, sz]),
    "ref_morph_blob": (sz, [vp, sz, sz, C.POINTER(sz), vp, vp]),
    "ref_morph_chain_count": (sz, [vp]),
    "ref_morph_chain": (sz, [vp, sz, C.POINTER(sz), C.POINTER(sz), vp]),
    "ref_morph_put_chain": (None, [vp, sz, vp]),
    "ref_morph_get_time": (f64, [vp, sz, sz]),
    "ref_morph_render": (None, [vp, f64, vp]),
    "ref_morph_bbox": (None, [vp, vp]),
    "ref_morph_get_pixel": (u32, [vp, sz, C.c_uint, C.c_uint]),
    "ref_morph_rng_state": (u64, [vp]),
    "ref_morph_frame_means": (None, [vp, sz, vp]),
    "ref_morph_match_energy": (None, [vp, vp]),
    "ref_morph_particle_count": (sz, [vp]),
    "ref_fluid_frame": (C.c_int, [vp, f64, C.c_uint, vp, vp, vp]),
    "ref_morph_background": (None, [vp, f64, vp]),
}


def _bind(path, protos):
    lib = C.CDLL(path)
    for name, (res, args) in protos.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_o = None
_r = None


def oracle():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError(f"{ORACLE_LIB} missing: run `make -C oracle`")
        _o = _bind(ORACLE_LIB, _ORACLE)
    return _o


def have_ref():
    return os.path.exists(REF_LIB)


def ref():
    global _r
    if _r is None:
        _r = _bind(REF_LIB, _REF)
    return _r


def ptr(a):
    return a.ctypes.data_as(vp)


# ----------------------------------------------------------------- helpers
def pack_points(x, y, xf=0, yf=0, flags=0):
    x, y, xf, yf, flags = (np.asarray(v, np.uint64) for v in (x, y, xf, yf, flags))
    return (x & 0xFFFF) | ((y & 0xFFFF) << 16) | ((xf & 0xFF) << 32) | ((yf & 0xFF) << 40) | ((flags & 0xFF) << 48)


def chain_cost(pts_atom_major, h):
    pts = np.ascontiguousarray(pts_atom_major, np.uint64)
    hi = u64(0)
    lo = oracle().or_chain_cost(ptr(pts), pts.size // h, h, C.byref(hi))
    return (hi.value << 64) | lo


def perlin_perm(seed):
    p = np.zeros(512, np.int32)
    oracle().or_perlin_perm(seed, ptr(p))
    return p


def segment(frame):
    f = np.ascontiguousarray(frame, np.uint8)
    h, w = f.shape[:2]
    lab = np.zeros(h * w, np.int32)
    n = oracle().or_segment(ptr(f), w, h, ptr(lab))
    return lab, n


def refine_rounds(pts_atom_major, h, y, rounds, seed, stream):
    """Returns (points, accepted, gain, attempts)."""
    pts = np.ascontiguousarray(pts_atom_major, np.uint64).copy()
    g = i64(0)
    att = u64(0)
    acc = oracle().or_refine_rounds(ptr(pts), pts.size // h, h, y, rounds, seed, stream, C.byref(g),
                                    C.byref(att))
    return pts, acc, g.value, att.value


def refine_pass(pts_atom_major, offsets, widths, h, cap, pas, rounds, seed):
    pts = np.ascontiguousarray(pts_atom_major, np.uint64).copy()
    off = np.ascontiguousarray(offsets, np.int64)
    wid = np.ascontiguousarray(widths, np.int32)
    g = i64(0)
    att = u64(0)
    acc = oracle().or_refine_pass(ptr(pts), ptr(off), ptr(wid), len(wid), h, cap, pas, rounds, seed,
                                  C.byref(g), C.byref(att))
    return pts, acc, g.value, att.value


def refine_knn_pass(pts_atom_major, offsets, widths, h, cap, pas, rounds, seed):
    pts = np.ascontiguousarray(pts_atom_major, np.uint64).copy()
    off = np.ascontiguousarray(offsets, np.int64)
    wid = np.ascontiguousarray(widths, np.int32)
    g = i64(0)
    att = u64(0)
    acc = oracle().or_refine_knn_pass(ptr(pts), ptr(off), ptr(wid), len(wid), h, cap, pas, rounds, seed,
                                      C.byref(g), C.byref(att))
    return pts, acc, g.value, att.value


def ingest(words, hsp=1):
    w = np.ascontiguousarray(words, np.uint32)
    st = np.zeros_like(w)
    sh = np.zeros_like(w)
    oracle().or_ingest(ptr(w), w.size, hsp, ptr(st), ptr(sh))
    return st, sh


def lap_min_cost(cost):
    c = np.ascontiguousarray(cost, np.int64)
    n = c.shape[0]
    a = np.zeros(n, np.int32)
    v = oracle().or_lap_min_cost(ptr(c), n, ptr(a))
    return v, a
