"""ctypes binding of am::morph (libatomorph.so, C shim in host/morph_c.cpp).

    m = Morph()
    m.set(seed=1, cycle_length=1 << 16)
    for i, frame in enumerate(frames):
        m.add_frame_rgba(i, frame)
    m.set_resolution(w, h)
    m.run(1)                # blob detection
    ...
    img = m.render(m.get_time(i, n))
"""
import ctypes as C
import os

import numpy as np

from . import _capi

STATE_BLOB_DETECTION, STATE_BLOB_UNIFICATION, STATE_BLOB_MATCHING, STATE_ATOM_MORPHING, STATE_DONE = range(5)
REFINE_PLAIN, REFINE_BATCHED, REFINE_KNN = 0, 1, 2

vp, sz, u64, i64 = C.c_void_p, C.c_size_t, C.c_uint64, C.c_int64

_PROTOS = {
    "am_morph_new": (vp, []),
    "am_morph_free": (None, [vp]),
    "am_morph_set": (C.c_int, [vp, C.c_char_p, C.c_double]),
    "am_morph_set_refine": (None, [vp, C.c_uint, C.c_uint, C.c_int, C.c_uint]),
    "am_morph_add_frame_rgba": (C.c_int, [vp, sz, vp, C.c_uint, C.c_uint, sz]),
    "am_morph_add_frame": (C.c_int, [vp, sz]),
    "am_morph_set_resolution": (None, [vp, C.c_uint, C.c_uint]),
    "am_morph_run": (C.c_int, [vp, sz]),
    "am_morph_compute": (None, [vp]),
    "am_morph_compute_seconds": (None, [vp, C.c_double]),
    "am_morph_iterate": (None, [vp, sz]),
    "am_morph_suspend": (None, [vp]),
    "am_morph_wait": (None, [vp]),
    "am_morph_restart": (None, [vp]),
    "am_morph_run_timed": (C.c_int, [vp, sz, C.POINTER(C.c_float)]),
    "am_morph_render_device": (C.c_int, [vp, vp, sz, vp, C.POINTER(C.c_float)]),
    "am_morph_suspend_timeout": (C.c_int, [vp, C.c_double]),
    "am_morph_is_busy": (C.c_int, [vp]),
    "am_morph_synchronize": (C.c_int, [vp]),
    "am_morph_next_state": (None, [vp]),
    "am_morph_state": (C.c_uint, [vp]),
    "am_morph_energy": (C.c_double, [vp]),
    "am_morph_cost": (C.c_int, [vp, C.POINTER(u64), C.POINTER(u64)]),
    "am_morph_counters": (None, [vp, C.POINTER(u64), C.POINTER(u64)]),
    "am_morph_fluid_particles": (C.c_size_t, [vp, vp]),
    "am_morph_set_learned": (None, [vp, vp]),
    "am_morph_warm_start": (C.c_int, [vp, C.POINTER(C.c_float)]),
    "am_morph_save_chains": (C.c_int, [vp, C.c_char_p]),
    "am_morph_load_chains": (C.c_int, [vp, C.c_char_p]),
    "am_morph_autotune": (C.c_int, [vp, C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_uint), vp, sz,
                                    C.POINTER(sz)]),
    "am_morph_learn_samples": (C.c_size_t, [vp, vp, vp]),
    "am_morph_get_time": (C.c_double, [vp, sz, sz]),
    "am_morph_render": (C.c_int, [vp, C.c_double, vp]),
    "am_morph_render_pixels": (C.c_int, [vp, C.c_double, vp, C.c_uint, C.c_uint]),
    "am_morph_blob_count": (sz, [vp, sz]),
    "am_morph_blob": (C.c_longlong, [vp, sz, sz, C.POINTER(sz), vp, vp]),
    "am_morph_chains": (sz, [vp, C.POINTER(sz), vp, vp]),
    "am_morph_chain_count": (sz, [vp]),
    "am_morph_put_chains": (C.c_int, [vp, vp, sz]),
    "am_morph_error": (C.c_char_p, [vp]),
    "am_morph_frame_key": (sz, [vp, C.c_double]),
    "am_morph_normalize_time": (C.c_double, [vp, C.c_double]),
    "am_morph_get_pixel": (C.c_uint32, [vp, sz, sz]),
    "am_morph_add_pixel": (C.c_int, [vp, sz, C.c_uint, C.c_uint, C.c_uint32]),
    "am_morph_interpolate_color": (C.c_uint32, [vp, C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double]),
    "am_morph_interpolate_point": (u64, [vp, u64, u64, C.c_double]),
    "am_morph_background": (C.c_uint32, [vp, C.c_uint, C.c_uint, C.c_double]),
    "am_morph_frame_count": (sz, [vp]),
    "am_morph_pixel_count": (sz, [vp, sz]),
    "am_blobify_frames": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.c_double, u64, u64, u64, u64,
                                    C.c_uint32, C.c_uint, vp, vp, vp, vp, vp]),
}

_lib = None


def load():
    global _lib
    if _lib is None:
        _capi.load()  # the HIP library first (it is a dependency)
        if not os.path.exists(_capi.HOST_LIB):
            raise _capi.AmhError(f"{_capi.HOST_LIB} is missing: build it first")
        lib = C.CDLL(_capi.HOST_LIB)
        for name, (res, args) in _PROTOS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


SIZE_MAX = (1 << 64) - 1


def blobify_frames(stored, threshold=1.0, max_size=SIZE_MAX, min_size=1, number=1, box_samples=10,
                   box_grip=65535, seed=0):
    """Greedy blob detection and dust unification on the host
    (thread.cpp:225-596), as the engine runs them for non-default blob
    options.

    stored: (frames, H, W) uint32 RGBA words of the key frames as the morph
    stores them (after the colour conversion); a pixel is present when its
    alpha is non-zero.  Returns (bop, blobs): bop (frames, H*W) int32, the
    blob ordinal of each pixel (-1: no pixel); blobs[f], frame f's blobs in
    list order as dicts of means (x, y, r, g, b, a), group (y*65536 + x of
    the pixel it grew from) and size."""
    lib = load()
    st = np.ascontiguousarray(stored, np.uint32)
    nf, H, W = st.shape
    npx = H * W
    bop = np.zeros((nf, npx), np.int32)
    nb = np.zeros(nf, np.int32)
    means = np.zeros((nf, npx, 6))
    groups = np.zeros((nf, npx), np.uint64)
    sizes = np.zeros((nf, npx), np.uint64)
    rc = lib.am_blobify_frames(st.ctypes.data_as(vp), nf, W, H, float(threshold), int(max_size), int(min_size),
                               int(number), int(box_samples), int(box_grip), int(seed), bop.ctypes.data_as(vp),
                               nb.ctypes.data_as(vp), means.ctypes.data_as(vp), groups.ctypes.data_as(vp),
                               sizes.ctypes.data_as(vp))
    if rc != 0:
        raise ValueError("am_blobify_frames: invalid argument")
    blobs = [[{"means": means[f, b].copy(), "group": int(groups[f, b]), "size": int(sizes[f, b])}
              for b in range(nb[f])] for f in range(nf)]
    return bop, blobs


class Morph:
    def __init__(self):
        self.lib = load()
        self.h = self.lib.am_morph_new()
        if not self.h:
            raise MemoryError("am_morph_new")
        self.width = self.height = 0

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.am_morph_free(self.h)
            self.h = None

    def error(self):
        return self.lib.am_morph_error(self.h).decode(errors="replace")

    def set(self, **opts):
        for k, v in opts.items():
            if self.lib.am_morph_set(self.h, k.encode(), float(v)) != 0:
                raise KeyError(k)

    def set_refine(self, mode, rounds=8, tile=2, graph_passes=0):
        self.lib.am_morph_set_refine(self.h, mode, rounds, tile, graph_passes)

    def add_frame_rgba(self, key, frame):
        f = np.ascontiguousarray(frame, np.uint8)
        hh, ww = f.shape[:2]
        if self.lib.am_morph_add_frame_rgba(self.h, key, f.ctypes.data_as(vp), ww, hh, ww * 4) != 0:
            raise ValueError("add_frame_rgba")

    def set_resolution(self, w, h):
        self.width, self.height = w, h
        self.lib.am_morph_set_resolution(self.h, w, h)

    def run(self, n):
        st = self.lib.am_morph_run(self.h, n)
        if st < 0:
            raise _capi.AmhError(self.error())
        return st

    def restart(self):
        """Back to blob detection at the next run (frames and options kept)."""
        self.lib.am_morph_restart(self.h)

    def run_timed(self, n):
        """n steps; returns (state, device milliseconds of those steps)."""
        ms = C.c_float()
        st = self.lib.am_morph_run_timed(self.h, n, C.byref(ms))
        if st < 0:
            raise _capi.AmhError(self.error())
        return st, ms.value

    def render_device(self, times, out_dev_ptr):
        """Render len(times) frames into a device buffer; returns device ms."""
        ts = np.ascontiguousarray(times, np.float64)
        ms = C.c_float()
        if self.lib.am_morph_render_device(self.h, ts.ctypes.data_as(vp), ts.size, C.c_void_p(out_dev_ptr),
                                           C.byref(ms)) != 0:
            raise _capi.AmhError(self.error())
        return ms.value

    def next_state(self):
        self.lib.am_morph_next_state(self.h)

    def state(self):
        return self.lib.am_morph_state(self.h)

    def energy(self):
        return self.lib.am_morph_energy(self.h)

    def cost(self):
        lo, hi = u64(), u64()
        if self.lib.am_morph_cost(self.h, C.byref(lo), C.byref(hi)) != 0:
            raise _capi.AmhError(self.error())
        return (hi.value << 64) | lo.value

    def counters(self):
        a, b = u64(), u64()
        self.lib.am_morph_counters(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def set_learned(self, params):
        """Learned warm start on (params: AMH_MLP_PARAMS float32) or off (None)."""
        if params is None:
            self.lib.am_morph_set_learned(self.h, None)
            self._mlp = None
            return
        self._mlp = np.ascontiguousarray(params, np.float32)
        if self._mlp.size != _capi.MLP_PARAMS:
            raise ValueError(f"expected {_capi.MLP_PARAMS} parameters, got {self._mlp.size}")
        self.lib.am_morph_set_learned(self.h, self._mlp.ctypes.data_as(vp))

    def warm_start(self):
        """Apply the learned warm start to the current chains; device ms."""
        ms = C.c_float()
        if self.lib.am_morph_warm_start(self.h, C.byref(ms)) != 0:
            raise _capi.AmhError(self.error())
        return ms.value

    def autotune(self, budget_ms=3.0, rounds=0):
        """Autotune the batched descent on the current chains (which it does
        not change): returns (tile, rounds, report), report one dict per
        candidate with its passes, device ms and exact gain at the budget.
        rounds > 0: only the tile is tuned, for that many rounds per pass."""
        t, r, n = C.c_int(), C.c_uint(rounds), sz()
        rep = np.zeros((64, 5), np.float64)
        if self.lib.am_morph_autotune(self.h, budget_ms, C.byref(t), C.byref(r), rep.ctypes.data_as(vp), 64,
                                      C.byref(n)) != 0:
            raise _capi.AmhError(self.error())
        rows = [{"tile": int(a), "rounds": int(b), "passes": int(c), "ms": float(d), "gain_at_budget": float(e)}
                for a, b, c, d, e in rep[:min(n.value, 64)]]
        return t.value, r.value, rows

    def save_chains(self, path):
        """Checkpoint the synchronized chains to a file."""
        if self.lib.am_morph_save_chains(self.h, str(path).encode()) != 0:
            raise _capi.AmhError(self.error())

    def load_chains(self, path):
        """Resume from a checkpoint (paused in ATOM_MORPHING, same chain shapes)."""
        if self.lib.am_morph_load_chains(self.h, str(path).encode()) != 0:
            raise _capi.AmhError(self.error())

    def learn_samples(self):
        """(features (n, 8), targets (n, 2)) float32 from the synchronized chains."""
        n = self.lib.am_morph_learn_samples(self.h, None, None)
        x = np.zeros((n, 8), np.float32)
        t = np.zeros((n, 2), np.float32)
        if n and self.lib.am_morph_learn_samples(self.h, x.ctypes.data_as(vp), t.ctypes.data_as(vp)) != n:
            raise _capi.AmhError(self.error())
        return x, t

    def fluid_particles(self):
        """Particles of the last fluid frame, (count, 20) FP64 (see morph.h)."""
        import numpy as np
        n = self.lib.am_morph_fluid_particles(self.h, None)
        out = np.zeros((n, 20))
        if n:
            self.lib.am_morph_fluid_particles(self.h, out.ctypes.data)
        return out

    def get_time(self, i, n):
        return self.lib.am_morph_get_time(self.h, i, n)

    def render(self, t):
        out = np.zeros((self.height, self.width, 4), np.uint8)
        if self.lib.am_morph_render(self.h, float(t), out.ctypes.data_as(vp)) != 0:
            raise _capi.AmhError(self.error())
        return out

    def render_pixels(self, t):
        out = np.zeros((self.height, self.width, 4), np.uint8)
        self.lib.am_morph_render_pixels(self.h, float(t), out.ctypes.data_as(vp), self.width, self.height)
        return out

    def blob_count(self, f):
        return self.lib.am_morph_blob_count(self.h, f)

    def blob(self, f, b):
        g = sz()
        means = np.zeros(6)
        n = self.lib.am_morph_blob(self.h, f, b, C.byref(g), means.ctypes.data_as(vp), None)
        if n < 0:
            return None
        surf = np.zeros(max(n, 1), np.uint64)
        self.lib.am_morph_blob(self.h, f, b, C.byref(g), None, surf.ctypes.data_as(vp))
        return {"group": g.value, "means": means, "surface": surf[:n]}

    def chains(self):
        """(points key-frame-major (h, atoms) uint64, widths int64)."""
        nch = self.lib.am_morph_chain_count(self.h)
        hh = sz()
        widths = np.zeros(max(nch, 1), np.int64)
        n = self.lib.am_morph_chains(self.h, C.byref(hh), widths.ctypes.data_as(vp), None)
        pts = np.zeros(max(n * hh.value, 1), np.uint64)
        self.lib.am_morph_chains(self.h, C.byref(hh), widths.ctypes.data_as(vp), pts.ctypes.data_as(vp))
        return pts[:n * hh.value].reshape(hh.value, n), widths[:nch]

    def put_chains(self, pts):
        p = np.ascontiguousarray(pts, np.uint64).ravel()
        if self.lib.am_morph_put_chains(self.h, p.ctypes.data_as(vp), p.size) != 0:
            raise _capi.AmhError("put_chains: " + self.error())
