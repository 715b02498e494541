"""CPU checks of the drop-in boundary: the C-ABI library loads and exports
every entry point include/*.h declares, the ctypes bindings cover them, and
the host-only parts of am::morph behave like the reference (no GPU calls)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from atomorph_amd import _capi as A
from oracle import oracle_py as op

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"\b(amh_[a-z0-9_]+)\s*\(", text))
    return names


def test_header_declares_entry_points():
    names = _declared()
    assert len(names) > 30
    for must in ("amh_segment", "amh_refine_chain", "amh_refine_batched", "amh_render_frame", "amh_chain_cost"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(A.HIP_LIB)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_prototypes_cover_the_header():
    assert _declared() <= set(A.PROTOTYPES), sorted(_declared() - set(A.PROTOTYPES))
    A.load()  # binds every prototype (raises on a missing symbol)


def test_host_library_exports_the_morph_shim():
    from atomorph_amd import morph as M
    lib = M.load()
    for name in M._PROTOS:
        assert hasattr(lib, name), name


def test_error_path_without_device_work():
    lib = A.load()
    # invalid arguments are rejected before any HIP call, with a message
    rc = lib.amh_segment(None, 0, 0, None, None)
    assert rc == -1 and b"invalid" in lib.amh_last_error()
    rc = lib.amh_refine_chain(None, 2, 10, 0, 5, 0, 1, 0, 0, None, None)
    assert rc == -1
    g = C.c_void_p()
    assert lib.amh_render_graph_create(C.byref(g), None, None, 0, None) == -1 and not g.value
    assert lib.amh_render_graph_frame(None, 0.5, None, None) == -1
    assert lib.amh_render_graph_destroy(None) == 0


@pytest.mark.parametrize("gsx,gsy,w,h", [(65535, 64, 64, 64), (64, 70000, 64, 64), (64, 64, 65536, 8),
                                         (64, 64, 8, 100000), (7, 64, 64, 64)])
def test_fluid_create_rejects_sizes_the_packing_cannot_hold(gsx, gsy, w, h):
    """Tap pixels pack x | y << 16 (0xFFFFFFFF = no tap) and cell keys
    (cx << 16) | cy, so sizes of 65535 and more are refused before any
    device allocation."""
    lib = A.load()
    f = C.c_void_p()
    rc = lib.amh_fluid_create(C.byref(f), gsx, gsy, 100, w, h, 2, None)
    assert rc == -1 and b"invalid" in lib.amh_last_error()
    assert f.value is None


def test_host_colour_matches_oracle():
    lib = A.load()
    rng = np.random.default_rng(0)
    o = op.oracle()
    for c in rng.integers(0, 1 << 32, 5000, dtype=np.uint64):
        c = int(c)
        assert lib.amh_rgb_to_hsp(c) == o.or_rgb_to_hsp(c)
        assert lib.amh_hsp_to_rgb(c) == o.or_hsp_to_rgb(c)


# --------------------------------------------- am::morph host-only members
@pytest.fixture()
def morph():
    from atomorph_amd import morph as M
    return M.Morph()


def test_morph_time_and_frame_keys(morph):
    lib, h = morph.lib, morph.h
    for k in (0, 3, 7):
        assert lib.am_morph_add_frame(h, k) == 0
    assert lib.am_morph_add_frame(h, 3) == -1  # already there (morph.cpp:288-296)
    assert lib.am_morph_frame_count(h) == 3
    assert lib.am_morph_frame_key(h, 0.0) == 0
    assert lib.am_morph_frame_key(h, 0.5) == 3
    assert lib.am_morph_frame_key(h, 0.9) == 7
    assert lib.am_morph_frame_key(h, -0.1) == 7
    assert lib.am_morph_normalize_time(h, -0.25) == 0.75
    assert morph.get_time(3, 16) == 3 / 16
    morph.set(finite=1)
    assert morph.get_time(15, 16) == 1.0 * (1 - 1 / 3)


def test_morph_pixels_are_hsp_round_tripped(morph):
    lib, h = morph.lib, morph.h
    o = op.oracle()
    c = 0xFF3366CC
    assert lib.am_morph_add_pixel(h, 0, 5, 7, c) == 0
    got = lib.am_morph_get_pixel(h, 0, 7 * 65536 + 5)
    assert got == o.or_hsp_to_rgb(o.or_rgb_to_hsp(c))
    assert lib.am_morph_get_pixel(h, 0, 7 * 65536 + 6) == 0  # void is transparent
    morph.set(blob_delimiter=0)
    assert lib.am_morph_get_pixel(h, 0, 7 * 65536 + 5) == c
    assert lib.am_morph_pixel_count(h, 0) == 1


def test_morph_interpolation_matches_reference_formulas(morph):
    lib, h = morph.lib, morph.h
    o = op.oracle()
    rng = np.random.default_rng(1)
    # the oracle's renderer uses the same fade; compare on random colours
    for _ in range(500):
        c1, c2 = (int(v) for v in rng.integers(0, 1 << 32, 2, dtype=np.uint64))
        lag, slope, w = rng.uniform(0, 1, 3)
        got = lib.am_morph_interpolate_color(h, c1, c2, lag, slope, w)
        s = (slope + 0.1) / 1.1
        l = (1.0 - s) * lag
        str_ = 0.0 if w <= l else (1.0 if w >= l + s else (-np.cos((w - l) * (3.14159265358 / s)) + 1.0) / 2.0)
        want = 0
        for k in range(4):
            a, b = (c1 >> (8 * k)) & 255, (c2 >> (8 * k)) & 255
            v = str_ * a + (1.0 - str_) * b
            want |= (int(np.floor(v + 0.5)) & 255) << (8 * k)
        assert abs(int(got) - int(want)) == 0 or all(
            abs(((got >> (8 * k)) & 255) - ((want >> (8 * k)) & 255)) <= 1 for k in range(4))
    # point interpolation truncates (morph.cpp:1501-1515)
    p1 = int(op.pack_points(10, 20, 0, 128, 0))
    p2 = int(op.pack_points(11, 21, 128, 0, 0))
    r = lib.am_morph_interpolate_point(h, p1, p2, 0.5)
    assert (r & 0xFFFF, (r >> 16) & 0xFFFF, (r >> 32) & 255, (r >> 40) & 255) == (10, 20, 192, 192)


def test_density_and_fluid_coupling(morph):
    """set_fluid(>0) forces density 1; set_density(>1) forces fluid 0
    (morph.cpp:1553-1570).  Observable through the restart identifier: a
    density change restarts, a fluid change alone does not."""
    morph.set(fluid=10)
    morph.set(density=2)  # turns fluid off
    morph.set(fluid=5)    # turns density back to 1


def test_group_workspace_is_linear_in_pixels_and_blobs():
    """Grouping pixels by blob needs O(pixels + blobs) workspace for any blob
    count (a dithered 2048^2 frame has about two million blobs)."""
    lib = A.load()
    for npx, nb in ((1024 * 1024, 150), (1024 * 1024, 524288), (2048 * 2048, 2097152), (2048 * 2048, 1200),
                    (64 * 64, 1), (1000, 999)):
        assert lib.amh_group_workspace_bytes(npx, nb) <= 24 * npx + 16 * nb + (1 << 20), (npx, nb)
