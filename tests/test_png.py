"""The command-line driver's PNG codec (CPU only): every colour type and bit
depth the PNG specification allows for non-interlaced images, all five
scanline filters, tRNS transparency, and RGBA8 encode -> decode round trips;
the driver's option parsing without touching a GPU."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from atomorph_amd import _capi as A
from tests import png_util as P

CLI = os.path.join(A.LIB_DIR, "atomorph_cli")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(A.HOST_LIB):
        pytest.fail("libatomorph.so not built: python -m atomorph_amd.build")
    L = C.CDLL(A.HOST_LIB)
    L.am_png_decode_file.argtypes = [C.c_char_p, C.c_void_p, C.c_size_t, C.POINTER(C.c_uint), C.POINTER(C.c_uint)]
    L.am_png_encode_file.argtypes = [C.c_char_p, C.c_void_p, C.c_uint, C.c_uint]
    L.am_png_error.restype = C.c_char_p
    return L


def _decode(lib, path):
    w, h = C.c_uint(), C.c_uint()
    assert lib.am_png_decode_file(str(path).encode(), None, 0, C.byref(w), C.byref(h)) == 0, lib.am_png_error()
    out = np.zeros((h.value, w.value, 4), np.uint8)
    assert lib.am_png_decode_file(str(path).encode(), out.ctypes.data_as(C.c_void_p), out.size, None, None) == 0
    return out


def _scale(v, depth):
    if depth == 16:
        return v >> 8
    return v * 255 // ((1 << depth) - 1)


ALL_FILTERS = (0, 1, 2, 3, 4)


@pytest.mark.parametrize("depth", [1, 2, 4, 8, 16])
def test_grey_with_colour_key(lib, tmp_path, depth):
    rng = np.random.default_rng(depth)
    s = rng.integers(0, 1 << depth, (9, 13, 1))
    key = int(s[0, 0, 0])
    P.write(tmp_path / "g.png", s, 0, depth, trns=key.to_bytes(2, "big"), filters=ALL_FILTERS)
    got = _decode(lib, tmp_path / "g.png")
    g = _scale(s[..., 0], depth)
    assert (got[..., 0] == g).all() and (got[..., 1] == g).all() and (got[..., 2] == g).all()
    assert (got[..., 3] == np.where(s[..., 0] == key, 0, 255)).all()


@pytest.mark.parametrize("depth", [8, 16])
def test_rgb_with_colour_key(lib, tmp_path, depth):
    rng = np.random.default_rng(10 + depth)
    s = rng.integers(0, 1 << depth, (7, 11, 3))
    s[3, 4] = s[0, 0]
    key = b"".join(int(v).to_bytes(2, "big") for v in s[0, 0])
    P.write(tmp_path / "c.png", s, 2, depth, trns=key, filters=ALL_FILTERS)
    got = _decode(lib, tmp_path / "c.png")
    assert (got[..., :3] == _scale(s, depth)).all()
    assert (got[..., 3] == np.where((s == s[0, 0]).all(axis=-1), 0, 255)).all()


@pytest.mark.parametrize("depth", [1, 2, 4, 8])
def test_palette_with_alpha_table(lib, tmp_path, depth):
    rng = np.random.default_rng(20 + depth)
    n = 1 << depth
    pal = rng.integers(0, 256, (n, 3))
    alpha = rng.integers(0, 256, n // 2 + 1)      # shorter than the palette: the rest are opaque
    idx = rng.integers(0, n, (10, 17, 1))
    P.write(tmp_path / "p.png", idx, 3, depth, palette=pal, trns=bytes(int(a) for a in alpha), filters=ALL_FILTERS)
    got = _decode(lib, tmp_path / "p.png")
    i = idx[..., 0]
    assert (got[..., :3] == pal[i]).all()
    full = np.concatenate([alpha, np.full(n - len(alpha), 255)])
    assert (got[..., 3] == full[i]).all()


@pytest.mark.parametrize("ctype,depth", [(4, 8), (4, 16), (6, 8), (6, 16)])
def test_alpha_types(lib, tmp_path, ctype, depth):
    rng = np.random.default_rng(30 + ctype + depth)
    s = rng.integers(0, 1 << depth, (8, 12, P.CHANNELS[ctype]))
    P.write(tmp_path / "a.png", s, ctype, depth, filters=ALL_FILTERS)
    got = _decode(lib, tmp_path / "a.png")
    v = _scale(s, depth)
    if ctype == 4:
        assert (got[..., 0] == v[..., 0]).all() and (got[..., 2] == v[..., 0]).all()
        assert (got[..., 3] == v[..., 1]).all()
    else:
        assert (got == v).all()


def test_encode_round_trip(lib, tmp_path):
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, (37, 53, 4), dtype=np.uint8)
    path = str(tmp_path / "o.png").encode()
    assert lib.am_png_encode_file(path, img.ctypes.data_as(C.c_void_p), 53, 37) == 0
    assert (P.read_rgba8(tmp_path / "o.png") == img).all()
    assert (_decode(lib, tmp_path / "o.png") == img).all()


def test_corrupt_files_are_rejected(lib, tmp_path):
    img = np.zeros((4, 4, 4), np.uint8)
    P.write_rgba8(tmp_path / "ok.png", img)
    data = bytearray(open(tmp_path / "ok.png", "rb").read())
    data[40] ^= 0xFF                                   # inside IDAT: CRC no longer matches
    (tmp_path / "bad.png").write_bytes(bytes(data))
    w, h = C.c_uint(), C.c_uint()
    assert lib.am_png_decode_file(str(tmp_path / "bad.png").encode(), None, 0, C.byref(w), C.byref(h)) == -1
    assert b"CRC" in lib.am_png_error()
    (tmp_path / "not.png").write_bytes(b"GIF89a")
    assert lib.am_png_decode_file(str(tmp_path / "not.png").encode(), None, 0, C.byref(w), C.byref(h)) == -1


def test_cli_help_and_missing_input():
    if not os.path.exists(CLI):
        pytest.fail("atomorph_cli not built: python -m atomorph_amd.build")
    r = subprocess.run([CLI, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "--morph-time" in r.stdout and "--learned" in r.stdout
    r = subprocess.run([CLI, "-i", "/nonexistent", "a.png"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "cannot open" in r.stderr
