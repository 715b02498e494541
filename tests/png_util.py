"""A small PNG writer/reader for the tests (zlib + the PNG scanline filters),
independent of the library's C++ codec it checks."""
import struct
import zlib

import numpy as np

CHANNELS = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}


def _chunk(kind, body):
    return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body) & 0xFFFFFFFF)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def _filter_row(kind, raw, prev, bpp):
    out = bytearray(len(raw))
    for i, v in enumerate(raw):
        left = raw[i - bpp] if i >= bpp else 0
        up = prev[i]
        ul = prev[i - bpp] if i >= bpp else 0
        pred = (0, left, up, (left + up) >> 1, _paeth(left, up, ul))[kind]
        out[i] = (v - pred) & 0xFF
    return bytes(out)


def _pack_row(samples, depth):
    if depth == 8:
        return bytes(int(s) for s in samples)
    if depth == 16:
        return b"".join(struct.pack(">H", int(s)) for s in samples)
    per = 8 // depth
    out = bytearray((len(samples) * depth + 7) // 8)
    for k, s in enumerate(samples):
        out[k // per] |= (int(s) & ((1 << depth) - 1)) << (8 - depth * (k % per + 1))
    return bytes(out)


def write(path, samples, ctype, depth, palette=None, trns=None, filters=(0,)):
    """samples: (H, W, channels) integer array of raw samples at `depth`
    bits.  Rows use the filter types in `filters`, cyclically."""
    s = np.asarray(samples)
    h, w = s.shape[:2]
    ch = CHANNELS[ctype]
    s = s.reshape(h, w, ch)
    bpp = max(1, ch * depth // 8)
    rows, prev = [], None
    for y in range(h):
        raw = _pack_row(s[y].reshape(-1), depth)
        if prev is None:
            prev = bytes(len(raw))
        kind = filters[y % len(filters)]
        rows.append(bytes([kind]) + _filter_row(kind, raw, prev, bpp))
        prev = raw
    out = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 0))
    if palette is not None:
        out += _chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).reshape(-1)))
    if trns is not None:
        out += _chunk(b"tRNS", bytes(trns))
    out += _chunk(b"IDAT", zlib.compress(b"".join(rows), 9)) + _chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(out)


def write_rgba8(path, rgba, filters=(0,)):
    write(path, rgba, 6, 8, filters=filters)


def read_rgba8(path):
    """Reads an 8-bit RGBA PNG (what the library writes)."""
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w, h = 8, b"", 0, 0
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        kind, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        assert zlib.crc32(kind + body) & 0xFFFFFFFF == struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]
        if kind == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            assert (depth, ctype) == (8, 6)
        elif kind == b"IDAT":
            idat += body
        pos += 12 + n
    raw = zlib.decompress(idat)
    stride = 4 * w + 1
    out = np.zeros((h, 4 * w), np.uint8)
    prev = bytes(4 * w)
    for y in range(h):
        kind, line = raw[y * stride], bytearray(raw[y * stride + 1:(y + 1) * stride])
        for i in range(len(line)):
            left = line[i - 4] if i >= 4 else 0
            up = prev[i]
            ul = prev[i - 4] if i >= 4 else 0
            line[i] = (line[i] + (0, left, up, (left + up) >> 1, _paeth(left, up, ul))[kind]) & 0xFF
        out[y] = np.frombuffer(bytes(line), np.uint8)
        prev = bytes(line)
    return out.reshape(h, w, 4)
