"""GPU parity of ingest (HSP), segmentation, blob statistics, pixel grouping
and chain build, through the C ABI.  Integer work: bit-exact."""
import ctypes as C

import numpy as np
import pytest

from atomorph_amd import _capi as A
from atomorph_amd import synth
from oracle import oracle_py as op
from tests import gpu_util as G

pytestmark = pytest.mark.gpu


def test_ingest_hsp_exhaustive():
    """Every 24-bit colour through rgb_to_hsp and back (FP64 on the GPU with
    the reference's operation order) equals the oracle, which equals the
    compiled reference."""
    T = G.torch()
    lib = G.lib()
    words = (np.arange(1 << 24, dtype=np.uint32) | np.uint32(0xFF000000))
    words[::97] &= np.uint32(0x00FFFFFF)  # some absent pixels (alpha 0)
    st_w, sh_w = op.ingest(words, 1)
    src = G.to_dev(words)
    st = G.zeros(words.size, T.int32)
    sh = G.zeros(words.size, T.int32)
    G.log("ingest 4096x4096 hsp=1")
    G.check(lib.amh_ingest_frame(G.dptr(src), 4096, 4096, 1, G.dptr(st), G.dptr(sh), G.stream()), "ingest")
    assert (G.to_host(st, np.uint32) == st_w).all()
    assert (G.to_host(sh, np.uint32) == sh_w).all()
    G.log("ingest hsp=0")
    G.check(lib.amh_ingest_frame(G.dptr(src), 4096, 4096, 0, G.dptr(st), G.dptr(sh), G.stream()), "ingest")
    st_w, sh_w = op.ingest(words, 0)
    assert (G.to_host(st, np.uint32) == st_w).all() and (G.to_host(sh, np.uint32) == sh_w).all()


def _segment(frame):
    T = G.torch()
    lib = G.lib()
    h, w = frame.shape[:2]
    src = G.to_dev(synth.to_words(frame))
    lab = G.zeros(h * w, T.int32)
    G.check(lib.amh_segment(G.dptr(src), w, h, G.dptr(lab), G.stream()), "amh_segment")
    return src, lab


def _frames():
    rng = np.random.default_rng(5)
    out = []
    for size, p in ((1, 0.9), (5, 0.6), (63, 0.55), (64, 0.6), (65, 0.5), (130, 0.58), (257, 0.6)):
        img = np.zeros((size, size + 3, 4), np.uint8)
        img[..., 3] = (rng.uniform(0, 1, (size, size + 3)) < p) * 255
        img[..., 0] = rng.integers(0, 256, (size, size + 3))
        out.append((f"random{size}", img))
    # worst case for tile merging: one serpentine component across tiles
    s = np.zeros((96, 200, 4), np.uint8)
    s[::2, :, 3] = 255
    s[1::4, -1, 3] = 255
    s[3::4, 0, 3] = 255
    out.append(("serpentine", s))
    out.append(("empty", np.zeros((17, 33, 4), np.uint8)))
    full = np.zeros((70, 70, 4), np.uint8)
    full[..., 3] = 1
    out.append(("full", full))
    return out


@pytest.mark.parametrize("name,frame", _frames(), ids=[n for n, _ in _frames()])
def test_segment_matches_oracle(name, frame):
    G.log(f"segment {name} {frame.shape}")
    _, lab = _segment(frame)
    want, n = op.segment(frame)
    assert (G.to_host(lab, np.int32) == want).all()


def test_segment_reference_blobs(golden_cfg1):
    for f in range(2):
        G.log(f"segment cfg1 frame {f}")
        _, lab = _segment(golden_cfg1["frames"][f])
        assert (G.to_host(lab, np.int32) == golden_cfg1["labels"][f]).all()


def test_segment_blobs512_full_size():
    frames = synth.workload("blobs512")
    for f, fr in enumerate(frames):
        G.log(f"segment blobs512 frame {f}")
        _, lab = _segment(fr)
        want, n = op.segment(fr)
        assert (G.to_host(lab, np.int32) == want).all()
        assert n == 40


def _stats_and_groups(frame):
    T = G.torch()
    lib = G.lib()
    h, w = frame.shape[:2]
    src, lab = _segment(frame)
    stored = G.zeros(h * w, T.int32)
    G.check(lib.amh_ingest_frame(G.dptr(src), w, h, 1, G.dptr(stored), None, G.stream()), "ingest")
    ws = G.zeros(lib.amh_blob_workspace_bytes(w, h) // 4 + 1, T.int32)
    nb = C.c_int32(0)
    G.check(lib.amh_blob_count(G.dptr(lab), w, h, G.dptr(ws), C.byref(nb), G.stream()), "count")
    nb = nb.value
    bop = G.zeros(h * w, T.int32)
    st = T.zeros(max(nb, 1) * C.sizeof(A.BlobStat), dtype=T.uint8, device="cuda")
    G.check(lib.amh_blob_stats(G.dptr(lab), G.dptr(stored), w, h, G.dptr(ws), nb, G.dptr(bop), G.dptr(st),
                               G.stream()), "stats")
    gws_b = lib.amh_group_workspace_bytes(h * w, nb)
    gws = T.zeros(max(gws_b, 4), dtype=T.uint8, device="cuda")
    first = G.zeros(nb + 1, T.int64)
    plist = G.zeros(h * w, T.int32)
    G.check(lib.amh_group_pixels(G.dptr(bop), h * w, nb, G.dptr(gws), gws_b, G.dptr(first), G.dptr(plist),
                                 G.stream()), "group")
    stats = (A.BlobStat * max(nb, 1)).from_buffer_copy(st.cpu().numpy().tobytes())
    return nb, G.to_host(bop, np.int32), stats, G.to_host(first, np.int64), G.to_host(plist, np.int32), stored


@pytest.mark.parametrize("which", ["cfg1", "blobs512"])
def test_blob_stats_and_grouping(which, golden_cfg1):
    frames = golden_cfg1["frames"] if which == "cfg1" else synth.workload("blobs512")
    for f, fr in enumerate(frames):
        G.log(f"blob stats {which} frame {f}")
        h, w = fr.shape[:2]
        nb, bop, stats, first, plist, stored_d = _stats_and_groups(fr)
        lab, n = op.segment(fr)
        assert nb == n
        roots = np.sort(np.unique(lab[lab >= 0]))
        stored, _ = op.ingest(synth.to_words(fr), 1)
        for b, r in enumerate(roots):
            idx = np.nonzero(lab == r)[0]
            assert (bop[idx] == b).all()
            s = stats[b]
            assert s.count == idx.size and s.first == r
            assert s.sum_x == int((idx % w).sum()) and s.sum_y == int((idx // w).sum())
            for k in range(4):
                assert s.sum_c[k] == int(((stored[idx] >> (8 * k)) & 255).sum())
            assert (s.x1, s.y1, s.x2, s.y2) == (int((idx % w).min()), int((idx // w).min()),
                                                int((idx % w).max()), int((idx // w).max()))
            # grouped pixel list: the blob's pixels in row-major order
            assert (plist[first[b]:first[b + 1]] == idx).all()
        assert first[nb] == int((lab >= 0).sum())
        assert (bop[lab < 0] == -1).all()


def _group(bop_host, nb):
    T = G.torch()
    lib = G.lib()
    n = bop_host.size
    gws_b = lib.amh_group_workspace_bytes(n, nb)
    gws = T.zeros(max(gws_b, 4), dtype=T.uint8, device="cuda")
    bop = G.to_dev(np.ascontiguousarray(bop_host, np.int32))
    first = G.zeros(nb + 1, T.int64)
    plist = G.zeros(n, T.int32)
    G.check(lib.amh_group_pixels(G.dptr(bop), n, nb, G.dptr(gws), gws_b, G.dptr(first), G.dptr(plist), G.stream()),
            "group")
    return G.to_host(first, np.int64), G.to_host(plist, np.int32), gws_b


@pytest.mark.parametrize("case", ["checkerboard512", "few_blobs", "many_blobs", "sparse_many", "one_blob"])
def test_group_pixels_any_blob_count(case):
    """Stable grouping of pixels by blob for blob counts from 1 to a
    checkerboard's 131,072 single-pixel blobs: the per-(block, blob) table
    for few blobs, the radix sort beyond; both give the pixels of each blob
    in row-major order, and the workspace stays O(pixels + blobs)."""
    rng = np.random.default_rng(9)
    if case == "checkerboard512":
        yy, xx = np.mgrid[0:512, 0:512]
        on = ((xx + yy) % 2 == 0).ravel()
        bop = -np.ones(512 * 512, np.int32)
        bop[on] = np.arange(int(on.sum()))
        nb = int(on.sum())
    elif case == "few_blobs":
        nb = 300
        bop = rng.integers(-1, nb, 256 * 256).astype(np.int32)
    elif case == "many_blobs":
        nb = 5000
        bop = rng.integers(-1, nb, 256 * 256).astype(np.int32)
    elif case == "sparse_many":
        nb = 70000
        bop = np.where(rng.uniform(size=1024 * 512) < 0.6, rng.integers(0, nb, 1024 * 512), -1).astype(np.int32)
    else:
        nb = 1
        bop = np.where(rng.uniform(size=333 * 77) < 0.5, 0, -1).astype(np.int32)
    G.log(f"group pixels {case}: {bop.size} pixels, {nb} blobs")
    first, plist, ws = _group(bop, nb)
    counts = np.bincount(bop[bop >= 0], minlength=nb)
    assert np.array_equal(first, np.concatenate([[0], np.cumsum(counts)]))
    idx = np.flatnonzero(bop >= 0)
    want = idx[np.argsort(bop[idx], kind="stable")]
    assert np.array_equal(plist[:idx.size], want)
    assert ws <= 24 * bop.size + 16 * nb + (1 << 20)


def test_engine_detects_checkerboard_alpha():
    """A dithered (checkerboard) alpha channel: every present pixel is its
    own blob, 131,072 per 512^2 key frame, through am::morph's detection."""
    from atomorph_amd import morph as M
    yy, xx = np.mgrid[0:512, 0:512]
    frames = []
    for k in range(2):
        f = np.zeros((512, 512, 4), np.uint8)
        on = (xx + yy + k) % 2 == 0
        f[on] = (200, 40 + 100 * k, 90, 255)
        frames.append(f)
    m = M.Morph()
    m.set(seed=1)
    for i, f in enumerate(frames):
        m.add_frame_rgba(i, f)
    m.set_resolution(512, 512)
    G.log("checkerboard alpha: detection")
    assert m.run(1) == M.STATE_BLOB_UNIFICATION, m.error()
    for k in range(2):
        assert m.blob_count(k) == 512 * 512 // 2
        for b in (0, 1, 77777, 131071):
            s = m.blob(k, b)["surface"]
            assert len(s) == 1
