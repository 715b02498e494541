"""End-to-end am::morph on the GPU: the reference's driver sequence
(set options, add frames, iterate states, render) through the C++ engine."""
import ctypes as C

import numpy as np
import pytest

from atomorph_amd import _capi as A
from atomorph_amd import morph as M
from atomorph_amd import synth
from oracle import oracle_py as op
from tests import gpu_util as G

pytestmark = pytest.mark.gpu


def _pipeline(frames, seed=1, cycle=1 << 16, match_steps=2000, **opts):
    m = M.Morph()
    m.set(seed=seed, cycle_length=cycle, **opts)
    for i, f in enumerate(frames):
        m.add_frame_rgba(i, f)
    hh, ww = frames[0].shape[:2]
    m.set_resolution(ww, hh)
    G.log("morph: detection")
    assert m.run(1) == M.STATE_BLOB_UNIFICATION
    G.log("morph: matching")
    st = m.run(match_steps + 1)
    assert st == M.STATE_BLOB_MATCHING
    m.next_state()
    G.log("morph: chain build")
    assert m.run(1) == M.STATE_ATOM_MORPHING
    return m


def test_detection_matches_reference_blobs(golden_cfg1):
    g = golden_cfg1
    m = M.Morph()
    m.set(seed=1)
    for i, f in enumerate(g["frames"]):
        m.add_frame_rgba(i, f)
    m.set_resolution(64, 64)
    G.log("cfg1 detection")
    assert m.run(1) == M.STATE_BLOB_UNIFICATION
    for f in range(2):
        n = m.blob_count(f)
        comps = set()
        for b in range(n):
            s = m.blob(f, b)["surface"]
            idx = (s // 65536) * 64 + (s % 65536)
            comps.add(tuple(sorted(idx.tolist())))
        lab = g["labels"][f]
        want = set()
        for r in np.unique(lab[lab >= 0]):
            want.add(tuple(np.nonzero(lab == r)[0].tolist()))
        assert comps == want


def test_chains_follow_init_morph(golden_cfg1):
    """Row y of every chain: the blob's pixels in row-major order (HAS_PIXEL |
    HAS_FLUID, zero fractions), then duplicates (HAS_PIXEL only) of those
    pixels; an empty blob gives one HAS_FLUID point and flag-less copies."""
    m = _pipeline(list(golden_cfg1["frames"]))
    pts, widths = m.chains()
    h = pts.shape[0]
    off = np.concatenate([[0], np.cumsum(widths)[:-1]])
    assert widths.sum() == pts.shape[1]
    for f in range(h):
        groups = {}
        for b in range(m.blob_count(f)):
            bl = m.blob(f, b)
            groups[bl["group"]] = bl
        for c, (o, w) in enumerate(zip(off, widths)):
            row = pts[f, o:o + w]
            surf = groups[c]["surface"]
            fl = (row >> 48) & 255
            xy = (row & 0xFFFFFFFF)
            pos = (xy >> 16) * 65536 + (xy & 0xFFFF)
            n = len(surf)
            if n:
                assert (pos[:n] == surf).all()
                assert (fl[:n] == 3).all() and (((row[:n] >> 32) & 0xFFFF) == 0).all()
                assert (fl[n:] == 1).all() and set(pos[n:].tolist()) <= set(surf.tolist())
            else:
                assert fl[0] == 2 and (fl[1:] == 0).all()
            assert w == max(len(groups_f[c]["surface"]) for groups_f in [
                {m.blob(ff, bb)["group"]: m.blob(ff, bb) for bb in range(m.blob_count(ff))} for ff in range(h)])


def test_refinement_plain_then_batched_then_graph(golden_cfg1):
    m = _pipeline(list(golden_cfg1["frames"]))
    c0 = m.cost()
    pts0, widths = m.chains()
    G.log(f"plain refine from cost {c0}")
    m.run(3 * len(widths))
    c1 = m.cost()
    assert c1 <= c0
    a1, acc1 = m.counters()
    assert a1 >= 3 * (1 << 16) * 0.99
    m.set_refine(M.REFINE_BATCHED, rounds=16, tile=2, graph_passes=0)
    G.log("batched refine")
    m.run(20)
    c2 = m.cost()
    assert c2 <= c1
    m.set_refine(M.REFINE_BATCHED, rounds=16, tile=3, graph_passes=8)
    G.log("graph refine")
    m.run(5)
    c3 = m.cost()
    assert c3 <= c2
    m.set_refine(M.REFINE_KNN, rounds=8, tile=1, graph_passes=0)
    G.log("knn refine")
    m.run(10)
    c4 = m.cost()
    assert c4 <= c3
    a3, _ = m.counters()
    assert a3 > a1
    c3 = c4
    # rows remain permutations of the initial rows
    pts3, _ = m.chains()
    off = np.concatenate([[0], np.cumsum(widths)[:-1]])
    for o, w in zip(off, widths):
        for y in range(pts3.shape[0]):
            assert (np.sort(pts3[y, o:o + w]) == np.sort(pts0[y, o:o + w])).all()
    # exact cost equals the oracle's recomputation
    am = G.atom_major(pts3.ravel(), pts3.shape[0])
    h = pts3.shape[0]
    assert c3 == sum(op.chain_cost(am[o * h:(o + w) * h], h) for o, w in zip(off, widths))
    G.log(f"costs: {c0} -> {c1} -> {c2} -> {c3}")


def test_render_matches_oracle_on_engine_chains(golden_cfg1):
    """Frames from am::morph equal the oracle's rendering of the engine's own
    chains and blob order within 1 LSB; the pixel-vector API gives the same
    image as the bulk one."""
    m = _pipeline(list(golden_cfg1["frames"]))
    m.run(6)
    pts, widths = m.chains()
    h = pts.shape[0]
    nb = m.blob_count(0)
    order = np.zeros((h, nb), np.int32)
    for f in range(h):
        for b in range(nb):
            order[f, b] = m.blob(f, b)["group"]
    off = np.concatenate([[0], np.cumsum(widths)[:-1]]).astype(np.int64)
    keys = np.ascontiguousarray(golden_cfg1["shown"]).view(np.uint8).reshape(h, 64, 64, 4)
    lag, slope = op.perlin_perm(1), op.perlin_perm(2)
    bbox = golden_cfg1["bbox"]
    cfg = op.RenderCfg(64, 64, h, int(bbox[0]), int(bbox[1]), int(bbox[2]), int(bbox[3]), 4, 6, 0, 1, 0, 0, 0, 0,
                       op.ptr(lag), op.ptr(slope))
    am = G.atom_major(pts.ravel(), h)
    w64 = widths.astype(np.int64)
    out = np.zeros((64, 64, 4), np.uint8)
    worst = 0
    for i in range(16):
        t = m.get_time(i, 16)
        G.log(f"engine render frame {i}")
        got = m.render(t)
        op.oracle().or_render_atoms(C.byref(cfg), op.ptr(keys), op.ptr(am), op.ptr(off), op.ptr(w64), len(widths),
                                    op.ptr(order), nb, None, float(t), op.ptr(out))
        worst = max(worst, int(np.abs(got.astype(int) - out.astype(int)).max()))
        if i % 5 == 0:
            assert (m.render_pixels(t) == got).all()
    assert worst <= 1


def test_blobs512_pipeline_runs():
    frames = synth.workload("blobs512")
    m = _pipeline(frames, cycle=1 << 14, match_steps=500)
    pts, widths = m.chains()
    assert len(widths) == 40 and pts.shape == (2, widths.sum())
    c0 = m.cost()
    m.set_refine(M.REFINE_BATCHED, rounds=32, tile=4, graph_passes=16)
    G.log("blobs512 graph refine")
    m.run(8)
    c1 = m.cost()
    assert c1 < c0
    img = m.render(0.25)
    assert img.shape == (512, 512, 4) and (img[..., 3] > 0).sum() > 50000


def test_chain_checkpoint_round_trip(tmp_path):
    """save_chains / load_chains: a second engine on the same key frames and
    seed resumes from the file with the same points, cost and frames; a
    corrupted file and chains of another shape are refused."""
    frames = synth.workload("ellipses64")
    a = _pipeline(frames)
    a.set_refine(M.REFINE_BATCHED, rounds=8, tile=2, graph_passes=0)
    G.log("checkpoint: refine and save")
    a.run(4)
    path = tmp_path / "chains.amc"
    a.save_chains(path)
    pts_a, widths = a.chains()
    b = _pipeline(frames)
    pts_b, _ = b.chains()
    assert not np.array_equal(pts_b, pts_a)
    G.log("checkpoint: load")
    b.load_chains(path)
    pts_b, widths_b = b.chains()
    assert np.array_equal(pts_b, pts_a) and np.array_equal(widths_b, widths)
    assert b.cost() == a.cost()
    assert (b.render(0.3) == a.render(0.3)).all()
    b.set_refine(M.REFINE_BATCHED, rounds=8, tile=2, graph_passes=0)
    b.run(2)
    assert b.cost() <= a.cost()
    data = bytearray(path.read_bytes())
    data[len(data) // 2] ^= 0x10
    bad = tmp_path / "bad.amc"
    bad.write_bytes(bytes(data))
    with pytest.raises(A.AmhError, match="checksum"):
        b.load_chains(bad)
    other = _pipeline(synth.ellipse_frames(seed=2, size=64, count=6, nframes=2))
    with pytest.raises(A.AmhError, match="shape|widths"):
        other.load_chains(path)


def _edge_frames(case):
    """Key frames for the edge cases, all on a 64 x 64 canvas."""
    if case == "three_keys":
        return synth.ellipse_frames(seed=3, size=64, count=5, nframes=3)
    if case == "uneven_blob_counts":   # 6 blobs, then 2: chains pad with empty blobs
        a = synth.ellipse_frames(seed=4, size=64, count=6, nframes=1)[0]
        b = synth.ellipse_frames(seed=5, size=64, count=2, nframes=1)[0]
        return [a, b]
    if case == "empty_key_frame":      # a key frame with no pixel at all
        a = synth.ellipse_frames(seed=6, size=64, count=4, nframes=1)[0]
        return [a, np.zeros_like(a), synth.ellipse_frames(seed=7, size=64, count=3, nframes=1)[0]]
    if case == "ragged_sizes":         # key frames of different sizes
        a = synth.ellipse_frames(seed=8, size=64, count=5, nframes=1)[0]
        b = synth.ellipse_frames(seed=9, size=48, count=4, nframes=1)[0][:40]
        return [a, b]
    raise KeyError(case)


@pytest.mark.parametrize("case", ["three_keys", "uneven_blob_counts", "empty_key_frame", "ragged_sizes"])
def test_edge_cases_render_like_the_oracle(case):
    """Through am::morph: rows stay permutations of the built chains while
    refining, and 12 frames render within 1 LSB of the oracle's rendering of
    the engine's chains and blob order."""
    frames = _edge_frames(case)
    m = M.Morph()
    m.set(seed=1, cycle_length=1 << 12)
    for i, f in enumerate(frames):
        m.add_frame_rgba(i, f)
    m.set_resolution(64, 64)
    G.log(f"edge case {case}: pipeline")
    assert m.run(1) == M.STATE_BLOB_UNIFICATION
    assert m.run(501) == M.STATE_BLOB_MATCHING
    m.next_state()
    assert m.run(1) == M.STATE_ATOM_MORPHING, m.error()
    pts0, widths = m.chains()
    h = pts0.shape[0]
    assert h == len(frames)
    m.set_refine(M.REFINE_BATCHED, rounds=8, tile=0, graph_passes=0)
    m.run(3)
    pts, widths2 = m.chains()
    assert np.array_equal(widths, widths2)
    cid = np.repeat(np.arange(len(widths)), widths)
    for y in range(h):
        a, b = np.lexsort((pts[y], cid)), np.lexsort((pts0[y], cid))
        assert np.array_equal(pts[y][a], pts0[y][b]), y
    # the oracle on the engine's chains: keys on the 64 x 64 canvas
    canvas = []
    for f in frames:
        c = np.zeros((64, 64, 4), np.uint8)
        c[:f.shape[0], :f.shape[1]] = f
        canvas.append(c)
    keys = np.stack([op.ingest(synth.to_words(c))[1] for c in canvas]).view(np.uint8).reshape(h, 64, 64, 4)
    present = np.any(np.stack([c[..., 3] > 0 for c in canvas]), axis=0)
    ys, xs = np.nonzero(present)
    nb = max(m.blob_count(f) for f in range(h))
    order = np.full((h, nb), -1, np.int32)
    for f in range(h):
        for b in range(m.blob_count(f)):
            order[f, b] = m.blob(f, b)["group"]
    lag, slope = op.perlin_perm(1), op.perlin_perm(2)
    cfg = op.RenderCfg(64, 64, h, int(xs.min()), int(ys.min()), int(xs.max()), int(ys.max()), 4, 6, 0, 1, 0, 0, 0,
                       0, op.ptr(lag), op.ptr(slope))
    off = np.concatenate([[0], np.cumsum(widths)[:-1]]).astype(np.int64)
    am = G.atom_major(pts.ravel(), h)
    w64 = widths.astype(np.int64)
    out = np.zeros((64, 64, 4), np.uint8)
    worst = 0
    for i in range(12):
        t = m.get_time(i, 12)
        got = m.render(t)
        op.oracle().or_render_atoms(C.byref(cfg), op.ptr(keys), op.ptr(am), op.ptr(off), op.ptr(w64), len(widths),
                                    op.ptr(order), nb, None, float(t), op.ptr(out))
        worst = max(worst, int(np.abs(got.astype(int) - out.astype(int)).max()))
    G.log(f"edge case {case}: worst channel difference {worst}")
    assert worst <= 1


def _blobify_fixture():
    import os
    return dict(np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "blobify.npz")))


@pytest.mark.parametrize("c", range(9))
def test_greedy_blob_options_through_the_engine(c):
    """Non-default blob options: detection runs the host greedy merge and
    dust unification over the device-ingested stored colours; the blobs the
    engine reports (order, running means, membership) are the compiled
    reference's, and the pipeline runs on to rendered frames."""
    fx = _blobify_fixture()
    W, H, nf = (int(v) for v in fx["bf_size"])
    mx, mn, num, samples, grip, delim, seed = (int(v) for v in fx["bf_opts"][c])
    opts = dict(blob_threshold=float(fx["bf_threshold"][c]), blob_min_size=mn, blob_number=num,
                blob_box_samples=samples, blob_box_grip=grip, blob_delimiter=delim)
    if mx != M.SIZE_MAX:
        opts["blob_max_size"] = mx
    m = M.Morph()
    m.set(seed=seed, cycle_length=1 << 12, **opts)
    for i in range(nf):
        m.add_frame_rgba(i, fx["bf_frames"][c, i])
    m.set_resolution(W, H)
    G.log(f"greedy blobs case {c}: {opts}")
    assert m.run(1) == M.STATE_BLOB_UNIFICATION
    base = int(fx["bf_nblobs"][:c].sum())
    for f in range(nf):
        n = int(fx["bf_nblobs"][c, f])
        assert m.blob_count(f) == n
        for b in range(n):
            k = base + int(fx["bf_nblobs"][c, :f].sum()) + b
            got = m.blob(f, b)
            assert np.array_equal(got["means"], fx["bf_means"][k]), (f, b)
            idx = np.flatnonzero(fx["bf_bop"][c, f] == b)
            want = ((idx // W) * 65536 + idx % W).astype(np.uint64)
            assert np.array_equal(np.sort(got["surface"]), want), (f, b)
    assert m.run(201) == M.STATE_BLOB_MATCHING
    m.next_state()
    assert m.run(1) == M.STATE_ATOM_MORPHING
    m.run(4)
    img = m.render_pixels(0.5)
    assert img.shape == (H, W, 4) and img[..., 3].any()
