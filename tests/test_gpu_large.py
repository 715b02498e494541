"""The 2048^2 power-law workloads end to end on the GPU: two key frames of
1200 separate blobs each, areas drawn from p(s) ~ s^-1.8 on [8, 60000] px
(about 0.2 M blob pixels per frame) and on [200, 60000] px (about 2 M per
frame, 4.1 M over the pair: the scale of a ~4 M-atom pair).

Blob detection equals the oracle's 4-connected labelling; the learned warm
start keeps every chain row a permutation and lowers the starting cost; the
autotuned batched descent (hipGraph replays) never raises the exact cost,
keeps the rows permutations and ends on the cost the oracle recomputes from
the chains; 120 frames render into HBM in one call and sampled frames equal
the oracle's rendering of the same chains within 1 LSB."""
import ctypes as C

import numpy as np
import pytest

from atomorph_amd import learned as L
from atomorph_amd import morph as M
from atomorph_amd import synth
from oracle import oracle_py as op
from tests import gpu_util as G

pytestmark = pytest.mark.gpu

FRAMES_OUT = 120


def _same_rows(a, b, widths):
    """Every chain row of a is a permutation of the same row of b."""
    cid = np.repeat(np.arange(len(widths)), widths)
    for y in range(a.shape[0]):
        ka = np.lexsort((a[y], cid))
        kb = np.lexsort((b[y], cid))
        if not np.array_equal(a[y][ka], b[y][kb]):
            return False
    return True


@pytest.mark.parametrize("name,sampled", [("powerlaw2048", (0, 37, 61, 119)), ("powerlaw2048_4m", (0, 61))])
def test_powerlaw2048_learned_autotuned_pipeline(name, sampled):
    frames = synth.workload(name)
    H, W = frames[0].shape[:2]
    params = L.load_default()
    if params is None:
        pytest.fail("trained parameters missing: run python -m atomorph_amd.learned")
    m = M.Morph()
    m.set(seed=1, cycle_length=1 << 16)
    for i, f in enumerate(frames):
        m.add_frame_rgba(i, f)
    m.set_resolution(W, H)
    G.log(f"{name}: detection")
    assert m.run(1) == M.STATE_BLOB_UNIFICATION
    for f in range(2):
        lab, n = op.segment(frames[f])
        assert m.blob_count(f) == n == 1200
        got = sorted(tuple(((s // 65536) * W + (s % 65536)).tolist())
                     for s in (m.blob(f, b)["surface"] for b in range(n)))
        order = np.argsort(lab, kind="stable")
        bounds = np.flatnonzero(np.diff(lab[order])) + 1
        want = sorted(tuple(np.sort(g).tolist()) for g in np.split(order, bounds) if lab[g[0]] >= 0)
        assert got == want, f"frame {f}"
    G.log(f"{name}: matching + chain build")
    assert m.run(2001) == M.STATE_BLOB_MATCHING
    m.next_state()
    assert m.run(1) == M.STATE_ATOM_MORPHING
    pts0, widths = m.chains()
    c_plain = m.cost()

    G.log(f"{name}: learned warm start ({int(widths.sum())} atoms, {len(widths)} chains)")
    m.set_learned(params)
    m.warm_start()
    m.set_learned(None)
    pts1, _ = m.chains()
    c_warm = m.cost()
    assert c_warm < c_plain
    assert _same_rows(pts1, pts0, widths)

    G.log(f"{name}: autotuned batched descent in hipGraph replays")
    m.set_refine(M.REFINE_BATCHED, rounds=0, tile=-1, graph_passes=16)
    costs = [c_warm]
    for _ in range(4):
        m.run(1)
        costs.append(m.cost())
    assert all(b <= a for a, b in zip(costs, costs[1:])) and costs[-1] < costs[0], costs
    pts2, _ = m.chains()
    assert _same_rows(pts2, pts0, widths)
    h = pts2.shape[0]
    assert op.chain_cost(G.atom_major(pts2.ravel(), h), h) == costs[-1]

    G.log(f"{name}: render {FRAMES_OUT} frames into HBM")
    T = G.torch()
    out = T.empty(FRAMES_OUT * H * W * 4, dtype=T.uint8, device="cuda")
    times = [m.get_time(i, FRAMES_OUT) for i in range(FRAMES_OUT)]
    ms = m.render_device(times, out.data_ptr())
    assert ms > 0
    G.log(f"{name}: {FRAMES_OUT} frames in {ms:.1f} ms")
    frames_dev = out.view(FRAMES_OUT, H, W, 4)

    # the oracle's rendering of the same chains and blob order
    nb = m.blob_count(0)
    order = np.zeros((h, nb), np.int32)
    for f in range(h):
        for b in range(nb):
            order[f, b] = m.blob(f, b)["group"]
    off = np.concatenate([[0], np.cumsum(widths)[:-1]]).astype(np.int64)
    keys = np.stack([op.ingest(synth.to_words(f))[1] for f in frames]).view(np.uint8).reshape(h, H, W, 4)
    ys, xs = np.nonzero(np.any(np.stack([f[..., 3] > 0 for f in frames]), axis=0))
    lag, slope = op.perlin_perm(1), op.perlin_perm(2)
    cfg = op.RenderCfg(W, H, h, int(xs.min()), int(ys.min()), int(xs.max()), int(ys.max()), 4, 6, 0, 1, 0, 0, 0, 0,
                       op.ptr(lag), op.ptr(slope))
    am = G.atom_major(pts2.ravel(), h)
    w64 = widths.astype(np.int64)
    want = np.zeros((H, W, 4), np.uint8)
    for i in sampled:
        G.log(f"{name}: oracle frame {i}")
        op.oracle().or_render_atoms(C.byref(cfg), op.ptr(keys), op.ptr(am), op.ptr(off), op.ptr(w64), len(widths),
                                    op.ptr(order), nb, None, float(times[i]), op.ptr(want))
        got = frames_dev[i].cpu().numpy()
        assert int(np.abs(got.astype(int) - want.astype(int)).max()) <= 1, f"frame {i}"
