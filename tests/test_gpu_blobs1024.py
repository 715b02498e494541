"""The 1024^2 full-pipeline workload on the GPU (4 key frames of 150
noise-thresholded blobs each, fluid on a 276 x 276 grid) and the learned
warm start evaluated on those frames, with full-size checks:

- blob detection equals the oracle's 4-connected labelling, frame by frame;
- the learned warm start and the batched descent in hipGraph replays keep
  every chain row a permutation of its points, and the exact cost only
  falls (and equals the oracle's cost of the downloaded chains);
- atom-path frames equal the oracle's rendering of the same chains within
  1 LSB;
- the fluid frames render, and one model step of all of the job's real
  particles agrees with the FP64 oracle's step."""
import ctypes as C

import numpy as np
import pytest

from atomorph_amd import learned as L
from atomorph_amd import morph as M
from atomorph_amd import synth
from oracle import oracle_py as op
from tests import gpu_util as G
from tests.test_gpu_fluid import _step_gpu
from tests.test_gpu_large import _same_rows

pytestmark = pytest.mark.gpu

SPARSENESS = 0.25


@pytest.fixture(scope="module")
def job():
    frames = synth.workload("blobs1024")
    H, W = frames[0].shape[:2]
    m = M.Morph()
    m.set(seed=1, cycle_length=1 << 16)
    for i, f in enumerate(frames):
        m.add_frame_rgba(i, f)
    m.set_resolution(W, H)
    G.log("blobs1024: detection")
    assert m.run(1) == M.STATE_BLOB_UNIFICATION
    return {"m": m, "frames": frames, "W": W, "H": H}


def test_blobs1024_detection_equals_oracle(job):
    m, frames, W = job["m"], job["frames"], job["W"]
    for f in range(len(frames)):
        lab, n = op.segment(frames[f])
        assert m.blob_count(f) == n == 150
        got = sorted(tuple(((s // 65536) * W + (s % 65536)).tolist())
                     for s in (m.blob(f, b)["surface"] for b in range(n)))
        order = np.argsort(lab, kind="stable")
        bounds = np.flatnonzero(np.diff(lab[order])) + 1
        want = sorted(tuple(np.sort(g).tolist()) for g in np.split(order, bounds) if lab[g[0]] >= 0)
        assert got == want, f"frame {f}"


def test_blobs1024_learned_warm_start_and_batched_descent(job):
    m, frames = job["m"], job["frames"]
    params = L.load_default()
    if params is None:
        pytest.fail("trained parameters missing: run python -m atomorph_amd.learned")
    G.log("blobs1024: matching (20,000 steps) and chain build")
    assert m.run(20001) == M.STATE_BLOB_MATCHING
    m.next_state()
    assert m.run(1) == M.STATE_ATOM_MORPHING, m.error()
    pts0, widths = m.chains()
    h = pts0.shape[0]
    c_plain = m.cost()
    assert op.chain_cost(G.atom_major(pts0.ravel(), h), h) == c_plain
    G.log(f"blobs1024: learned warm start on {int(widths.sum())} atoms x {h} key frames, {len(widths)} chains")
    m.set_learned(params)
    m.warm_start()
    m.set_learned(None)
    pts1, _ = m.chains()
    c_warm = m.cost()
    assert _same_rows(pts1, pts0, widths)
    assert np.array_equal(pts1[0], pts0[0])  # row 0 keeps its order
    assert op.chain_cost(G.atom_major(pts1.ravel(), h), h) == c_warm
    G.log(f"blobs1024: cost plain start {c_plain}, learned start {c_warm}")
    assert c_warm < c_plain
    G.log("blobs1024: batched descent, autotuned tile, hipGraph replays of 16 passes")
    m.set_refine(M.REFINE_BATCHED, rounds=32, tile=-1, graph_passes=16)
    costs = [c_warm]
    for _ in range(3):
        m.run(1)
        costs.append(m.cost())
    assert all(b <= a for a, b in zip(costs, costs[1:])) and costs[-1] < costs[0], costs
    pts2, _ = m.chains()
    assert _same_rows(pts2, pts0, widths)
    assert op.chain_cost(G.atom_major(pts2.ravel(), h), h) == costs[-1]
    job["pts"], job["widths"] = pts2, widths


def test_blobs1024_atom_frames_equal_oracle(job):
    if "pts" not in job:
        pytest.skip("needs the chains of the previous test")
    m, frames, W, H = job["m"], job["frames"], job["W"], job["H"]
    pts, widths = job["pts"], job["widths"]
    h = pts.shape[0]
    F = 60
    T = G.torch()
    out = T.empty(F * H * W * 4, dtype=T.uint8, device="cuda")
    times = [m.get_time(i, F) for i in range(F)]
    G.log("blobs1024: 60 atom-path frames into HBM")
    assert m.render_device(times, out.data_ptr()) > 0
    frames_dev = out.view(F, H, W, 4)
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
    am = G.atom_major(pts.ravel(), h)
    w64 = widths.astype(np.int64)
    want = np.zeros((H, W, 4), np.uint8)
    for i in (0, 11, 29, 52):
        G.log(f"blobs1024: oracle frame {i}")
        op.oracle().or_render_atoms(C.byref(cfg), op.ptr(keys), op.ptr(am), op.ptr(off), op.ptr(w64), len(widths),
                                    op.ptr(order), nb, None, float(times[i]), op.ptr(want))
        got = frames_dev[i].cpu().numpy()
        assert int(np.abs(got.astype(int) - want.astype(int)).max()) <= 1, f"frame {i}"
        assert (want[..., 3] > 0).sum() > 100000


def test_blobs1024_fluid_frames_and_step_parity(job):
    if "pts" not in job:
        pytest.skip("needs the chains of the previous tests")
    m, W, H = job["m"], job["W"], job["H"]
    m.set(fluid=10, fluid_sparseness=SPARSENESS)
    F = 60
    T = G.torch()
    out = T.empty(F * H * W * 4, dtype=T.uint8, device="cuda")
    times = [m.get_time(i, F) for i in range(F)]
    G.log("blobs1024: 60 fluid frames (10 steps each) into HBM")
    assert m.render_device(times, out.data_ptr()) > 0
    cover = [(out.view(F, H, W, 4)[i, :, :, 3] > 0).sum().item() for i in (0, 20, 40, 59)]
    assert all(c > 100000 for c in cover), cover
    P = np.ascontiguousarray(m.fluid_particles())
    active = int(P[:, 16].sum())
    assert active > 100000 and np.isfinite(P[:, :4]).all()
    gs = int(W * SPARSENESS) + 20
    radius = np.sqrt(W * W + H * H) * SPARSENESS / 4.0 * 0.5
    G.log(f"blobs1024: fluid step parity, {P.shape[0]} particles ({active} active) on a {gs}x{gs} grid")
    outs = _step_gpu(P, gs, gs, 2, radius)
    Q = P.copy()
    o = op.oracle()
    for k in range(2):
        o.or_fluid_step(op.ptr(Q), Q.shape[0], gs, gs, 1 - k, radius, 0)
        for col in (0, 1, 2, 3):
            dd = np.abs(outs[k][:, col] - Q[:, col])
            d = dd.max()
            worst = int(np.argmax(dd))
            G.log(f"  step {k} field {col}: max difference {d:.2e} (particle {worst}, at {Q[worst, :2]}), "
                  f"99.9% {np.quantile(dd, 0.999):.2e}, 99% {np.quantile(dd, 0.99):.2e}, "
                  f"median {np.median(dd):.2e}")
            assert d <= 1e-4, (k, col, d)  # FP32 grid against the FP64 oracle
        assert (outs[k][:, 16] == Q[:, 16]).all()
