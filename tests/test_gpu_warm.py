"""The learned warm start on the GPU: sort-and-match keeps every chain row a
permutation of its points, and the trained network lowers the transport
cost the refinement starts from (and ends at for a fixed budget) against the
plain start the reference uses (rows in pixel order) and against the same
sort-and-match driven by an untrained network."""
import numpy as np
import pytest

from atomorph_amd import learned as L
from atomorph_amd import morph as M
from atomorph_amd import synth
from tests import gpu_util as G

pytestmark = pytest.mark.gpu


def _fresh(frames, seed=1):
    return L.build(frames, seed=seed, match_steps=2000)


def _rows(pts, widths):
    off = np.concatenate([[0], np.cumsum(widths)[:-1]])
    return [(o, w) for o, w in zip(off, widths)]


def test_warm_start_rows_stay_permutations():
    frames = synth.workload("blobs512")
    m = _fresh(frames)
    before, widths = m.chains()
    m.set_learned(L.init_params(seed=3))
    G.log("warm start, random network, blobs512")
    ms = m.warm_start()
    after, w2 = m.chains()
    assert np.array_equal(widths, w2) and ms > 0.0
    h = before.shape[0]
    assert np.array_equal(after[0], before[0])  # row 0 keeps its order
    moved = 0
    for y in range(1, h):
        for o, w in _rows(before, widths):
            a, b = np.sort(after[y, o:o + w]), np.sort(before[y, o:o + w])
            assert np.array_equal(a, b), (y, o, w)
            moved += int(np.count_nonzero(after[y, o:o + w] != before[y, o:o + w]))
    assert moved > 0


def test_trained_network_gains_over_plain_and_untrained_start():
    p = L.load_default()
    if p is None:
        pytest.fail("trained parameters missing: run python -m atomorph_amd.learned")
    frames = synth.workload("blobs512")
    base = _fresh(frames)
    c_plain = base.cost()
    pts0, widths = base.chains()
    costs = {}
    for name, params in (("trained", p), ("untrained", L.init_params(seed=3))):
        m = _fresh(frames)
        m.put_chains(pts0)
        m.set_learned(params)
        G.log(f"warm start {name}")
        m.warm_start()
        costs[name] = m.cost()
    G.log(f"start costs: plain {c_plain}, {costs}")
    assert costs["trained"] < 0.97 * c_plain
    assert costs["trained"] < costs["untrained"]
    # after the same batched budget the warm-started chains stay ahead
    ends = {}
    for name, params in (("plain", None), ("trained", p)):
        m = _fresh(frames)
        m.put_chains(pts0)
        if params is not None:
            m.set_learned(params)
            m.warm_start()
        m.set_refine(M.REFINE_BATCHED, rounds=32, tile=4, graph_passes=0)
        m.run(4)
        ends[name] = m.cost()
    G.log(f"after 4 batched passes: {ends}")
    assert ends["trained"] <= ends["plain"]
