"""The batched-descent autotuner on the GPU: every candidate races from the
same start (the caller's chains stay untouched), its recorded result equals
the oracle's replay of the same passes, the winner is the best measured
candidate, and the engine's automatic schedule runs exactly the tuned one."""
import ctypes as C

import numpy as np
import pytest

from atomorph_amd import _capi as A
from atomorph_amd import learned as L
from atomorph_amd import morph as M
from atomorph_amd import synth
from oracle import oracle_py as op
from tests import gpu_util as G
from tests.test_gpu_refine import _chains, _tiles

pytestmark = pytest.mark.gpu

ROUND_CHOICES = (4, 8, 16, 32, 64)


def test_autotune_races_every_candidate_from_the_same_start():
    T = G.torch()
    lib = G.lib()
    rng = np.random.default_rng(11)
    h = 2
    pts, off, wid = _chains(rng, [2, 100, 300, 1000, 2000, 3000], h, lim=300)
    wtot = int(wid.sum())
    dev = G.to_dev(G.key_major(pts, h))
    before = G.to_host(dev, np.uint64).copy()
    work = G.zeros(wtot * h, T.int64)
    off_d, w_d = G.to_dev(off), G.to_dev(wid.astype(np.int32))
    w_host = np.ascontiguousarray(wid.astype(np.int32))
    tile, rounds, n = C.c_int32(-1), C.c_int32(0), C.c_int32(0)
    rep = (A.TuneEntry * 64)()
    budget = 0.3
    G.log(f"autotune wtot={wtot} budget={budget} ms")
    G.check(lib.amh_refine_autotune(G.dptr(dev), G.dptr(work), h, wtot, G.dptr(off_d), G.dptr(w_d),
                                    w_host.ctypes.data_as(C.c_void_p), len(wid), budget, 4242, 0, C.byref(tile),
                                    C.byref(rounds), rep, 64, C.byref(n), G.stream()), "amh_refine_autotune")
    G.sync()
    assert np.array_equal(G.to_host(dev, np.uint64), before)        # the caller's chains are untouched
    tiles = _tiles()
    assert n.value == len(tiles) * len(ROUND_CHOICES)
    rows = [rep[i] for i in range(n.value)]
    assert [(r.tile, r.rounds) for r in rows] == [(t, r) for t, _, _ in tiles for r in ROUND_CHOICES]
    for r in rows:
        assert r.passes >= 1 and (r.ms >= budget or r.passes == 512)
        assert 0 <= r.gain_at_budget <= r.gain
    best = max(rows, key=lambda r: r.gain_at_budget)
    assert (tile.value, rounds.value) == (best.tile, best.rounds)
    # each candidate ran its own passes from the same start: the oracle's
    # replay of that many passes gives the same exact gain
    caps = {t: cap for t, cap, _ in tiles}
    for r in (rows[0], best):
        want, gain = pts.copy(), 0
        for p in range(r.passes):
            want, _, g, _ = op.refine_pass(want, off, wid, h, caps[r.tile], p, r.rounds, 4242)
            gain += g
        assert gain == r.gain, (r.tile, r.rounds, r.passes)


def test_engine_automatic_schedule_runs_the_tuned_one():
    frames = synth.workload("blobs512")
    m = L.build(frames, seed=1, match_steps=2000)
    G.log("engine autotune, blobs512")
    tile, rounds, rows = m.autotune(2.0)
    assert rows and (tile, rounds) == max(((r["tile"], r["rounds"]) for r in rows),
                                          key=lambda tr: next(r["gain_at_budget"] for r in rows
                                                              if (r["tile"], r["rounds"]) == tr))
    m.set_refine(M.REFINE_BATCHED, rounds=0, tile=-1, graph_passes=0)   # automatic: the tuned schedule
    m.run(3)
    auto, _ = m.chains()
    ref = L.build(frames, seed=1, match_steps=2000)
    ref.set_refine(M.REFINE_BATCHED, rounds=rounds, tile=tile, graph_passes=0)
    ref.run(3)
    explicit, _ = ref.chains()
    assert np.array_equal(auto, explicit)
    assert m.cost() == ref.cost()
    # a fixed rounds count: one candidate per kernel variant, all at it
    tile16, rounds16, rows16 = m.autotune(1.0, rounds=16)
    assert rounds16 == 16 and len(rows16) == len(_tiles()) and all(r["rounds"] == 16 for r in rows16)
    assert tile16 == max(rows16, key=lambda r: r["gain_at_budget"])["tile"]
