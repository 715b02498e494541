"""Statistical parity of the round schedule with the reference descent at
h >= 3 key frames and at the chain widths of the 512^2 workload (hundreds to
thousands of atoms).  The reference's morph_asynch (restated bit for bit in
the oracle, pinned to the compiled reference) draws one row per call and
`repeat` random pairs on it (thread.cpp:989-1042); the GPU's schedule
visits the rows in turn with fold rounds of disjoint pairs.  After the same
number of attempts from the same row-major start (the reference's
init_morph order, thread.cpp:809-822) the two costs agree within 0.5% and
the round schedule keeps at least 95% of the reference's gain."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle_py as op


def _word(x, y):
    # x u16 | y u16 << 16 | zero fractions | HAS_PIXEL flag (atomorph.h:275-284)
    return x.astype(np.uint64) | (y.astype(np.uint64) << np.uint64(16)) | (np.uint64(1) << np.uint64(48))


def _chain(w, h, seed):
    """h key-frame rows of w distinct pixels each, one blob per key frame at
    its own place and shape (w of the pixels of an ellipse of about twice
    that area), every row in row-major order; atom-major."""
    rng = np.random.default_rng(seed)
    rows = []
    for _ in range(h):
        cx, cy = rng.uniform(100, 400, 2)
        ax = np.sqrt(2.0 * w / np.pi) * rng.uniform(0.7, 1.4)
        ay = 2.0 * w / (np.pi * ax) + 1.0
        ys, xs = np.mgrid[-int(ay) - 1:int(ay) + 2, -int(ax) - 1:int(ax) + 2]
        inside = (xs / ax) ** 2 + (ys / ay) ** 2 <= 1.0
        cand = np.stack([xs[inside] + int(cx), ys[inside] + int(cy)], 1)
        pick = cand[rng.choice(len(cand), w, replace=False)]
        pick = pick[np.lexsort((pick[:, 0], pick[:, 1]))]  # row-major: y, then x
        rows.append(_word(pick[:, 0], pick[:, 1]))
    return np.ascontiguousarray(np.stack(rows, 1).ravel())


@pytest.mark.parametrize("h,w", [(3, 500), (3, 2500), (4, 500), (4, 2500)])
def test_round_schedule_matches_reference_descent_at_more_key_frames(h, w):
    o = op.oracle()
    p0 = _chain(w, h, 7 * w + h)
    c0 = op.chain_cost(p0, h)
    budget, calls = 1 << 16, 64
    for s in range(2):
        q = p0.copy()
        g = C.c_double(0.0)
        for k in range(calls):  # one row per call, as morph_asynch draws it
            o.or_morph_asynch(op.ptr(q), w, h, budget // calls, 1000 * s + k, C.byref(g))
        ref = op.chain_cost(q, h)
        q, att, rnd = p0.copy(), 0, 0
        while att < budget:
            q, _, _, a = op.refine_rounds(q, h, rnd % h, 1, seed=99 + s, stream=rnd)
            att += a
            rnd += 1
        mine = op.chain_cost(q, h)
        assert ref < c0 and mine < c0
        assert abs(mine - ref) <= 0.005 * ref, (h, w, s, mine, ref)
        assert c0 - mine >= 0.95 * (c0 - ref), (h, w, s, c0, mine, ref)
