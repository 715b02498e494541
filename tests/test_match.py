"""Blob matching in the oracle (or_match, restating thread::match,
thread.cpp:598-738) against the compiled reference (tests/golden/match.npz):
list order after the padding shuffle, every blob's group, the energies and
the fix_volatiles positions, bit for bit after 1 .. 3000 steps; and the
descent's result against an exhaustive search on small frames."""
import ctypes as C
import itertools
import os

import numpy as np
import pytest

from oracle import oracle_py as op

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "match.npz")


@pytest.fixture(scope="module")
def mfx():
    return dict(np.load(GOLDEN))


def case_inputs(d, c):
    p = f"m{c}_"
    nb = d[p + "nblobs0"].astype(np.int32)
    opts = d[p + "opts"]
    weights = np.array(opts[6:9], np.uint8)
    return dict(h=len(nb), nblobs=nb, sizes=d[p + "sizes0"].astype(np.int64),
                means=np.ascontiguousarray(d[p + "means0"]), fmeans=np.ascontiguousarray(d[p + "fmeans"]),
                bbox=d[p + "bbox"].astype(np.int32), weights=weights, degeneration=int(opts[5]),
                rng=int(d[p + "rng0"]))


def run_or_match(inp, steps, fix=False):
    h, bw = inp["h"], int(inp["nblobs"].max())
    order = np.zeros((h, bw), np.int32)
    group = np.zeros((h, bw), np.int32)
    energy = np.zeros(3)
    fixed = np.zeros((h, bw, 2))
    rng_out = C.c_uint64()
    n = op.oracle().or_match(h, op.ptr(inp["nblobs"]), op.ptr(inp["sizes"]), op.ptr(inp["means"]),
                             op.ptr(inp["fmeans"]), op.ptr(inp["bbox"]), op.ptr(inp["weights"]),
                             inp["degeneration"], inp["rng"], steps, 1 if fix else 0, op.ptr(order),
                             op.ptr(group), op.ptr(energy), op.ptr(fixed), C.byref(rng_out))
    assert n == bw
    return order, group, energy, fixed


@pytest.mark.parametrize("c", range(6))
def test_or_match_equals_reference_groups_and_energies(mfx, c):
    inp = case_inputs(mfx, c)
    p = f"m{c}_"
    for k, steps in enumerate(mfx["match_steps"]):
        order, group, energy, _ = run_or_match(inp, int(steps))
        assert np.array_equal(order, mfx[p + "order"][k]), (c, steps)
        assert np.array_equal(group, mfx[p + "group"][k]), (c, steps)
        assert np.array_equal(energy, mfx[p + "energy"][k]), (c, steps, energy, mfx[p + "energy"][k])


@pytest.mark.parametrize("c", range(6))
def test_or_match_fix_volatiles_equals_reference(mfx, c):
    inp = case_inputs(mfx, c)
    p = f"m{c}_"
    _, _, _, fixed = run_or_match(inp, int(mfx["match_steps"][-1]), fix=True)
    assert np.array_equal(fixed, mfx[p + "fixed_xy"]), c


def test_fixture_exercises_the_branches(mfx):
    """The cases cover ties of the degeneration path (current energy above
    the best), padding blobs, an empty key frame and non-default weights."""
    degenerate = [c for c in range(6) if mfx[f"m{c}_opts"][5] > 0]
    assert any(mfx[f"m{c}_energy"][-1][0] > mfx[f"m{c}_energy"][-1][1] for c in degenerate)
    assert any((mfx[f"m{c}_order"][-1] == -1).any() for c in range(6))
    assert any((mfx[f"m{c}_nblobs0"] == 0).any() for c in range(6))


def _padded(inp, order):
    """Padded per-frame lists in list order: sizes[h][bw], means[h][bw][6]."""
    h, bw = order.shape
    sizes = np.zeros((h, bw), np.int64)
    means = np.zeros((h, bw, 6))
    base = np.concatenate([[0], np.cumsum(inp["nblobs"])])
    for y in range(h):
        for b in range(bw):
            i = order[y, b]
            if i >= 0:
                sizes[y, b] = inp["sizes"][base[y] + i]
                means[y, b] = inp["means"][base[y] + i]
            else:
                means[y, b] = inp["fmeans"][y]
    return sizes, np.ascontiguousarray(means)


def test_match_energy_closed_form(mfx):
    """or_match_energy of the reference's final groups equals the best energy
    it reports, up to the running sum's rounding."""
    for c in range(6):
        inp = case_inputs(mfx, c)
        order = mfx[f"m{c}_order"][-1]
        group = np.ascontiguousarray(mfx[f"m{c}_group"][-1], np.int32)
        sizes, means = _padded(inp, order)
        h, bw = order.shape
        e = op.oracle().or_match_energy(h, bw, op.ptr(sizes), op.ptr(means), op.ptr(inp["bbox"]),
                                        op.ptr(inp["weights"]), op.ptr(group))
        best = mfx[f"m{c}_energy"][-1][1]
        assert abs(e - best) <= 1e-9 * max(1.0, best), (c, e, best)


def _small_case(seed, h, nb):
    rng = np.random.default_rng(seed)
    nbl = np.array([rng.integers(max(1, nb - 2), nb + 1) for _ in range(h)], np.int32)
    nbl[rng.integers(0, h)] = nb
    tot = int(nbl.sum())
    sizes = rng.integers(1, 400, tot).astype(np.int64)
    means = np.zeros((tot, 6))
    means[:, 0] = rng.uniform(0, 63, tot)
    means[:, 1] = rng.uniform(0, 63, tot)
    means[:, 2:] = rng.uniform(0, 1, (tot, 4))
    fmeans = np.tile(np.array([32.0, 32.0, 0.5, 0.5, 0.5, 1.0]), (h, 1))
    return dict(h=h, nblobs=nbl, sizes=sizes, means=means, fmeans=fmeans, bbox=np.array([0, 0, 63, 63], np.int32),
                weights=np.array([1, 1, 1], np.uint8), degeneration=0, rng=int(rng.integers(1, 2 ** 31 - 1)))


def test_match_brute_force_is_exhaustive():
    """or_match_brute against an enumeration written here from the energy
    alone: every assignment of frame 1 (h = 2) evaluated through
    or_match_energy."""
    inp = _small_case(5, 2, 5)
    order, _, _, _ = run_or_match(inp, 1)
    sizes, means = _padded(inp, order)
    bw = order.shape[1]
    o = op.oracle()
    best = np.inf
    for perm in itertools.permutations(range(bw)):
        g = np.array([list(range(bw)), list(perm)], np.int32)
        best = min(best, o.or_match_energy(2, bw, op.ptr(sizes), op.ptr(means), op.ptr(inp["bbox"]),
                                           op.ptr(inp["weights"]), op.ptr(g)))
    got = o.or_match_brute(2, bw, op.ptr(sizes), op.ptr(means), op.ptr(inp["bbox"]), op.ptr(inp["weights"]),
                           None)
    assert got == best


def test_match_descent_reaches_a_swap_local_optimum_near_the_exhaustive_one():
    """The reference's matching is a random pair-swap descent: after enough
    steps no single swap in any key frame lowers the energy, and on frames
    of at most 6 blobs it usually lands on the exhaustive optimum."""
    o = op.oracle()
    hits, cases = 0, 0
    for seed in range(40):
        h = 2 if seed % 2 == 0 else 3
        inp = _small_case(100 + seed, h, 6 if h == 2 else 5)
        order, group, energy, _ = run_or_match(inp, 20000)
        sizes, means = _padded(inp, order)
        bw = order.shape[1]
        args = (op.ptr(sizes), op.ptr(means), op.ptr(inp["bbox"]), op.ptr(inp["weights"]))
        e = o.or_match_energy(h, bw, *args, op.ptr(group))
        # 2-opt local optimality in every frame
        for y in range(h):
            for a, b in itertools.combinations(range(bw), 2):
                g = group.copy()
                ia, ib = np.flatnonzero(g[y] == a)[0], np.flatnonzero(g[y] == b)[0]
                g[y, ia], g[y, ib] = b, a
                assert o.or_match_energy(h, bw, *args, op.ptr(g)) >= e - 1e-12, (seed, y, a, b)
        best = o.or_match_brute(h, bw, *args, None)
        assert best <= e + 1e-12
        hits += int(e <= best + 1e-12)
        cases += 1
    assert hits >= 0.7 * cases, (hits, cases)
