"""Greedy blob detection and dust unification on the host (CPU only), the
path the engine takes for non-default blob options (thread.cpp:225-596).

- The blob lists of every key frame -- order, running means, starting pixel,
  sizes and membership -- equal the compiled reference's (fixture
  tests/golden/blobify.npz) for nine option sets: thresholds below 1, a
  size cap, minimum sizes (dust unification and the threshold-free retry of
  undersized blobs), preferred blob counts, a grip box, RGB and HSP stored
  colours, and a key frame without pixels (first or middle).
- At threshold 1 with the other options at their defaults the partition is
  the 4-connected components of the foreground (scipy's labelling), whatever
  the seed: the reduction the device segmentation relies on.
- Without dust unification every blob is 4-connected, and every pixel
  belongs to exactly one blob.
"""
import numpy as np
import pytest
from scipy import ndimage

from atomorph_amd import morph as M

CASES = range(9)


@pytest.fixture(scope="module")
def fx():
    import os
    p = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "blobify.npz")
    return dict(np.load(p))


def _run(fx, c, **over):
    mx, mn, num, samples, grip, _, seed = (int(v) for v in fx["bf_opts"][c])
    W, H, nf = (int(v) for v in fx["bf_size"])
    kw = dict(threshold=float(fx["bf_threshold"][c]), max_size=mx, min_size=mn, number=num, box_samples=samples,
              box_grip=grip, seed=seed)
    kw.update(over)
    return M.blobify_frames(fx["bf_stored"][c].reshape(nf, H, W), **kw)


@pytest.mark.parametrize("c", CASES)
def test_greedy_blobs_equal_the_reference(fx, c):
    bop, blobs = _run(fx, c)
    nf = int(fx["bf_size"][2])
    base = int(fx["bf_nblobs"][:c].sum())
    for f in range(nf):
        n = int(fx["bf_nblobs"][c, f])
        assert len(blobs[f]) == n
        assert np.array_equal(bop[f], fx["bf_bop"][c, f].astype(np.int32))
        for b in range(n):
            k = base + int(fx["bf_nblobs"][c, :f].sum()) + b
            assert blobs[f][b]["group"] == int(fx["bf_groups"][k])
            assert blobs[f][b]["size"] == int(fx["bf_sizes"][k])
            # running means bit for bit: same merge order, same arithmetic
            assert np.array_equal(blobs[f][b]["means"], fx["bf_means"][k]), (f, b)


def _partition(lab):
    return {tuple(np.flatnonzero(lab == v)) for v in np.unique(lab[lab >= 0])}


@pytest.mark.parametrize("seed", [0, 1, 77])
def test_threshold_one_gives_connected_components(fx, seed):
    W, H, nf = (int(v) for v in fx["bf_size"])
    for c in CASES:
        bop, _ = _run(fx, c, threshold=1.0, max_size=M.SIZE_MAX, min_size=1, number=1, seed=seed)
        for f in range(nf):
            present = (fx["bf_stored"][c, f] >> 24) != 0
            lab, _ = ndimage.label(present.reshape(H, W), structure=[[0, 1, 0], [1, 1, 1], [0, 1, 0]])
            want = lab.reshape(-1) - 1
            assert _partition(bop[f]) == _partition(want)


@pytest.mark.parametrize("threshold", [0.0, 0.05, 0.2])
def test_blobs_are_connected_and_cover_the_foreground(fx, threshold):
    W, H, nf = (int(v) for v in fx["bf_size"])
    rng = np.random.default_rng(3)
    frames = fx["bf_stored"][rng.integers(0, len(CASES))]
    bop, blobs = M.blobify_frames(frames.reshape(nf, H, W), threshold=threshold, seed=5)
    for f in range(nf):
        present = (frames[f] >> 24) != 0
        assert np.array_equal(bop[f] >= 0, present)
        sizes = np.bincount(bop[f][bop[f] >= 0], minlength=len(blobs[f]))
        assert [b["size"] for b in blobs[f]] == sizes.tolist()
        for b in range(len(blobs[f])):
            _, n = ndimage.label((bop[f] == b).reshape(H, W), structure=[[0, 1, 0], [1, 1, 1], [0, 1, 0]])
            assert n == 1
