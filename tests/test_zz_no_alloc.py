"""The allocation-free frame render (amh_render_frame_noalloc) against the
workspace form (amh_render_frame) through the C ABI: identical frames on two
calls in a row, the second with a smaller input, and again after it; sizes
beyond the fixed workspace are refused with an error."""
import ctypes as C

import numpy as np
import pytest

from atomorph_amd import _capi as A
from oracle import oracle_py as op
from tests import gpu_util as G
from tests.test_gpu_render import _renderer_for

pytestmark = pytest.mark.gpu


def _noalloc(rend, t):
    out = G.zeros(rend.W * rend.H, G.torch().int32)
    G.check(G.lib().amh_render_frame_noalloc(C.byref(rend.args), float(t), G.dptr(out), G.stream()),
            "amh_render_frame_noalloc")
    return G.to_host(out, np.uint32).view(np.uint8).reshape(rend.H, rend.W, 4).copy()


class SyntheticRenderer:
    """Random chains over W x H key frames (two keys, TEXTURE, PERLIN fades)."""

    def __init__(self, W, H, nchains, seed, feather=0):
        T = G.torch()
        lib = G.lib()
        rng = np.random.default_rng(seed)
        h = 2
        widths = rng.integers(50, 4000, nchains).astype(np.int64)
        wtot = int(widths.sum())
        off = np.concatenate([[0], np.cumsum(widths)[:-1]]).astype(np.int64)
        keys = rng.integers(0, 1 << 32, (h, H * W), dtype=np.uint64).astype(np.uint32)
        cx = rng.integers(20, W - 20, nchains)
        cy = rng.integers(20, H - 20, nchains)
        pts = np.zeros((h, wtot), np.uint64)
        for y in range(h):
            xs = np.clip(np.repeat(cx, widths) + rng.integers(-15, 16, wtot), 0, W - 1)
            ys = np.clip(np.repeat(cy, widths) + rng.integers(-15, 16, wtot), 0, H - 1)
            pts[y] = op.pack_points(xs, ys, rng.integers(0, 256, wtot), rng.integers(0, 256, wtot), 3)
        rank = np.tile(np.arange(nchains, dtype=np.int32), (h, 1))
        lag, slope = op.perlin_perm(seed), op.perlin_perm(seed + 1)
        self.W, self.H, self.wtot = W, H, wtot
        self.keep = [G.to_dev(keys.ravel()), G.to_dev(pts.ravel()), G.to_dev(off), G.to_dev(widths.astype(np.int32)),
                     G.to_dev(rank.ravel()), G.to_dev(lag), G.to_dev(slope)]
        k, p, o, w, r, lg, sl = self.keep
        self.args = A.RenderArgs(W, H, h, nchains, 0, 0, W - 1, H - 1, 4, 6, 0, 1, feather, 1, 0, nchains, wtot,
                                 k.data_ptr(), p.data_ptr(), o.data_ptr(), w.data_ptr(), r.data_ptr(), None,
                                 lg.data_ptr(), sl.data_ptr())
        self.wsb = lib.amh_render_workspace_bytes(W, H, wtot)
        self.ws = T.zeros(self.wsb, dtype=T.uint8, device="cuda")
        G.check(lib.amh_render_workspace_init(G.dptr(self.ws), self.wsb, W, H, wtot, G.stream()), "workspace init")
        self.out = G.zeros(W * H, T.int32)

    def frame(self, t):
        G.check(G.lib().amh_render_frame(C.byref(self.args), float(t), G.dptr(self.ws), self.wsb, G.dptr(self.out),
                                         G.stream()), "amh_render_frame")
        return G.to_host(self.out, np.uint32).view(np.uint8).reshape(self.H, self.W, 4).copy()


@pytest.mark.parametrize("feather", [0, 2])
def test_noalloc_render_equals_workspace_render(golden_cfg1, feather):
    lib = G.lib()
    big = SyntheticRenderer(512, 384, 60, seed=7, feather=feather)
    assert lib.amh_render_workspace_bytes(512, 384, big.wtot) <= lib.amh_render_noalloc_max_bytes()
    small, _ = _renderer_for(golden_cfg1, "spline_perlin_bg_feather2" if feather else "default")
    G.log(f"noalloc: big 512x384 {big.wtot} atoms, then cfg1 64x64, feather {feather}")
    for t in (0.3, 0.8):
        want_big = big.frame(t)
        got_big = _noalloc(big, t)
        assert (got_big == want_big).all(), ("big", t, int((got_big != want_big).sum()))
        assert want_big[..., 3].any()
        want_small = small.frame(t)
        got_small = _noalloc(small, t)  # second call, smaller input, same scratch
        assert (got_small == want_small).all(), ("small", t, int((got_small != want_small).sum()))
    # the larger input again after the smaller one: nothing left over
    assert (_noalloc(big, 0.3) == big.frame(0.3)).all()


def test_noalloc_render_refuses_sizes_beyond_its_workspace(golden_cfg1):
    lib = G.lib()
    rend, _ = _renderer_for(golden_cfg1, "default")
    args = A.RenderArgs.from_buffer_copy(rend.args)
    args.width, args.height = 4096, 4096
    out = G.zeros(64 * 64, G.torch().int32)
    rc = lib.amh_render_frame_noalloc(C.byref(args), 0.5, G.dptr(out), G.stream())
    assert rc == -1
    assert b"workspace" in lib.amh_last_error()
