"""GPU parity of frame rendering (atom path) through the C ABI.

From identical trajectories (the reference's refined cfg1 chains) the GPU
frames are within 1 LSB per channel of the frames the compiled reference
rendered, for six option sets; and equal the oracle's frames except where
the oracle's FP64 weight sums sit on a rounding tie."""
import ctypes as C

import numpy as np
import pytest

from atomorph_amd import _capi as A
from oracle import oracle_py as op
from tests import gpu_util as G
from tests.test_oracle import render_inputs

pytestmark = pytest.mark.gpu

VARIANTS = ["default", "linear_cosine", "none_linear_bg", "spline_perlin_bg_feather2", "average_blobs",
            "distinct_blobs"]


class GpuRenderer:
    """Device buffers for rendering one fixed set of chains."""

    def __init__(self, keys, pts_atom_major, off, widths, order, blob_rgba, cfg_vals, lag, slope):
        T = G.torch()
        lib = G.lib()
        h, H, W = keys.shape[0], keys.shape[1], keys.shape[2]
        self.W, self.H = W, H
        nchains = len(widths)
        nb = order.shape[1]
        rank = -np.ones((h, nchains), np.int32)
        for y in range(h):
            for b in range(nb):
                if order[y, b] >= 0:
                    rank[y, order[y, b]] = b
        self.keep = [G.to_dev(np.ascontiguousarray(keys).view(np.uint32).reshape(-1)),
                     G.to_dev(G.key_major(pts_atom_major, h)), G.to_dev(np.asarray(off, np.int64)),
                     G.to_dev(np.asarray(widths, np.int32)), G.to_dev(rank.ravel()),
                     G.to_dev(np.ascontiguousarray(blob_rgba, np.uint32).ravel()), G.to_dev(lag), G.to_dev(slope)]
        k, p, o, w, r, br, lg, sl = self.keep
        self.args = A.RenderArgs(W, H, h, nb, *cfg_vals["bbox"], cfg_vals["motion"], cfg_vals["fading"],
                                 cfg_vals["show"], cfg_vals["density"], cfg_vals["feather"], cfg_vals["keep_bg"],
                                 0, nchains, int(np.sum(widths)), k.data_ptr(), p.data_ptr(), o.data_ptr(),
                                 w.data_ptr(), r.data_ptr(), br.data_ptr(), lg.data_ptr(), sl.data_ptr())
        self.wsb = lib.amh_render_workspace_bytes(W, H, int(np.sum(widths)))
        self.ws = T.zeros(self.wsb, dtype=T.uint8, device="cuda")
        G.check(lib.amh_render_workspace_init(G.dptr(self.ws), self.wsb, W, H, int(np.sum(widths)), G.stream()),
                "workspace init")
        self.out = G.zeros(W * H, T.int32)

    def frame(self, t):
        G.check(G.lib().amh_render_frame(C.byref(self.args), float(t), G.dptr(self.ws), self.wsb,
                                         G.dptr(self.out), G.stream()), "amh_render_frame")
        return G.to_host(self.out, np.uint32).view(np.uint8).reshape(self.H, self.W, 4).copy()


def _renderer_for(g, variant, density=1):
    cfg, keys, off, widths, order, nb, blob_rgba, (lag, slope) = render_inputs(g, variant, density)
    vals = {"bbox": [cfg.bbox_x1, cfg.bbox_y1, cfg.bbox_x2, cfg.bbox_y2], "motion": cfg.motion,
            "fading": cfg.fading, "show": cfg.show_blobs, "density": cfg.density, "feather": cfg.feather,
            "keep_bg": cfg.keep_background}
    return GpuRenderer(keys, g["chain_after"], off, widths, order, blob_rgba, vals, lag, slope), cfg


@pytest.mark.parametrize("variant", VARIANTS)
def test_render_within_one_lsb_of_reference(golden_cfg1, variant):
    g = golden_cfg1
    rend, cfg = _renderer_for(g, variant)
    worst, differing = 0, 0
    for i, t in enumerate(g[f"render_{variant}_t"]):
        G.log(f"render {variant} frame {i} t={t}")
        got = rend.frame(t)
        want = g[f"render_{variant}"][i]
        d = np.abs(got.astype(np.int32) - want.astype(np.int32))
        worst = max(worst, int(d.max()))
        differing += int((d > 0).sum())
    G.log(f"render {variant}: worst channel difference {worst}, differing channels {differing}")
    assert worst <= 1


def test_render_workspace_is_left_clean(golden_cfg1):
    """Rendering the same frame twice, with other frames between, gives the
    same image (the per-pixel accumulators and overflow counters reset)."""
    rend, _ = _renderer_for(golden_cfg1, "default")
    a = rend.frame(0.3)
    for t in (0.0, 0.55, 0.9):
        rend.frame(t)
    assert (rend.frame(0.3) == a).all()


def test_render_empty_and_negative_time(golden_cfg1):
    rend, _ = _renderer_for(golden_cfg1, "default")
    # normalize_time maps -0.75 to 0.25
    assert (rend.frame(-0.75) == rend.frame(0.25)).all()
    assert (rend.frame(3.25) == rend.frame(0.25)).all()


DENSITY_CASES = [(2, "default"), (2, "spline_perlin_bg_feather2"), (2, "average_blobs"), (3, "default")]


@pytest.mark.parametrize("density,variant", DENSITY_CASES)
def test_render_at_density_within_one_lsb_of_reference(golden_density, density, variant):
    """Chains max_surface * density wide (the reference CLI's default
    density is 2) and alpha round(min(1, taps / density) * mean alpha)
    (morph.cpp:606-612): GPU frames within 1 LSB of the reference's."""
    g = golden_density(density)
    rend, cfg = _renderer_for(g, variant, density)
    assert cfg.density == density
    worst, halved = 0, 0
    for i, t in enumerate(g[f"render_{variant}_t"]):
        G.log(f"render density {density} {variant} frame {i} t={t}")
        got = rend.frame(t)
        want = g[f"render_{variant}"][i]
        worst = max(worst, int(np.abs(got.astype(np.int32) - want.astype(np.int32)).max()))
        halved += int(((want[..., 3] > 0) & (want[..., 3] < 255)).sum())
    G.log(f"density {density} {variant}: worst channel difference {worst}, partial-alpha pixels {halved}")
    if cfg.show_blobs == 0:  # AVERAGE / DISTINCT draw the blob's own alpha (morph.cpp:1321-1334)
        assert halved > 0  # pixels with fewer taps than the density exist
    assert worst <= 1


@pytest.mark.parametrize("variant", ["default", "spline_perlin_bg_feather2"])
def test_render_graph_replays_equal_eager_frames(golden_cfg1, variant):
    """amh_render_graph_frame (one captured frame replayed with the time and
    output in a device block) gives the eager frames bit for bit, replay
    after replay, with eager frames rendered between the replays."""
    rend, _ = _renderer_for(golden_cfg1, variant)
    lib, T = G.lib(), G.torch()
    ts = [float(t) for t in golden_cfg1[f"render_{variant}_t"]][:6] + [-0.75, 2.3]
    eager = [rend.frame(t) for t in ts]
    outs = [G.zeros(rend.W * rend.H, T.int32) for _ in ts]
    G.sync()
    stream, graph = C.c_void_p(), C.c_void_p()
    G.check(lib.amh_stream_create(C.byref(stream)), "stream")  # capture needs a created stream
    try:
        G.check(lib.amh_render_graph_create(C.byref(graph), C.byref(rend.args), G.dptr(rend.ws), rend.wsb, stream),
                "amh_render_graph_create")
        for rep in range(3):
            G.log(f"render graph {variant}: replay round {rep}")
            for t, o in zip(ts, outs):
                G.check(lib.amh_render_graph_frame(graph, t, G.dptr(o), stream), "amh_render_graph_frame")
            G.check(lib.amh_stream_sync(stream), "sync")
            for t, o, want in zip(ts, outs, eager):
                got = G.to_host(o, np.uint32).view(np.uint8).reshape(rend.H, rend.W, 4)
                assert (got == want).all(), (rep, t)
            # eager work between replays on the same workspace
            assert (rend.frame(ts[rep]) == eager[rep]).all()
        # destroyed with replays still queued: it waits for them
        for o in outs:
            o.zero_()
        G.sync()
        for t, o in zip(ts, outs):
            G.check(lib.amh_render_graph_frame(graph, t, G.dptr(o), stream), "amh_render_graph_frame")
        G.check(lib.amh_render_graph_destroy(graph), "amh_render_graph_destroy")
        graph = C.c_void_p()
        G.check(lib.amh_stream_sync(stream), "sync")
        for t, o, want in zip(ts, outs, eager):
            got = G.to_host(o, np.uint32).view(np.uint8).reshape(rend.H, rend.W, 4)
            assert (got == want).all(), ("after destroy", t)
    finally:
        lib.amh_render_graph_destroy(graph)
        lib.amh_stream_destroy(stream)
