"""GPU fluid path through the C ABI: FluidModel::step against the reference's
own outputs (via the fixture) and the FP64 oracle, and the frame path's
particle bookkeeping and rendering."""
import ctypes as C

import numpy as np
import pytest

from atomorph_amd import _capi as A
from atomorph_amd import morph as M
from oracle import oracle_py as op
from tests import gpu_util as G
from tests.test_gpu_render import GpuRenderer
from tests.test_oracle import render_inputs

pytestmark = pytest.mark.gpu

# FP32 grid against FP64: positions within 1e-4 cell, velocities 1e-4, colours 1e-4
TOL = {0: 1e-4, 1: 1e-4, 2: 1e-4, 3: 1e-4, 11: 1e-4, 12: 1e-4, 13: 1e-4, 14: 1e-4}


def _model(n, gsx, gsy, W=64, H=64):
    lib = G.lib()
    f = C.c_void_p()
    G.check(lib.amh_fluid_create(C.byref(f), gsx, gsy, n, W, H, 2, G.stream()), "amh_fluid_create")
    return f


def _step_gpu(P, gsx, gsy, steps, radius):
    lib = G.lib()
    n = P.shape[0]
    f = _model(n, gsx, gsy)
    try:
        Q = np.ascontiguousarray(P.copy())
        G.check(lib.amh_fluid_set_particles(f, op.ptr(Q), G.stream()), "set")
        outs = []
        for k in range(steps):
            G.check(lib.amh_fluid_step(f, steps - 1 - k, radius, 0, G.stream()), "step")
            G.check(lib.amh_fluid_get_particles(f, op.ptr(Q), G.stream()), "get")
            outs.append(Q.copy())
        return outs
    finally:
        lib.amh_fluid_destroy(f)


def test_fluid_step_matches_reference_fixture(golden_functions):
    g = golden_functions
    gx, gy, fr = (int(g["fluid_meta"][0]), int(g["fluid_meta"][1]), float(g["fluid_meta"][2]))
    G.log("fluid step vs fixture")
    outs = _step_gpu(g["fluid_in"], gx, gy, 3, fr)
    for k in range(3):
        want = g["fluid_out"][k]
        for col, tol in TOL.items():
            d = np.abs(outs[k][:, col] - want[:, col]).max()
            assert d <= tol * (k + 1), (k, col, d)


@pytest.mark.parametrize("layout", ["uniform", "clustered", "cell_sorted"])
def test_fluid_step_matches_oracle_many_particles(layout):
    """uniform: every block's particles span the grid (direct global
    scatter); clustered: runs of 256 particles around one centre, as chain
    atoms are (the LDS-tiled scatter), with a few wide runs mixed in;
    cell_sorted: the clustered particles ordered by cell, so runs of lanes
    share a stencil and are summed across lanes before they scatter."""
    rng = np.random.default_rng(4)
    n, gsx, gsy = 20480, 276, 276
    P = np.zeros((n, 20))
    if layout == "uniform":
        P[:, 0] = rng.uniform(8, 268, n)
        P[:, 1] = rng.uniform(8, 268, n)
    else:
        centres = rng.uniform(20, 256, (n // 256, 2))
        spread = np.where(np.arange(n // 256) % 10 == 0, 60.0, 3.0)
        P[:, 0] = np.clip(np.repeat(centres[:, 0], 256) + rng.normal(0, 1, n) * np.repeat(spread, 256), 8, 268)
        P[:, 1] = np.clip(np.repeat(centres[:, 1], 256) + rng.normal(0, 1, n) * np.repeat(spread, 256), 8, 268)
    P[:, 2:4] = rng.normal(0, 0.2, (n, 2))
    P[:, 4] = P[:, 0] + rng.normal(0, 3, n)
    P[:, 5] = P[:, 1] + rng.normal(0, 3, n)
    P[:, 6] = 1.0
    P[:, 7:15] = rng.uniform(0, 1, (n, 8))
    P[:, 15] = rng.choice([0.1, 1.0], n)
    P[:, 16] = (rng.uniform(0, 1, n) < 0.9).astype(float)
    P[:, 17] = rng.integers(0, 2, n)
    if layout == "cell_sorted":
        cx, cy = (P[:, 0] - 0.5).astype(int), (P[:, 1] - 0.5).astype(int)
        P = P[np.lexsort((cy, cx))]
    G.log(f"fluid step 20k particles, {layout}")
    outs = _step_gpu(P, gsx, gsy, 3, 20.0)
    Q = P.copy()
    o = op.oracle()
    for k in range(3):
        o.or_fluid_step(op.ptr(Q), n, gsx, gsy, 2 - k, 20.0, 0)
        for col in (0, 1, 2, 3):
            d = np.abs(outs[k][:, col] - Q[:, col]).max()
            assert d <= 2e-3, (k, col, d)
        assert (outs[k][:, 16] == Q[:, 16]).all()


def _clustered(rng, n, lo, hi, shift=0.0):
    """Particles in runs of 256 around random centres in [lo, hi)^2 (moved by
    shift cells), with velocities, attractors and colours."""
    P = np.zeros((n, 20))
    centres = rng.uniform(lo, hi, (n // 256, 2)) + shift
    P[:, 0] = np.clip(np.repeat(centres[:, 0], 256) + rng.normal(0, 3, n), 8, hi + 20)
    P[:, 1] = np.clip(np.repeat(centres[:, 1], 256) + rng.normal(0, 3, n), 8, hi + 20)
    P[:, 2:4] = rng.normal(0, 0.2, (n, 2))
    P[:, 4] = P[:, 0] + rng.normal(0, 3, n)
    P[:, 5] = P[:, 1] + rng.normal(0, 3, n)
    P[:, 6] = 1.0
    P[:, 7:15] = rng.uniform(0, 1, (n, 8))
    P[:, 15] = rng.choice([0.1, 1.0], n)
    P[:, 16] = 1.0
    P[:, 17] = rng.integers(0, 2, n)
    return P


@pytest.mark.parametrize("grid", [276, 600])
def test_fluid_steps_after_new_particles_match_oracle(grid):
    """Work items exist only for tiles with entries, plus one zeroing item
    for each tile that had entries the step before: after a step of one
    particle set, a second set a few cells away (sharing the first one's
    node halos) steps as the oracle steps it from an empty grid.  276^2 has
    1,225 tiles (counts scanned in LDS), 600^2 has 5,625 (scanned on the
    device first)."""
    rng = np.random.default_rng(11)
    n = 20480
    A = _clustered(rng, n, 30, grid - 60)
    B = _clustered(rng, n, 30, grid - 60, shift=4.0)
    lib = G.lib()
    f = _model(n, grid, grid)
    try:
        G.log(f"fluid grid {grid}: one step of the first set, then the second")
        QA = np.ascontiguousarray(A.copy())
        G.check(lib.amh_fluid_set_particles(f, op.ptr(QA), G.stream()), "set A")
        G.check(lib.amh_fluid_step(f, 0, 20.0, 0, G.stream()), "step A")
        Q = np.ascontiguousarray(B.copy())
        G.check(lib.amh_fluid_set_particles(f, op.ptr(Q), G.stream()), "set B")
        outs = []
        for k in range(3):
            G.check(lib.amh_fluid_step(f, 2 - k, 20.0, 0, G.stream()), "step B")
            G.check(lib.amh_fluid_get_particles(f, op.ptr(Q), G.stream()), "get")
            outs.append(Q.copy())
    finally:
        lib.amh_fluid_destroy(f)
    R = B.copy()
    for k in range(3):
        op.oracle().or_fluid_step(op.ptr(R), n, grid, grid, 2 - k, 20.0, 0)
        for col in (0, 1, 2, 3, 11, 12, 13, 14):
            d = np.abs(outs[k][:, col] - R[:, col]).max()
            assert d <= 2e-3, (k, col, d)


def _fluid_frame(g, variant, t, fluidsteps, sp=1.0, reset=True, fluid=None, rend=None, atom_keys=None):
    """Render one fluid frame of the cfg1 chains through the C ABI (atom_keys:
    a device copy of the key points atom-major for the update, or None)."""
    lib = G.lib()
    cfg, keys, off, widths, order, nb, blob_rgba, (lag, slope) = render_inputs(g, variant)
    if rend is None:
        vals = {"bbox": [cfg.bbox_x1, cfg.bbox_y1, cfg.bbox_x2, cfg.bbox_y2], "motion": cfg.motion,
                "fading": cfg.fading, "show": cfg.show_blobs, "density": cfg.density, "feather": cfg.feather,
                "keep_bg": cfg.keep_background}
        rend = GpuRenderer(keys, g["chain_after"], off, widths, order, blob_rgba, vals, lag, slope)
    h = 2
    wtot = int(np.sum(widths))
    pts = g["chain_after"].reshape(-1, h)
    nch = len(widths)
    # activation ranks and blob sizes per key segment, as the engine builds them
    rank = np.zeros((h, wtot), np.int32)
    surf = np.zeros((h, nch), np.int32)
    for y in range(h):
        yn = (y + 1) % h
        for c, (o_, w) in enumerate(zip(off, widths)):
            f1 = (pts[o_:o_ + w, y] >> 48) & 2
            f2 = (pts[o_:o_ + w, yn] >> 48) & 2
            cls = np.where((f1 > 0) & (f2 > 0), 0, np.where(f1 > 0, 1, np.where(f2 > 0, 2, 3)))
            order_x = np.lexsort((np.arange(w), cls))
            rank[y, o_ + order_x] = np.arange(w)
            surf[y, c] = int((f1 > 0).sum())
    keep = [G.to_dev(rank.ravel()), G.to_dev(surf.ravel())]
    fa = A.FluidFrame(C.pointer(rend.args), keep[0].data_ptr(), keep[1].data_ptr(), None, sp, fluidsteps, 0,
                      None if atom_keys is None else atom_keys.data_ptr())
    if fluid is None:
        fluid = C.c_void_p()
        G.check(lib.amh_fluid_create(C.byref(fluid), int(64 * sp) + 20, int(64 * sp) + 20, wtot, 64, 64, h,
                                     G.stream()), "create")
        G.check(lib.amh_fluid_reset(fluid, sp, G.stream()), "reset")
    G.check(lib.amh_fluid_render(fluid, C.byref(fa), float(t), G.dptr(rend.out), G.stream()), "render")
    img = G.to_host(rend.out, np.uint32).view(np.uint8).reshape(64, 64, 4).copy()
    return img, fluid, rend, surf, rank


def test_fluid_frame_without_steps_reproduces_key_frame(golden_cfg1):
    """fluidsteps = 0 at t = 0: each of the |A| active particles of a chain
    sits on its source pixel (grid = pixel + 10, weight 1), so the frame is
    key frame 0 as get_pixel returns it."""
    g = golden_cfg1
    G.log("fluid frame t=0, no steps")
    img, fluid, rend, surf, rank = _fluid_frame(g, "default", 0.0, 0)
    try:
        key = np.ascontiguousarray(g["shown"][0]).view(np.uint8).reshape(64, 64, 4)
        assert (img == key).all()
        # active particles per chain = round((1-t)|A| + t|B|) at t = 0
        lib = G.lib()
        n = int(rank.shape[1])
        Q = np.zeros((n, 20))
        G.check(lib.amh_fluid_get_particles(fluid, op.ptr(Q), G.stream()), "get")
        widths = g["chain_w"]
        off = np.concatenate([[0], np.cumsum(widths)[:-1]])
        for c, (o_, w) in enumerate(zip(off, widths)):
            assert int(Q[o_:o_ + w, 16].sum()) == surf[0, c]
    finally:
        G.lib().amh_fluid_destroy(fluid)


def test_fluid_frames_advance_and_count(golden_cfg1):
    """Frames in order with 4 model steps each: the active count follows the
    interpolated blob areas and the image stays covered and opaque where the
    fluid is."""
    g = golden_cfg1
    fluid = rend = None
    widths = g["chain_w"]
    off = np.concatenate([[0], np.cumsum(widths)[:-1]])
    lib = G.lib()
    try:
        for i in range(16):
            t = i / 16
            G.log(f"fluid frame {i}")
            img, fluid, rend, surf, rank = _fluid_frame(g, "default", t, 4, fluid=fluid, rend=rend)
            n = int(rank.shape[1])
            Q = np.zeros((n, 20))
            G.check(lib.amh_fluid_get_particles(fluid, op.ptr(Q), G.stream()), "get")
            fk = int(t * 2)
            tl = (t - fk / 2) * 2
            for c, (o_, w) in enumerate(zip(off, widths)):
                want = int(np.floor((1 - tl) * surf[fk, c] + tl * surf[(fk + 1) % 2, c] + 0.5))
                assert int(Q[o_:o_ + w, 16].sum()) == want
            cover = int((img[..., 3] > 0).sum())
            assert cover > 0.5 * min(surf[0].sum(), surf[1].sum())
            assert np.isfinite(Q[:, :4]).all()
    finally:
        if fluid is not None:
            lib.amh_fluid_destroy(fluid)


def _renderer(g, variant, swap_tables=False):
    cfg, keys, off, widths, order, nb, blob_rgba, (lag, slope) = render_inputs(g, variant)
    vals = {"bbox": [cfg.bbox_x1, cfg.bbox_y1, cfg.bbox_x2, cfg.bbox_y2], "motion": cfg.motion,
            "fading": cfg.fading, "show": cfg.show_blobs, "density": cfg.density, "feather": cfg.feather,
            "keep_bg": cfg.keep_background}
    if swap_tables:
        lag, slope = slope, lag
    return GpuRenderer(keys, g["chain_after"], off, widths, order, blob_rgba, vals, lag, slope)


def test_fluid_perlin_lag_slope_kept_between_frames(golden_cfg1):
    """A model keeps each atom's PERLIN lag and slope from frame to frame of
    a key segment and computes them again when the noise tables change: the
    colour targets (R, G, B, A) its update sets equal those of a fresh model
    at the same time, and they depend on the tables."""
    g = golden_cfg1
    variant = "spline_perlin_bg_feather2"
    assert int(g[f"render_{variant}_opts"][1]) == 6  # PERLIN
    lib = G.lib()
    n = int(np.sum(g["chain_w"]))
    made = []

    def particles(f):
        Q = np.zeros((n, 20))
        G.check(lib.amh_fluid_get_particles(f, op.ptr(Q), G.stream()), "get")
        return Q

    def same_targets(QA, QB):
        both = (QA[:, 16] > 0) & (QB[:, 16] > 0)
        assert both.sum() > 0.5 * (QB[:, 16] > 0).sum()
        return (QA[both, 7:11] == QB[both, 7:11]).all()

    try:
        rend = _renderer(g, variant)
        swapped = _renderer(g, variant, swap_tables=True)
        fa = None
        for t in (0.0, 0.1, 0.2, 0.3):  # key segment 0: the lag and slope of t = 0 are kept
            G.log(f"perlin memo: frame t={t}")
            _, fa, _, _, _ = _fluid_frame(g, variant, t, 2, fluid=fa, rend=rend)
            if not made:
                made.append(fa)
        _, fb, _, _, _ = _fluid_frame(g, variant, 0.3, 2, rend=rend)
        made.append(fb)
        assert same_targets(particles(fa), particles(fb))
        # other tables: the kept values no longer hold
        _, fa, _, _, _ = _fluid_frame(g, variant, 0.35, 2, fluid=fa, rend=swapped)
        _, fc, _, _, _ = _fluid_frame(g, variant, 0.35, 2, rend=swapped)
        made.append(fc)
        _, fd, _, _, _ = _fluid_frame(g, variant, 0.35, 2, rend=rend)
        made.append(fd)
        assert same_targets(particles(fa), particles(fc))
        assert not same_targets(particles(fd), particles(fc))  # the tables matter at t = 0.35
    finally:
        for f in made:
            lib.amh_fluid_destroy(f)


@pytest.mark.parametrize("variant", ["default", "spline_perlin_bg_feather2"])
def test_fluid_update_from_atom_major_keys(golden_cfg1, variant):
    """amh_fluid_atom_keys transposes the key points atom-major, and frames
    whose update reads an atom's key points from that copy equal frames
    that read them row by row: the same active particles with the same
    colour targets, and (as between any two runs: tiles split over several
    work items add their node sums with FP32 atomics, in no fixed order)
    images within 1 LSB and positions and velocities within 1e-4."""
    g = golden_cfg1
    lib = G.lib()
    T = G.torch()
    h, n = 2, int(np.sum(g["chain_w"]))
    rend = _renderer(g, variant)
    keys = G.zeros(n * h, T.int64)
    G.check(lib.amh_fluid_atom_keys(G.dptr(rend.keep[1]), h, n, G.dptr(keys), G.stream()), "atom keys")
    want = np.ascontiguousarray(G.key_major(g["chain_after"], h)).view(np.uint64).reshape(h, n).T.ravel()
    assert np.array_equal(G.to_host(keys, np.uint64), want)
    made = []
    try:
        fa = fb = None
        for t in (0.0, 0.3, 0.55, 0.8):
            G.log(f"atom-major keys: {variant} t={t}")
            ia, fa, _, _, _ = _fluid_frame(g, variant, t, 2, fluid=fa, rend=rend)
            if len(made) == 0:
                made.append(fa)
            ib, fb, _, _, _ = _fluid_frame(g, variant, t, 2, fluid=fb, rend=rend, atom_keys=keys)
            if len(made) == 1:
                made.append(fb)
            assert np.abs(ia.astype(np.int32) - ib.astype(np.int32)).max() <= 1, t
            QA, QB = np.zeros((n, 20)), np.zeros((n, 20))
            G.check(lib.amh_fluid_get_particles(fa, op.ptr(QA), G.stream()), "get")
            G.check(lib.amh_fluid_get_particles(fb, op.ptr(QB), G.stream()), "get")
            assert np.array_equal(QA[:, 16], QB[:, 16]), t
            act = QA[:, 16] > 0
            assert act.sum() > 0
            assert np.array_equal(QA[act, 7:11], QB[act, 7:11]), t
            assert np.abs(QA[:, :4] - QB[:, :4]).max() <= 1e-4, t
    finally:
        for f in made:
            lib.amh_fluid_destroy(f)


def test_engine_fluid_render():
    """am::morph with set_fluid: frames render through the fluid path and
    differ from the atom path."""
    from atomorph_amd import synth
    frames = synth.workload("ellipses64")
    m = M.Morph()
    m.set(seed=1, cycle_length=1 << 14)
    for i, f in enumerate(frames):
        m.add_frame_rgba(i, f)
    m.set_resolution(64, 64)
    m.run(1)
    m.run(300)
    m.next_state()
    assert m.run(1) == M.STATE_ATOM_MORPHING
    m.run(4)
    atoms = m.render(0.3)
    m.set(fluid=5, feather=1)
    imgs = [m.render(i / 16) for i in range(16)]
    assert all((im[..., 3] > 0).sum() > 100 for im in imgs)
    assert not (imgs[5] == atoms).all()


def test_fluid_activation_ranks():
    """amh_fluid_ranks: per key segment and chain the ranks are a
    permutation of the slots, classes first (HAS_FLUID in both rows, only the
    source, only the destination, neither), and within a class the order of
    the keyed permutation -- equal to a host replay of the definition."""
    T = G.torch()
    lib = G.lib()
    rng = np.random.default_rng(21)
    widths = np.array([1, 2, 3, 64, 255, 256, 257, 700, 1500], np.int64)
    h = 3
    off = np.concatenate([[0], np.cumsum(widths)[:-1]]).astype(np.int64)
    wtot = int(widths.sum())
    flags = rng.integers(0, 4, (h, wtot))
    pts = op.pack_points(rng.integers(0, 64, (h, wtot)), rng.integers(0, 64, (h, wtot)), 0, 0, flags)
    seed = 0x1234ABCD
    # every device buffer stays referenced until the kernel has run: a
    # temporary freed early hands its memory to the next allocation
    dev = G.to_dev(pts.ravel())
    off_d = G.to_dev(off)
    w_d = G.to_dev(widths.astype(np.int32))
    rank = G.zeros(h * wtot, T.int32)
    assert off_d.data_ptr() != w_d.data_ptr()
    G.log(f"fluid ranks: {len(widths)} chains, wtot {wtot}, {h} key frames")
    G.check(lib.amh_fluid_ranks(G.dptr(dev), h, wtot, G.dptr(off_d), G.dptr(w_d), len(widths), seed, G.dptr(rank),
                                G.stream()), "amh_fluid_ranks")
    got = G.to_host(rank, np.int32).reshape(h, wtot)
    orc = op.oracle()
    for y in range(h):
        yn = (y + 1) % h
        for c, (o, w) in enumerate(zip(off, widths)):
            f1 = (flags[y, o:o + w] & 2) > 0
            f2 = (flags[yn, o:o + w] & 2) > 0
            cls = np.where(f1 & f2, 0, np.where(f1, 1, np.where(f2, 2, 3)))
            key = orc.or_hash64((seed ^ (c << 32) ^ y) & ((1 << 64) - 1))
            base = np.concatenate([[0], np.cumsum(np.bincount(cls, minlength=4))[:-1]])
            want = np.zeros(w, np.int64)
            for p in range(w):
                x = orc.or_perm_index(p, int(w), key)
                want[x] = base[cls[x]]
                base[cls[x]] += 1
            seg = got[y, o:o + w]
            assert np.array_equal(np.sort(seg), np.arange(w)), (y, c)
            assert np.array_equal(seg, want), (y, c)


# ------------------------------------------------------------ fluid frames
def _fluid_fixture():
    import os
    return dict(np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fluid.npz")))


def test_fluid_draw_within_one_lsb_of_reference():
    """The drawing half of draw_fluid (colour relax, bilinear splat, FP64
    resolve, outward feather, background) on the particles the compiled
    reference had after step_fluid, at six frames of the cfg1 fluid morph
    with feather 0-3 and keep_background on and off: frames within 1 LSB of
    the reference's, relaxed colours and maturity equal to its particles."""
    fx = _fluid_fixture()
    lib = G.lib()
    T = G.torch()
    n = fx["fl_pre"].shape[1]
    keys = G.to_dev(np.ascontiguousarray(fx["fl_shown"]).ravel())
    lag, slope = G.to_dev(op.perlin_perm(1)), G.to_dev(op.perlin_perm(2))
    out = G.zeros(64 * 64, T.int32)
    bb = [int(v) for v in fx["fl_bbox"]]
    fluid = C.c_void_p()
    G.check(lib.amh_fluid_create(C.byref(fluid), 84, 84, n, 64, 64, 2, G.stream()), "create")
    try:
        for k, i in enumerate(fx["fl_index"]):
            feather, keep = (int(v) for v in fx["fl_opts"][k])
            t = float(fx["fl_t"][k])
            G.log(f"fluid draw: frame {i} t={t} feather {feather} keep_background {keep}")
            args = A.RenderArgs(64, 64, 2, 0, *bb, 4, 6, 0, 1, feather, keep, 0, 0, 0, keys.data_ptr(), None, None,
                                None, None, None, lag.data_ptr(), slope.data_ptr())
            fa = A.FluidFrame(C.pointer(args), None, None, None, 1.0, 3, 0)
            P = np.ascontiguousarray(fx["fl_pre"][k])
            G.check(lib.amh_fluid_set_particles(fluid, op.ptr(P), G.stream()), "set")
            G.check(lib.amh_fluid_draw(fluid, C.byref(fa), t, G.dptr(out), G.stream()), "amh_fluid_draw")
            img = G.to_host(out, np.uint32).view(np.uint8).reshape(64, 64, 4)
            want = fx["fl_img"][k]
            d = np.abs(img.astype(np.int32) - want.astype(np.int32))
            G.log(f"  worst channel difference {int(d.max())}, differing channels {int((d > 0).sum())}, "
                  f"drawn pixels {int((want[..., 3] > 0).sum())}")
            assert d.max() <= 1, (i, int(d.max()))
            Q = np.zeros((n, 20))
            G.check(lib.amh_fluid_get_particles(fluid, op.ptr(Q), G.stream()), "get")
            post = fx["fl_post"][k]
            assert np.allclose(Q[:, 11:15], post[:, :4], rtol=0, atol=1e-12), i
            assert (Q[:, 17] == post[:, 4]).all(), i
            assert (Q[:, :2] == P[:, :2]).all()  # drawing moves nothing
    finally:
        lib.amh_fluid_destroy(fluid)


@pytest.mark.parametrize("steps", [4, 3])
def test_fluid_frames_replayed_from_a_graph_equal_eager_frames(golden_cfg1, steps):
    """amh_fluid_render on a stream of its own captures a frame's launches
    into a hipGraph once and replays it, the per-frame values coming from
    the parameter block: 12 frames (with other GPU work and host syncs
    between the replays) equal the same frames launched one by one within
    1 LSB, and the particles agree within 1e-4.  With an odd number of
    steps per frame the bins' count parity alternates from frame to frame,
    so the frames replay two graphs in turn."""
    g = golden_cfg1
    lib = G.lib()
    T = G.torch()
    cfg, keys, off, widths, order, nb, blob_rgba, (lag, slope) = render_inputs(g, "spline_perlin_bg_feather2")
    vals = {"bbox": [cfg.bbox_x1, cfg.bbox_y1, cfg.bbox_x2, cfg.bbox_y2], "motion": cfg.motion,
            "fading": cfg.fading, "show": cfg.show_blobs, "density": cfg.density, "feather": cfg.feather,
            "keep_bg": cfg.keep_background}
    rend = GpuRenderer(keys, g["chain_after"], off, widths, order, blob_rgba, vals, lag, slope)
    h, wtot = 2, int(np.sum(widths))
    pts = g["chain_after"].reshape(-1, h)
    rank = np.zeros((h, wtot), np.int32)
    surf = np.zeros((h, len(widths)), np.int32)
    for y in range(h):
        for c, (o_, w) in enumerate(zip(off, widths)):
            f1 = (pts[o_:o_ + w, y] >> 48) & 2
            f2 = (pts[o_:o_ + w, (y + 1) % h] >> 48) & 2
            cls = np.where((f1 > 0) & (f2 > 0), 0, np.where(f1 > 0, 1, np.where(f2 > 0, 2, 3)))
            rank[y, o_ + np.lexsort((np.arange(w), cls))] = np.arange(w)
            surf[y, c] = int((f1 > 0).sum())
    keep = [G.to_dev(rank.ravel()), G.to_dev(surf.ravel())]
    fa = A.FluidFrame(C.pointer(rend.args), keep[0].data_ptr(), keep[1].data_ptr(), None, 1.0, steps, 0)
    stream = C.c_void_p()
    G.check(lib.amh_stream_create(C.byref(stream)), "stream")
    models = []
    try:
        for graphs in (1, 0):
            f = C.c_void_p()
            G.check(lib.amh_fluid_create(C.byref(f), 84, 84, wtot, 64, 64, h, stream), "create")
            G.check(lib.amh_fluid_reset(f, 1.0, stream), "reset")
            G.check(lib.amh_fluid_use_graphs(f, graphs), "graphs")
            models.append(f)
        outs = [T.zeros(12 * 64 * 64, dtype=T.int32, device="cuda") for _ in range(2)]
        G.sync()
        for i in range(12):
            t = i / 12
            G.log(f"fluid graph replay frame {i} ({steps} steps)")
            for f, o in zip(models, outs):
                G.check(lib.amh_fluid_render(f, C.byref(fa), t, C.c_void_p(o.data_ptr() + 4 * 64 * 64 * i), stream),
                        "render")
            G.check(lib.amh_stream_sync(stream), "sync")
            (T.ones(1 << 20, device="cuda") * 2.0).sum().item()  # other work between the replays
        a = G.to_host(outs[0], np.uint32).view(np.uint8).reshape(12, 64, 64, 4)
        b = G.to_host(outs[1], np.uint32).view(np.uint8).reshape(12, 64, 64, 4)
        for i in range(12):
            d = np.abs(a[i].astype(np.int32) - b[i].astype(np.int32))
            G.log(f"  frame {i}: differing channels {int((d > 0).sum())}")
            assert d.max() <= 1, (i, int(d.max()))
            assert (a[i][..., 3] > 0).sum() > 500
        # the float atomics of the model steps may add in another order from
        # one run to the next, so the particles agree to rounding, not bits
        Q = [np.zeros((wtot, 20)) for _ in range(2)]
        for f, q in zip(models, Q):
            G.check(lib.amh_fluid_get_particles(f, op.ptr(q), stream), "get")
        assert np.abs(Q[0][:, :4] - Q[1][:, :4]).max() <= 1e-4
        assert np.array_equal(Q[0][:, 16], Q[1][:, 16])
    finally:
        for f in models:
            lib.amh_fluid_destroy(f)
        lib.amh_stream_destroy(stream)
