#!/usr/bin/env python3
"""Benchmark of the MI355X morph engine on one GPU.

One step = one full morph job on the 1024^2 workload (4 RGBA8 key frames of
150 noise-thresholded blobs each): blob detection, blob matching, chain build,
learned warm start (MLP + sort-and-match), batched optimal-transport
refinement (hipGraph replays) and 60 output frames rendered into HBM with the
fluid model on (10 steps per frame on a 256^2 + 20 FP32 grid).  The key
frames are uploaded before the timed region.

Printed (one JSON line): frames/s of the whole job, plus
  time_to_cost  plain GPU descent (one chain at a time) run to convergence vs
                the batched persistent descent, the learned warm start +
                batched descent, and the batched descent with nearest-
                neighbour partners, timed until they reach the plain
                descent's final cost;
  roofline      the dominant kernel's achieved bytes/s against 8 TB/s, from
                per-kernel device times of one extra profiled job;
  mlp           the warm start's matrix-core forward against the BF16 peak;
  cpu_baseline  the compiled reference (oracle/_ref) on this host, one core,
                on a bounded sample of the same job, scaled to frames/s.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FRAMES_OUT = 60
PROFILED = ["k_blob_stats", "warm_start", "mlp_forward", "refine_graph", "k_refine_pass", "render_frame", "k_prepare",
            "k_splat", "k_composite", "fluid_render", "k_fluid_update", "fluid_sort", "k_fluid_splat", "fluid_step",
            "fluid_bins", "k_p2g", "k_forces", "k_g2p", "k_advect"]
MLP_FLOP_PER_ROW = 2 * (8 * 128 + 3 * 128 * 128 + 128 * 2)
MFMA_BF16_PEAK_TFLOPS = 2500.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_morph(M, frames, seed=1):
    m = M.Morph()
    m.set(seed=seed, cycle_length=1 << 16)
    for i, f in enumerate(frames):
        m.add_frame_rgba(i, f)
    h, w = frames[0].shape[:2]
    m.set_resolution(w, h)
    return m


def to_morphing(m, M, match_steps):
    m.run(1)                        # blob detection (+ unification)
    m.run(match_steps + 1)          # matching
    m.next_state()
    st = m.run(1)                   # chain build
    if st != M.STATE_ATOM_MORPHING:
        raise RuntimeError(f"engine did not reach ATOM_MORPHING ({st}): {m.error()}")


def spin_up():
    """Keeps the GPU busy for a few ms so that its clocks are up when a timed
    leg starts, as they are inside a job (the legs start after host work)."""
    import torch
    torch.cuda._sleep(20_000_000)
    torch.cuda.synchronize()


def time_to_cost(m, M, learned=None, log_every=20):
    """Plain descent to convergence, then the batched descent, the learned
    warm start followed by the batched descent, and the nearest-neighbour
    descent, from the same initial chains until they reach the plain final
    cost.  Device time of the warm start and refinement steps only (HIP
    events on the engine's compute stream); each leg is run for one
    untimed step first, so one-time costs (kernel code loading, hipGraph
    instantiation) stay out of it, and starts with the GPU clocks up."""
    pts0, widths = m.chains()
    nch = len(widths)
    c0 = m.cost()
    # plain: the reference's schedule on the GPU, one chain per visit,
    # cycle_length (2^16) attempts per visit, each visit one launch
    m.set_refine(M.REFINE_PLAIN, rounds=0)
    m.run_timed(1)                            # untimed: kernels loaded
    m.put_chains(pts0)
    spin_up()
    t_plain, hist = 0.0, [(0.0, c0)]
    for sweep in range(2000):
        _, ms = m.run_timed(nch)
        t_plain += ms
        hist.append((t_plain, m.cost()))
        gain = c0 - hist[-1][1]
        # converged: the last 20 sweeps gained < 0.1% of the gain so far
        if len(hist) > 21 and hist[-21][1] - hist[-1][1] <= 1e-3 * gain:
            break
        if t_plain > 60000.0:
            break
        if sweep % log_every == 0:
            log(f"  plain sweep {sweep}: {t_plain:.1f} ms cost {hist[-1][1]}")
    c_plain = hist[-1][1]
    gain = c0 - c_plain
    out = {"c0": c0, "plain_cost": c_plain, "plain_ms": round(t_plain, 3), "plain_sweeps": len(hist) - 1}

    def race(setup, params=None):
        # one untimed step of the same leg first: its kernels loaded and its
        # hipGraph instantiated, as inside a job; then the same start again
        m.put_chains(pts0)
        setup()
        if params is not None:
            m.set_learned(params)
            m.warm_start()
        m.run_timed(1)
        m.put_chains(pts0)
        if params is not None:
            m.set_learned(params)
        spin_up()
        t, warm_ms = 0.0, 0.0
        if params is not None:  # learned warm start, timed with the descent
            warm_ms = m.warm_start()
            t += warm_ms
            m.set_learned(None)
        c = c_start = m.cost()
        reach_lit = t if c <= 1.005 * c_plain else None
        reach_gain = t if c <= c_plain + 0.005 * gain else None
        for _ in range(4000):
            if reach_gain is not None and reach_lit is not None:
                break
            _, ms = m.run_timed(1)
            t += ms
            c = m.cost()
            if reach_lit is None and c <= 1.005 * c_plain:
                reach_lit = t
            if reach_gain is None and c <= c_plain + 0.005 * gain:
                reach_gain = t
            if t > 60000.0:
                break
        return {"start_cost": c_start, "warm_start_ms": round(warm_ms, 4),
                "ms_to_within_0.5pct_of_plain_cost": reach_lit, "ms_to_99.5pct_of_plain_gain": reach_gain,
                "final_cost": c}

    batched = lambda: m.set_refine(M.REFINE_BATCHED, rounds=32, tile=-1, graph_passes=4)  # autotuned tile
    out["batched"] = race(batched)
    if learned is not None:
        out["learned_batched"] = race(batched, learned)
    # the non-learned proposal baseline: nearest-neighbour partners
    out["knn_batched"] = race(lambda: m.set_refine(M.REFINE_KNN, rounds=16, tile=3, graph_passes=0))
    # primary measure: time to 99.5% of the plain descent's gain (the 0.5%
    # cost band is reached by any start within 0.5% of the optimum)
    out["primary"] = "ms_to_99.5pct_of_plain_gain"
    for k in ("batched", "learned_batched", "knn_batched"):
        if k in out:
            b = out[k]["ms_to_99.5pct_of_plain_gain"]
            out[f"ratio_{k}_vs_plain"] = round(b / t_plain, 5) if (b is not None and t_plain > 0) else None
            b = out[k]["ms_to_within_0.5pct_of_plain_cost"]
            out[f"ratio_{k}_vs_plain_0.5pct_cost"] = round(b / t_plain, 5) if (b is not None and t_plain > 0) else None
    return out


def greedy_blobs_beside_reference(M, synth):
    """Blob detection with non-default options (colour threshold 0.1, dust
    below 4 px): the engine (device ingest, greedy merge and unification on
    the host, device pixel grouping) and the compiled reference on the same
    2 x 256^2 noisy frames; wall times and whether the blob lists (order,
    running means, membership) are identical."""
    from oracle import oracle_py as op
    import numpy as np
    frames = synth.noisy_ellipse_frames(np.random.default_rng(77), 2, 256, 256, nell=12)
    opts = {"seed": 3, "blob_threshold": 0.1, "blob_min_size": 4}
    m = M.Morph()
    m.set(**opts)
    for i, f in enumerate(frames):
        m.add_frame_rgba(i, f)
    m.set_resolution(256, 256)
    m.run(0)
    t0 = time.perf_counter()
    m.run(1)
    t_engine = time.perf_counter() - t0
    got = [[m.blob(f, b) for b in range(m.blob_count(f))] for f in range(2)]
    r = op.ref()
    rm = r.ref_morph_new()
    for k, v in dict(opts, threads=0).items():
        r.ref_morph_set_option(rm, k.encode(), float(v))
    for i, f in enumerate(frames):
        a = np.ascontiguousarray(f)
        r.ref_morph_add_rgba(rm, i, op.ptr(a), 256, 256)
    r.ref_morph_set_resolution(rm, 256, 256)
    r.ref_morph_run(rm, 0)
    t0 = time.perf_counter()
    r.ref_morph_run_until(rm, 2, 1 << 40)
    t_ref = time.perf_counter() - t0
    same = True
    for f in range(2):
        nb = r.ref_morph_blob_count(rm, f)
        same &= nb == len(got[f])
        for b in range(min(nb, len(got[f]))):
            g = C.c_size_t()
            mm = np.zeros(6)
            cnt = r.ref_morph_blob(rm, f, b, C.byref(g), op.ptr(mm), None)
            surf = np.zeros(cnt, np.uint64)
            r.ref_morph_blob(rm, f, b, C.byref(g), None, op.ptr(surf))
            same &= bool(np.array_equal(mm, got[f][b]["means"]) and
                         np.array_equal(np.sort(got[f][b]["surface"]), surf))
    r.ref_morph_free(rm)
    return {"sample": "2 x 256^2 noisy ellipses, blob_threshold 0.1, blob_min_size 4, seed 3",
            "blobs": [len(g) for g in got], "engine_s": round(t_engine, 4), "reference_s": round(t_ref, 3),
            "identical_blobs": bool(same)}


def cpu_baseline(frames, gpu_attempts, match_steps, fluid_steps):
    """The compiled reference on this host, one core (threads = 0), on a
    bounded sample of the same job: full blob detection, matching and chain
    init; a timed refinement sample scaled to the GPU job's attempt count;
    3 rendered frames scaled to 60.  With fluid on, the reference's fluid grid
    is (W+20)x(H+20) FP64 (its sparseness is fixed at 1)."""
    from oracle import oracle_py as op
    if not op.have_ref():
        return None
    import numpy as np
    r = op.ref()
    m = r.ref_morph_new()
    for k, v in {"seed": 1, "threads": 0, "cycle_length": 1 << 16, "fluid": fluid_steps}.items():
        r.ref_morph_set_option(m, k.encode(), float(v))
    for i, f in enumerate(frames):
        a = np.ascontiguousarray(f)
        r.ref_morph_add_rgba(m, i, op.ptr(a), f.shape[1], f.shape[0])
    H, W = frames[0].shape[:2]
    r.ref_morph_set_resolution(m, W, H)
    st = r.ref_morph_run(m, 0)
    t0 = time.perf_counter()
    while st == 0:
        st = r.ref_morph_run(m, 50000)
    t_detect = time.perf_counter() - t0
    t0 = time.perf_counter()
    r.ref_morph_run(m, match_steps + 1)
    r.ref_morph_next_state(m)
    r.ref_morph_run(m, 1)
    t_match_init = time.perf_counter() - t0
    n = r.ref_morph_chain_count(m)
    t0 = time.perf_counter()
    r.ref_morph_run(m, n)           # one 2^16-attempt visit per chain
    t_visits = time.perf_counter() - t0
    rate = n * (1 << 16) / t_visits
    # the reference's racing-thread mode (thread.cpp:1051-1062): every chain
    # visit runs `threads` workers of 2^16 attempts each on the same chain
    nthr = max(1, min(16, os.cpu_count() or 1))
    r.ref_morph_set_option(m, b"threads", float(nthr))
    t0 = time.perf_counter()
    r.ref_morph_run(m, n)
    rate_thr = n * nthr * (1 << 16) / (time.perf_counter() - t0)
    r.ref_morph_set_option(m, b"threads", 0.0)
    r.ref_morph_next_state(m)
    r.ref_morph_run(m, 1)
    out = np.zeros((H, W, 4), np.uint8)
    nr = 3
    t0 = time.perf_counter()
    for i in range(nr):
        r.ref_morph_render(m, r.ref_morph_get_time(m, i * 20 + 7, FRAMES_OUT), op.ptr(out))
    t_frame = (time.perf_counter() - t0) / nr
    r.ref_morph_free(m)
    total = t_detect + t_match_init + gpu_attempts / rate + FRAMES_OUT * t_frame
    total_thr = t_detect + t_match_init + gpu_attempts / rate_thr + FRAMES_OUT * t_frame
    cpu_model = None
    try:
        with open("/proc/cpuinfo") as fh:
            cpu_model = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), None)
    except OSError:
        pass
    return {"value": round(FRAMES_OUT / total, 5), "unit": "frames/s", "cores": 1, "kind": "reference",
            "host": {"cpu_model": cpu_model, "nproc": os.cpu_count()},
            "sample": (f"compiled reference on this host, threads=0: full blob detection ({t_detect:.1f} s), "
                       f"{match_steps} matching steps + chain init ({t_match_init:.1f} s), {n} chain visits of 2^16 "
                       f"attempts ({rate / 1e6:.2f} M attempts/s, scaled to the GPU job's {gpu_attempts} attempts), "
                       f"{nr} rendered frames ({t_frame:.2f} s each, scaled to {FRAMES_OUT}); job estimate "
                       f"{total:.1f} s"),
            "phases_s": {"detect": round(t_detect, 3), "match_init": round(t_match_init, 3),
                         "refine": round(gpu_attempts / rate, 3), "render": round(FRAMES_OUT * t_frame, 3)},
            "threaded": {"cores": nthr, "value": round(FRAMES_OUT / total_thr, 5),
                         "refine_attempts_per_s": round(rate_thr, 1),
                         "note": "refinement in the reference's racing-thread mode; detection, matching and "
                                 "rendering are single-threaded in the reference"}}


def counter_calibration():
    """Kernels of known HBM traffic for calibrating FETCH_SIZE / WRITE_SIZE
    and the read-request-size counters: a read-only reduction of 1 GiB,
    reads of one 4-byte word per 128 B and per 256 B line (each line read
    once), and a 256 MiB device copy.  Under a counter profiler each shows as
    its own dispatch."""
    import torch
    torch.cuda.set_device(0)
    gib = 1 << 30
    x = torch.ones(gib // 4, dtype=torch.float32, device="cuda")  # float: the sums read it as it is
    torch.cuda.synchronize()
    for what, fn in (("read-only sum of 1 GiB", lambda: x.sum()),
                     ("one word per 128 B line, 1 GiB span", lambda: x.view(-1, 32)[:, 0].sum()),
                     ("one word per 256 B line, 1 GiB span", lambda: x.view(-1, 64)[:, 0].sum())):
        log(what)
        fn()
        torch.cuda.synchronize()
    a = torch.ones(256 << 20, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    log("device copy of 256 MiB")
    b.copy_(a)
    torch.cuda.synchronize()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="blobs1024")
    ap.add_argument("--match-steps", type=int, default=20000)
    ap.add_argument("--graphs", type=int, default=8, help="refinement hipGraph replays per job (16 passes each)")
    ap.add_argument("--fluid-steps", type=int, default=10, help="fluid model steps per output frame (0 = atom path)")
    ap.add_argument("--fluid-sparseness", type=float, default=0.25,
                    help="fluid grid scale: 0.25 gives a 256x256 grid (+ 20-node border) at 1024^2")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-time-to-cost", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="job without the learned warm start")
    ap.add_argument("--counter-calibration", action="store_true",
                    help="only run reads and a copy of known byte counts (for calibrating HBM counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); this build runs the morph engine on a "
              f"single MI355X. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.counter_calibration:
        counter_calibration()
        return

    import numpy as np
    import torch
    from atomorph_amd import learned as L
    from atomorph_amd import morph as M
    from atomorph_amd import synth

    torch.cuda.set_device(0)
    params = None if args.no_learned else L.load_default()
    if params is None and not args.no_learned:
        log("no trained warm-start parameters (atomorph_amd/data/warmstart_mlp.npz): learned path off")
    frames = synth.workload(args.workload)
    H, W = frames[0].shape[:2]
    m = build_morph(M, frames)
    out = torch.empty(FRAMES_OUT * H * W * 4, dtype=torch.uint8, device="cuda")
    times = [m.get_time(i, FRAMES_OUT) for i in range(FRAMES_OUT)]

    # time-to-cost (once, before the throughput steps); the batched kernel
    # variant is autotuned for 32 rounds per pass on these chains first
    # (untimed, like a kernel-selection cache; the engine reuses it for every
    # job on chains of the same shapes)
    ttc = None
    tuned = None
    if not args.no_time_to_cost:
        m.run(0)
        to_morphing(m, M, args.match_steps)
        tile, _, rep = m.autotune(3.0, rounds=32)
        tuned = {"tile": tile, "candidates": rep}
        log(f"autotuned tile {tile}: {rep}")
        log("time-to-cost: plain descent to convergence, then batched and learned + batched")
        ttc = time_to_cost(m, M, learned=params)
        log(f"  {ttc}")

    m.set(cycle_length=1 << 16, fluid=args.fluid_steps, fluid_sparseness=args.fluid_sparseness)
    m.set_learned(params)  # the warm start runs right after the chains are built
    m.set_refine(M.REFINE_BATCHED, rounds=32, tile=-1, graph_passes=16)
    phase = {"detect_ms": [], "refine_gpu_ms": [], "render_gpu_ms": [], "job_ms": []}
    attempts_job = accepted_job = 0

    def job(record):
        nonlocal attempts_job, accepted_job
        m.restart()
        m.run(0)                                   # key frames into HBM (untimed)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, ms_det = m.run_timed(1)
        m.run(args.match_steps + 1)
        m.next_state()
        if m.run(1) != M.STATE_ATOM_MORPHING:
            raise RuntimeError(m.error())
        a0, acc0 = m.counters()
        _, ms_ref = m.run_timed(args.graphs)
        a1, acc1 = m.counters()
        ms_ren = m.render_device(times, out.data_ptr())
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        if record:
            phase["detect_ms"].append(ms_det)
            phase["refine_gpu_ms"].append(ms_ref)
            phase["render_gpu_ms"].append(ms_ren)
            phase["job_ms"].append(dt * 1e3)
            attempts_job, accepted_job = a1 - a0, acc1 - acc0
        return dt

    for _ in range(args.warmup):
        job(False)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    total = 0.0
    for _ in range(args.steps):
        total += job(True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ms_per_step = 1e3 * total / args.steps
    value = FRAMES_OUT * args.steps / total

    pts, widths = m.chains()
    h = pts.shape[0]
    passes = 16 * args.graphs

    # the same job with the 60 frames read back to host memory (pinned), as
    # a caller of morph::get_pixels receives them
    host = torch.empty(FRAMES_OUT * H * W * 4, dtype=torch.uint8, pin_memory=True)
    readback = []
    for _ in range(2):
        dt = job(False)
        t0 = time.perf_counter()
        host.copy_(out)
        torch.cuda.synchronize()
        readback.append(dt * 1e3 + (time.perf_counter() - t0) * 1e3)
    job_readback_ms = float(np.mean(readback))

    # like for like with the reference's fluid: its grid is (W+20) x (H+20)
    # (sparseness 1); the same job on that grid
    like = None
    if args.fluid_steps and args.fluid_sparseness != 1.0:
        m.set(fluid_sparseness=1.0)
        job(False)
        lk = [job(False) for _ in range(2)]
        m.set(fluid_sparseness=args.fluid_sparseness)
        job(False)
        like_ms = 1e3 * float(np.mean(lk))
        like = {"fluid_sparseness": 1.0, "grid": f"{W + 20}x{H + 20}", "ms_per_job": round(like_ms, 3),
                "value": round(FRAMES_OUT / (like_ms / 1e3), 3), "unit": "frames/s"}

    # one more job with per-kernel device timing (HIP events on the engine's
    # stream around each launch); not part of the timed steps
    from atomorph_amd import _capi
    lib = _capi.load()
    lib.amh_profile_reset()
    lib.amh_profile_enable(1)
    job(False)
    torch.cuda.synchronize()
    kernels = {}
    for name in PROFILED:
        n, ms = C.c_uint64(0), C.c_double(0.0)
        lib.amh_profile_read(name.encode(), C.byref(n), C.byref(ms))
        if n.value:
            kernels[name] = {"launches": int(n.value), "ms": round(ms.value, 4),
                             "avg_us": round(1e3 * ms.value / n.value, 3)}
    lib.amh_profile_enable(0)
    lib.amh_profile_reset()
    log("kernels (profiled job): " + json.dumps(kernels))

    # Algorithmic HBM bytes per launch of the hot kernels (what each must move
    # at least once), divided by their measured average launch time.
    #   k_refine_pass  per attempt the two slots' points in rows y-1, y, y+1
    #                  (8 B each; rows y-1 = y+1 when h = 2), 16 B written per
    #                  accepted swap; the passes run in hipGraph replays of 16.
    #   fluid kernels  per active particle A the FP64/int state they read and
    #                  write, per grid node N the FP32 fields:
    #                  k_p2g    A (7x8 + 4 + 4) read, N 8x4 written
    #                  k_forces A (2x8 + 4) read, N 3x4 read + 2x4 written
    #                  k_g2p    A (6x8 + 4 + 4) read + 2x8 written, N 3x4 + 2x4
    #                  k_advect A (15x8 + 2x4) read + 8x8 written, N 8x4 read
    per_attempt = 8 * 2 * (2 + (1 if h > 2 else 0))
    refine_bytes = attempts_job * per_attempt + accepted_job * 16
    rooflines = {}
    g = kernels.get("refine_graph")
    if g:
        launch_s = g["ms"] / 1e3 / (16 * g["launches"])
    else:
        launch_s = float(np.mean(phase["refine_gpu_ms"])) / 1e3 / passes
    rooflines["k_refine_pass"] = (refine_bytes / passes, launch_s, passes)
    if args.fluid_steps:
        P = m.fluid_particles()
        A = int((P[:, 16] > 0).sum()) if P.size else 0
        N = (int(W * args.fluid_sparseness) + 20) * (int(H * args.fluid_sparseness) + 20)
        per = {"k_p2g": A * 64 + N * 32, "k_forces": A * 20 + N * 20, "k_g2p": A * 72 + N * 20,
               "k_advect": A * 192 + N * 32}
        for k, b in per.items():
            if k in kernels:
                rooflines[k] = (b, kernels[k]["avg_us"] / 1e6, kernels[k]["launches"])
    # the dominant kernel: the largest device time in the profiled job
    def total_ms(k):
        if k == "k_refine_pass":
            return g["ms"] if g else float(np.mean(phase["refine_gpu_ms"]))
        return kernels[k]["ms"]
    dom = max(rooflines, key=total_ms)
    dom_bytes, dom_s, _ = rooflines[dom]
    achieved = dom_bytes / dom_s / 1e9
    roof_all = {k: {"bytes_per_launch": int(b), "launch_us": round(t * 1e6, 3), "GBs": round(b / t / 1e9, 1),
                    "total_ms": round(total_ms(k), 3)} for k, (b, t, _) in rooflines.items()}
    # measured HBM bytes per launch of the dominant kernel: rocprofv3
    # counter passes on this workload (reads from the L2's read requests by
    # size, writes from WRITE_SIZE), kept as data beside the profile
    # summaries (null when absent for this workload)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"{args.workload}_hbm_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tdata = json.load(f)
        if tdata.get("workload") == args.workload:
            # kernel names as the profiler prints them (templates with their
            # arguments); the launch of this build's variant of `dom`
            for name, b in tdata.get("bytes_per_launch", {}).items():
                if name == dom or name.split("<")[0] == dom:
                    traffic = int(b)
                    break
    # rendering, for comparison: per atom and frame h points + 2 colours read,
    # per output pixel 4 B written
    render_bytes = FRAMES_OUT * (int(widths.sum()) * (8 * h + 8) + H * W * 4)
    render_gbs = render_bytes / (float(np.mean(phase["render_gpu_ms"])) / 1e3) / 1e9

    # matrix-core use of the warm start's MLP inference (one forward of all
    # atoms per key segment)
    mlp_stats = None
    f = kernels.get("mlp_forward")
    if f and f["ms"] > 0:
        flop = MLP_FLOP_PER_ROW * int(widths.sum()) * f["launches"]
        tf = flop / (f["ms"] / 1e3) / 1e12
        mlp_stats = {"kernel": "k_mlp_fused (v_mfma_f32_32x32x16_bf16, 5 layers chained in registers)",
                     "bound": "mfma",
                     "flop_per_launch": int(flop / f["launches"]), "avg_us": f["avg_us"],
                     "achieved": round(tf, 2), "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(tf / MFMA_BF16_PEAK_TFLOPS, 5)}

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline: compiled reference, bounded sample")
        cpu = cpu_baseline(frames, attempts_job, args.match_steps, args.fluid_steps)
        if cpu is not None:
            log("greedy blob options beside the reference")
            cpu["greedy_blobs"] = greedy_blobs_beside_reference(M, synth)
            cpu["gpu_over_cpu"] = round(value / cpu["value"], 1)
            if like is not None:
                cpu["gpu_over_cpu_same_fluid_grid"] = round(like["value"] / cpu["value"], 1)

    res = {
        "metric": "time to equal OT cost, learned+batched vs plain GPU, 1024²; frames/s",
        "value": round(value, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: 4 key frames 1024x1024 RGBA8, 150 fBm-thresholded blobs each, "
                               f"atom trajectories (SPLINE, PERLIN fade), "
                               + (f"fluid on ({args.fluid_steps} steps/frame, "
                                  f"{int(W * args.fluid_sparseness) + 20}x{int(H * args.fluid_sparseness) + 20} "
                                  f"FP32 grid)" if args.fluid_steps else "fluid off")
                               + f", {FRAMES_OUT} frames rendered",
                   "atoms": int(widths.sum()), "chains": int(len(widths)), "key_frames": h,
                   "blob_pixels": int(sum(int(np.count_nonzero(f[..., 3])) for f in frames)),
                   "match_steps": args.match_steps, "refine_passes": passes, "refine_attempts": attempts_job},
        "phases_ms": {k: round(float(np.mean(v)), 3) for k, v in phase.items()},
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "bytes_per_launch": int(dom_bytes), "launch_us": round(dom_s * 1e6, 3)},
        "rooflines": roof_all,
        "render_gbs": round(render_gbs, 1),
        "kernels": kernels,
        "mlp": mlp_stats,
        "job_with_readback_ms": round(job_readback_ms, 3),
        "like_for_like_fluid_grid": like,
        "time_to_cost": ttc,
        "autotune": tuned,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
