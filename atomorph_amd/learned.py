"""Learned warm start for the transport refinement: training data and
training of the proposal MLP.

The data are converged plain-descent assignments on synthetic 256 x 256
frame pairs (noise-thresholded blobs with random colours): for every atom of
a chain row, the engine's GPU feature kernel gives the 8 MLP inputs and the
normalised position the same slot holds in the next key frame
(Morph.learn_samples).  Training runs entirely on the library's own kernels
through the C ABI -- forward and backward products on the matrix cores,
MSE, Adam -- with torch only holding device buffers.

    python -m atomorph_amd.learned --pairs 2000 --out atomorph_amd/data/warmstart_mlp.npz
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

from . import _capi
from . import morph as M
from . import synth

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_PARAMS = os.path.join(HERE, "data", "warmstart_mlp.npz")
DIMS = [8, 128, 128, 128, 128, 2]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pair_frames(i, size=256):
    """Training pair i: two independent noise-blob frames, 8-16 blobs each,
    18-40% coverage (seeded by i)."""
    r = synth.SplitMix64(0xC0FFEE + 7919 * i)
    nb = r.randint(8, 16)
    cov = r.uniform(0.18, 0.40)
    return [synth.noise_blob_frame(size, nb, cov, seed=1_000_000 + 2 * i + f)[0] for f in range(2)]


def build(frames, seed=1, match_steps=3000):
    """A Morph paused in ATOM_MORPHING on freshly built chains."""
    m = M.Morph()
    m.set(seed=seed, cycle_length=1 << 16)
    for k, f in enumerate(frames):
        m.add_frame_rgba(k, f)
    h, w = frames[0].shape[:2]
    m.set_resolution(w, h)
    m.run(0)
    m.run(1)
    m.run(match_steps + 1)
    m.next_state()
    if m.run(1) != M.STATE_ATOM_MORPHING:
        raise RuntimeError(f"no chains: {m.error()}")
    return m


def converge_plain(m, rounds=4, window=20, tol=1e-3, max_sweeps=600):
    """Plain descent (one chain at a time) until the last `window` sweeps
    gained less than tol of the gain so far.  Returns (start, end) cost."""
    _, widths = m.chains()
    c0 = m.cost()
    m.set_refine(M.REFINE_PLAIN, rounds=rounds)
    hist = [c0]
    for _ in range(max_sweeps):
        m.run(len(widths))
        hist.append(m.cost())
        gain = c0 - hist[-1]
        if len(hist) > window + 1 and hist[-window - 1] - hist[-1] <= tol * max(gain, 1):
            break
    return c0, hist[-1]


def pair_samples(i, keep=4096, rng=None):
    m = build(pair_frames(i), seed=1 + i)
    converge_plain(m)
    x, t = m.learn_samples()
    if rng is not None and x.shape[0] > keep:
        sel = rng.choice(x.shape[0], keep, replace=False)
        x, t = x[sel], t[sel]
    return x, t


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    parts = []
    for din, dout in zip(DIMS[:-1], DIMS[1:]):
        parts.append(rng.normal(0, np.sqrt(1.0 / din), din * dout))
        parts.append(np.zeros(dout))
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == _capi.MLP_PARAMS
    return p


class Trainer:
    """The MLP handle plus device copies of a dataset."""

    def __init__(self, x, t, params, batch=16384):
        import torch
        self.torch = torch
        self.lib = _capi.load()
        self.stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        self.h = C.c_void_p()
        _capi.check(self.lib.amh_mlp_create(C.byref(self.h), batch, self.stream), "mlp create")
        self.batch = batch
        self.set_params(params)
        self.x = torch.from_numpy(np.ascontiguousarray(x, np.float32)).cuda()
        self.t = torch.from_numpy(np.ascontiguousarray(t, np.float32)).cuda()
        self.loss = torch.zeros(1, dtype=torch.float32, device="cuda")

    def close(self):
        if self.h:
            self.lib.amh_mlp_destroy(self.h)
            self.h = None

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _capi.check(self.lib.amh_mlp_set_params(self.h, p.ctypes.data_as(C.c_void_p), self.stream), "set params")

    def params(self):
        p = np.zeros(_capi.MLP_PARAMS, np.float32)
        _capi.check(self.lib.amh_mlp_get_params(self.h, p.ctypes.data_as(C.c_void_p), self.stream), "get params")
        return p

    def step(self, idx, lr):
        xb = self.x.index_select(0, idx).contiguous()
        tb = self.t.index_select(0, idx).contiguous()
        self.loss.zero_()
        _capi.check(self.lib.amh_mlp_gradients(self.h, C.c_void_p(xb.data_ptr()), C.c_void_p(tb.data_ptr()),
                                               int(idx.numel()), None, C.c_void_p(self.loss.data_ptr()),
                                               self.stream), "gradients")
        _capi.check(self.lib.amh_mlp_adam_step(self.h, C.c_float(lr), self.stream), "adam")
        return self.loss

    def predict(self, x):
        T = self.torch
        xd = x if isinstance(x, T.Tensor) else T.from_numpy(np.ascontiguousarray(x, np.float32)).cuda()
        out = T.empty((xd.shape[0], 2), dtype=T.float32, device="cuda")
        _capi.check(self.lib.amh_mlp_forward(self.h, C.c_void_p(xd.data_ptr()), int(xd.shape[0]),
                                             C.c_void_p(out.data_ptr()), self.stream), "forward")
        return out


def mse(pred, t):
    return float(((pred - t) ** 2).sum(dim=1).mean().item())


def train(x, t, xv, tv, steps=3000, batch=16384, lr=1e-3, seed=0, log_every=250):
    import torch
    tr = Trainer(x, t, init_params(seed), batch=batch)
    try:
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)
        n = tr.x.shape[0]
        hist = []
        t0 = time.time()
        for s in range(steps):
            lr_s = lr * 0.5 * (1.0 + np.cos(np.pi * s / steps))  # cosine decay
            idx = torch.randint(0, n, (batch,), device="cuda", generator=g)
            loss = tr.step(idx, max(lr_s, lr * 1e-3))
            if s % log_every == 0 or s == steps - 1:
                hist.append((s, float(loss.item())))
                log(f"  step {s}: loss {hist[-1][1]:.5f} ({time.time() - t0:.1f} s)")
        torch.cuda.synchronize()
        secs = time.time() - t0
        p = tr.params()
        xvd = torch.from_numpy(np.ascontiguousarray(xv)).cuda()
        tvd = torch.from_numpy(np.ascontiguousarray(tv)).cuda()
        val = mse(tr.predict(xvd), tvd)
        ident = mse(xvd[:, 6:8], tvd)  # the untrained proposal: keep the relative position
    finally:
        tr.close()
    return p, {"steps": steps, "batch": batch, "lr": lr, "train_seconds": round(secs, 2),
               "loss_history": hist, "val_mse": val, "val_mse_identity": ident}


def load_default():
    """The committed trained parameters, or None when absent."""
    if not os.path.exists(DEFAULT_PARAMS):
        return None
    with np.load(DEFAULT_PARAMS, allow_pickle=False) as z:
        return np.ascontiguousarray(z["params"], np.float32)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=2000)
    ap.add_argument("--keep", type=int, default=4096, help="samples kept per pair")
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=DEFAULT_PARAMS)
    args = ap.parse_args()
    import torch
    torch.cuda.set_device(0)
    rng = np.random.default_rng(0)
    xs, ts = [], []
    t0 = time.time()
    for i in range(args.pairs):
        x, t = pair_samples(i, args.keep, rng)
        xs.append(x)
        ts.append(t)
        if i % 100 == 0:
            log(f"pair {i}: {x.shape[0]} samples ({time.time() - t0:.1f} s)")
    gen_s = time.time() - t0
    nval = max(1, args.pairs // 20)
    x, t = np.concatenate(xs[nval:]), np.concatenate(ts[nval:])
    xv, tv = np.concatenate(xs[:nval]), np.concatenate(ts[:nval])
    log(f"{x.shape[0]} training samples, {xv.shape[0]} validation ({gen_s:.1f} s)")
    p, meta = train(x, t, xv, tv, steps=args.steps, batch=args.batch, lr=args.lr)
    meta.update({"pairs": args.pairs, "validation_pairs": nval, "samples": int(x.shape[0]),
                 "data_seconds": round(gen_s, 1), "frames": "256x256 noise blobs, 8-16 per frame, 18-40% cover"})
    log(json.dumps(meta))
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.savez(args.out, params=p, meta=np.array(json.dumps(meta)))
    print(json.dumps(meta), flush=True)


if __name__ == "__main__":
    main()
