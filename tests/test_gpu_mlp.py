"""GPU parity of the learned proposal's MLP (matrix-core kernels) against the
CPU oracle, which rounds the operands of every product to bfloat16 the same
way; tolerances cover FP32-vs-FP64 accumulation and the rare activation
that rounds to the other bfloat16 neighbour."""
import ctypes as C

import numpy as np
import pytest

from atomorph_amd import _capi as A
from oracle import oracle_py as op
from tests import gpu_util as G
from tests.test_oracle import _mlp_params

pytestmark = pytest.mark.gpu


def bf16(a):
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def _mlp(batch, params=None):
    lib = G.lib()
    h = C.c_void_p()
    G.check(lib.amh_mlp_create(C.byref(h), batch, G.stream()), "create")
    if params is not None:
        G.check(lib.amh_mlp_set_params(h, params.ctypes.data_as(C.c_void_p), G.stream()), "set")
    return h


@pytest.mark.parametrize("M,N,K", [(64, 64, 32), (70, 50, 40), (256, 128, 128), (128, 2, 128), (33, 128, 8),
                                   (128, 128, 3000)])
def test_gemm_matches_bf16_product(M, N, K):
    """The matrix-core tile kernel (operand lane layout, accumulator layout,
    edges) against numpy on bfloat16-rounded operands."""
    rng = np.random.default_rng(M * 7 + N + K)
    a = rng.normal(0, 1, (M, K)).astype(np.float32)
    b = rng.normal(0, 1, (K, N)).astype(np.float32)
    want = bf16(a).astype(np.float64) @ bf16(b).astype(np.float64)
    da, db, dc = G.to_dev(a), G.to_dev(b), G.zeros(M * N, G.torch().float32)
    G.log(f"gemm {M}x{N}x{K}")
    G.check(G.lib().amh_mlp_gemm_test(G.dptr(da), G.dptr(db), G.dptr(dc), M, N, K, G.stream()), "gemm")
    got = G.to_host(dc, np.float32).reshape(M, N).astype(np.float64)
    err = np.abs(got - want).max() / max(1.0, np.abs(want).max())
    assert err < 1e-5, err


def test_forward_matches_oracle():
    rng = np.random.default_rng(11)
    p = _mlp_params(rng)
    n = 1000
    x = rng.normal(0, 1, (n, 8)).astype(np.float32)
    want = np.zeros((n, 2))
    op.oracle().or_mlp_forward(op.ptr(p), op.ptr(x), n, 1, op.ptr(want))
    h = _mlp(256, p)  # several batches
    try:
        dx, dout = G.to_dev(x), G.zeros(n * 2, G.torch().float32)
        G.log("mlp forward 1000 rows")
        G.check(G.lib().amh_mlp_forward(h, G.dptr(dx), n, G.dptr(dout), G.stream()), "forward")
        got = G.to_host(dout, np.float32).reshape(n, 2)
    finally:
        G.lib().amh_mlp_destroy(h)
    scale = np.abs(want).max()
    assert np.abs(got - want).max() <= 2e-3 * scale, np.abs(got - want).max() / scale
    assert np.corrcoef(got.ravel(), want.ravel())[0, 1] > 0.99999


def test_gradients_match_oracle():
    """Backward kernels against or_mlp_grad (itself pinned by finite
    differences): per layer, relative L2 error of weight and bias
    gradients."""
    rng = np.random.default_rng(12)
    p = _mlp_params(rng)
    n = 700
    x = rng.normal(0, 1, (n, 8)).astype(np.float32)
    t = rng.normal(0, 1, (n, 2)).astype(np.float32)
    want = np.zeros(op.MLP_PARAMS)
    loss_want = op.oracle().or_mlp_grad(op.ptr(p), op.ptr(x), op.ptr(t), n, 1, op.ptr(want))
    h = _mlp(1024, p)
    try:
        T = G.torch()
        dx, dt = G.to_dev(x), G.to_dev(t)
        dg, dl = G.zeros(op.MLP_PARAMS, T.float32), G.zeros(1, T.float32)
        G.log("mlp gradients 700 rows")
        G.check(G.lib().amh_mlp_gradients(h, G.dptr(dx), G.dptr(dt), n, G.dptr(dg), G.dptr(dl), G.stream()), "grad")
        got = G.to_host(dg, np.float32).astype(np.float64)
        loss = float(G.to_host(dl, np.float32)[0])
    finally:
        G.lib().amh_mlp_destroy(h)
    assert abs(loss - loss_want) <= 1e-3 * loss_want
    o_ = 0
    for din, dout in zip([8, 128, 128, 128, 128], [128, 128, 128, 128, 2]):
        for part, size in (("W", din * dout), ("b", dout)):
            g, w = got[o_:o_ + size], want[o_:o_ + size]
            rel = np.linalg.norm(g - w) / max(np.linalg.norm(w), 1e-12)
            assert rel < 1e-2, (din, dout, part, rel)
            o_ += size


def test_adam_training_reduces_loss():
    """A few hundred Adam steps on a smooth target (the identity on two
    inputs plus a bend) cut the loss by well over an order of magnitude."""
    rng = np.random.default_rng(13)
    p = _mlp_params(rng, scale=1.0)
    n = 4096
    x = rng.uniform(-1, 1, (n, 8)).astype(np.float32)
    t = np.stack([x[:, 6], x[:, 7] + 0.5 * x[:, 6] ** 2], axis=1).astype(np.float32)
    h = _mlp(n, p)
    T = G.torch()
    try:
        dx, dt, dl = G.to_dev(x), G.to_dev(t), G.zeros(1, T.float32)
        losses = []
        G.log("mlp adam 300 steps")
        for step in range(300):
            dl.zero_()
            G.check(G.lib().amh_mlp_gradients(h, G.dptr(dx), G.dptr(dt), n, None, G.dptr(dl), G.stream()), "grad")
            G.check(G.lib().amh_mlp_adam_step(h, C.c_float(2e-3), G.stream()), "adam")
            if step % 50 == 0 or step == 299:
                losses.append(float(dl.item()))
        q = np.zeros(op.MLP_PARAMS, np.float32)
        G.check(G.lib().amh_mlp_get_params(h, q.ctypes.data_as(C.c_void_p), G.stream()), "get")
    finally:
        G.lib().amh_mlp_destroy(h)
    assert losses[-1] < 0.05 * losses[0], losses
    assert np.isfinite(q).all() and not np.array_equal(q, p)
