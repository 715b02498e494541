"""Device-memory helpers for the GPU tests: torch tensors as HBM buffers,
their raw pointers passed through the C ABI on torch's current stream."""
import ctypes as C
import sys

import numpy as np

from atomorph_amd import _capi as A


def lib():
    return A.load()


def torch():
    import torch as T
    return T


def to_dev(arr):
    """numpy -> device tensor (unsigned types viewed as signed of the same width)."""
    a = np.ascontiguousarray(arr)
    view = {np.dtype(np.uint64): np.int64, np.dtype(np.uint32): np.int32, np.dtype(np.uint8): np.uint8}
    if a.dtype in view:
        a = a.view(view[a.dtype])
    return torch().from_numpy(a.copy()).cuda()


def to_host(t, dtype):
    a = t.cpu().numpy()
    return a.view(dtype)


def dptr(t):
    return C.c_void_p(t.data_ptr())


def stream():
    return C.c_void_p(torch().cuda.current_stream().cuda_stream)


def sync():
    torch().cuda.synchronize()


def check(rc, what):
    A.check(rc, what)


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def zeros(n, dtype):
    T = torch()
    return T.zeros(n, dtype=dtype, device="cuda")


def key_major(pts_atom_major, h):
    """Chains concatenated atom-major (wtot*h) -> device layout (h, wtot)."""
    return np.ascontiguousarray(np.asarray(pts_atom_major, np.uint64).reshape(-1, h).T).ravel()


def atom_major(pts_key_major, h):
    return np.ascontiguousarray(np.asarray(pts_key_major, np.uint64).reshape(h, -1).T).ravel()
