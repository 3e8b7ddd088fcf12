"""Device-side helpers for the GPU parity tests (test infrastructure)."""
import ctypes as C

import numpy as np

import galfast_amd as G
from galfast_amd import _abi as A


def torch():
    import torch as T
    return T


def stream_ptr():
    T = torch()
    return C.c_void_p(T.cuda.current_stream().cuda_stream)


class DevRng:
    """Device copy of a host stream table."""

    def __init__(self, host_state):
        T = torch()
        self.ns = host_state.size // 3
        self.t = T.from_numpy(host_state.astype(np.uint32).view(np.int32).copy()).cuda()
        self.rng = A.gf_rng(self.t.data_ptr(), self.ns, 0)

    def state(self):
        torch().cuda.synchronize()
        return self.t.cpu().numpy().view(np.uint32)


def dev_array(a):
    T = torch()
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        return T.from_numpy(a.view(np.int32).copy()).cuda()
    return T.from_numpy(a.copy()).cuda()


def to_np(t, dtype=None):
    torch().cuda.synchronize()
    a = t.cpu().numpy()
    if dtype is not None:
        a = a.view(dtype)
    return a


def sync():
    torch().cuda.synchronize()
