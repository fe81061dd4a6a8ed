"""rhllc_mlp_grad_noalloc: the allocation-free variant of rhllc_mlp_grad.

It takes device arrays and keeps its packed weights and per-wave partials in
the library's per-GPU scratch array; its gradient and loss must be the same
bits as rhllc_mlp_grad's (which stages through hipMalloc'd buffers and
reduces the partials on the host in the same order).  Two calls in a row, the
second with a smaller batch, so the scratch keeps the first call's larger
contents.  Device memory for the arrays comes from the HIP runtime the
library links against (libamdhip64), through ctypes.
"""
import ctypes as C
import os

import numpy as np
import pytest

from radrhd_solver_amd import _lib, capi

pytestmark = pytest.mark.gpu

SEED = 20261202
MAX_N = 65536  # RHLLC_MLP_NOALLOC_MAX_N
ERR_ARG = 1  # RHLLC_ERR_ARG


class _Hip:
    def __init__(self):
        path = "/opt/rocm/lib/libamdhip64.so"
        if not os.path.exists(path):
            path = "libamdhip64.so"
        self.L = C.CDLL(path)
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.bufs = []

    def to_dev(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 4)) == 0
        self.bufs.append(p)
        if a.nbytes:
            assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def to_host(self, p, shape, dtype):
        out = np.zeros(shape, dtype)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _batch(rng, n):
    X = rng.normal(size=(n, 3)).astype(np.float32)
    y = rng.normal(size=n).astype(np.float32)
    ok = (rng.uniform(size=n) > 0.1).astype(np.float32)
    return X, y, ok


def test_grad_noalloc_matches_grad_on_two_calls():
    lib = _lib.load_hip_lib()
    model, _ = capi.mlp_train(steps=50, batch=4096, lr=3e-3, seed=SEED)
    rng = np.random.default_rng(SEED + 5)
    hip = _Hip()
    try:
        dmodel = hip.to_dev(model)
        dgrad = hip.to_dev(np.zeros(capi.MLP_PARAMS, np.float32))
        dloss = hip.to_dev(np.zeros(1, np.float32))
        for n in (16384, 2048):
            print("no-alloc grad, n =", n, flush=True)
            X, y, ok = _batch(rng, n)
            g_ref, l_ref = capi.mlp_grad(model, X, y, ok)
            dX, dy, dok = hip.to_dev(X), hip.to_dev(y), hip.to_dev(ok)
            rc = lib.rhllc_mlp_grad_noalloc(0, dmodel, n, dX, dy, dok, dgrad, dloss, None)
            assert rc == 0, lib.rhllc_last_error().decode()
            assert hip.L.hipStreamSynchronize(None) == 0
            g = hip.to_host(dgrad, capi.MLP_PARAMS, np.float32)
            loss = hip.to_host(dloss, 1, np.float32)[0]
            assert np.array_equal(g, g_ref), np.abs(g - g_ref).max()
            assert loss == np.float32(l_ref)
        # beyond the scratch: a clear error, nothing launched
        rc = lib.rhllc_mlp_grad_noalloc(0, dmodel, MAX_N + 256, dX, dy, dok, dgrad, dloss, None)
        assert rc == ERR_ARG and "exceeds the no-alloc scratch" in lib.rhllc_last_error().decode()
    finally:
        hip.free()
