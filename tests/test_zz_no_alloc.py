"""hcc_mlp_gradients_noalloc: the indicator network's gradient on device-resident data with its workgroup partials
in the library's static scratch (no allocation in the call), against hcc_mlp_gradients on the same batch, bit for
bit. Two calls in a row, the second with a smaller batch, so that the scratch holds the first call's partials when
the second starts."""
import ctypes as C

import numpy as np
import pytest

from hcc_hip import MLP
from hcc_hip import _lib as L
from oracle import mlp as NN

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return _hip


class DevBuf:
    """Caller-owned device memory (the caller allocates; the entry point under test does not)."""

    def __init__(self, nbytes):
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), max(int(nbytes), 1)) == 0
        self.n = int(nbytes)

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.n
        assert hip().hipMemcpy(self.p, a.ctypes.data, a.nbytes, 1) == 0      # hipMemcpyHostToDevice

    def download(self, dtype, count):
        out = np.empty(count, dtype)
        assert hip().hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0  # hipMemcpyDeviceToHost
        return out

    def free(self):
        if self.p:
            hip().hipFree(self.p)
            self.p = None


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, NN.F)).astype(np.float32)
    y = (X[:, 0] - 0.5 * X[:, 5] + 0.3 * X[:, 9] ** 2 > 0.4).astype(np.uint8)
    return L.bf16_rows(X), y


def test_noalloc_gradients_equal_the_allocating_entry_point(ctx):
    m = MLP(ctx, NN.init_params(11))
    print("case: hcc_mlp_gradients_noalloc, batches 40000 then 1000 rows", flush=True)
    sizes = [40000, 1000]          # 1250 tiles (all 512 partial rows in use), then 32 tiles
    nmax = max(sizes)
    bX, by, bg = DevBuf(nmax * NN.F * 2), DevBuf(nmax), DevBuf((L.MLP_PARAMS + 1) * 4)
    try:
        for k, n in enumerate(sizes):
            Xb, y = _data(n, 20 + k)
            bX.upload(Xb)
            by.upload(y)
            rc = L.lib().hcc_mlp_gradients_noalloc(m.h, bX.p, by.p, n, C.c_float(3.0), bg.p)
            L.check(rc, ctx.h, "hcc_mlp_gradients_noalloc")
            L.check(L.lib().hcc_ctx_synchronize(ctx.h), ctx.h, "sync")
            out = bg.download(np.float32, L.MLP_PARAMS + 1)
            loss, g = m.gradients(Xb, y, 3.0)
            assert np.array_equal(out[:L.MLP_PARAMS].view(np.uint32), g.view(np.uint32)), f"call {k}: gradients differ"
            assert np.float32(loss) == out[L.MLP_PARAMS], f"call {k}: loss {out[L.MLP_PARAMS]} vs {loss}"
            assert np.all(np.isfinite(g)) and np.any(g != 0)
    finally:
        for b in (bX, by, bg):
            b.free()
        m.close()


def test_noalloc_rejects_bad_arguments(ctx):
    m = MLP(ctx, NN.init_params(1))
    try:
        assert L.lib().hcc_mlp_gradients_noalloc(m.h, None, None, 10, C.c_float(1.0), None) == -1
        assert L.lib().hcc_mlp_gradients_noalloc(m.h, None, None, 0, C.c_float(1.0), None) == -1
    finally:
        m.close()
