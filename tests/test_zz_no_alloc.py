"""amca_mlp_forward_device_noalloc, the entry point that allocates nothing,
against amca_mlp_forward_device (the same network on device-resident rows),
both called through the C ABI on device buffers: equal bit for bit on two
calls in a row (the second smaller), no change in free device memory across
the calls, and a clear refusal beyond the row maximum."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_ROWS = 32768  # AMCA_MLP_NOALLOC_MAX_ROWS


class _Hip:
    """Device buffers through the HIP runtime the library itself links."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        self.bufs = []

    def ok(self, st, what):
        assert st == 0, f"{what}: hip status {st}"

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self.ok(self.L.hipMalloc(C.byref(p), a.nbytes), "hipMalloc")
        self.bufs.append(p)
        self.ok(self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        self.ok(self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def free_bytes(self):
        self.ok(self.L.hipDeviceSynchronize(), "hipDeviceSynchronize")
        f, t = C.c_size_t(), C.c_size_t()
        self.ok(self.L.hipMemGetInfo(C.byref(f), C.byref(t)), "hipMemGetInfo")
        return f.value

    def close(self):
        for p in self.bufs:
            self.L.hipFree(p)


def test_noalloc_forward_equals_forward_device(amca):
    from amca3d_amd import surrogate as S
    L = S.lib()
    net = S.MLP(seed=11)
    hip = _Hip()
    rng = np.random.default_rng(5)
    p = net.params()
    p += rng.normal(0.0, 0.05, p.size).astype(np.float32)  # non-zero biases
    net.set_params(p)

    for n in (MAX_ROWS, 1237):  # the second call smaller
        X = rng.uniform(-2.0, 2.0, (n, 5)).astype(np.float32)
        nan = np.full(n, np.nan, np.float32)
        dX, dref, dgot = hip.upload(X), hip.upload(nan), hip.upload(nan)
        assert L.amca_mlp_forward_device(net.h, n, dX, dref) == 0, L.amca_mlp_last_error()
        free_before = hip.free_bytes()
        st = L.amca_mlp_forward_device_noalloc(net.h, n, dX, dgot)
        assert st == 0, L.amca_mlp_last_error()
        assert hip.free_bytes() == free_before
        ref, got = hip.download(dref, nan), hip.download(dgot, nan)
        assert np.isfinite(ref).all() and np.ptp(ref) > 0
        np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))
    # one row past the maximum is refused before anything runs
    st = L.amca_mlp_forward_device_noalloc(net.h, MAX_ROWS + 1, dX, dgot)
    assert st != 0 and b"maximum" in L.amca_mlp_last_error()
    hip.close()
    net.close()
