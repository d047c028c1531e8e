"""Temperature surrogate (amca_mlp_* C ABI): a 5-128-128-128-128-1 SiLU MLP on
the matrix cores, trained on the GPU."""
import ctypes as C

import numpy as np

from ._lib import AmcaError, load

F32P = C.POINTER(C.c_float)

SIGNATURES = {
    "amca_mlp_param_count": (C.c_int32, []),
    "amca_mlp_create": (C.c_int, [C.c_uint64, C.POINTER(C.c_void_p)]),
    "amca_mlp_destroy": (C.c_int, [C.c_void_p]),
    "amca_mlp_set_params": (C.c_int, [C.c_void_p, F32P]),
    "amca_mlp_get_params": (C.c_int, [C.c_void_p, F32P]),
    "amca_mlp_forward": (C.c_int, [C.c_void_p, C.c_int32, F32P, F32P]),
    "amca_mlp_gradients": (C.c_int, [C.c_void_p, C.c_int32, F32P, F32P, F32P, F32P]),
    "amca_mlp_train_device": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int32,
                                        C.c_int32, C.c_float, C.c_uint64, F32P]),
    "amca_mlp_forward_device": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "amca_mlp_gemm_selftest": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                         F32P, F32P, F32P]),
    "amca_mlp_last_error": (C.c_char_p, []),
}

_bound = False


def lib():
    global _bound
    L = load()
    if not _bound:
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _bound = True
    return L


def _check(st, where):
    if st != 0:
        raise AmcaError(f"{where}: status {st}: {lib().amca_mlp_last_error().decode()}")


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def _p(a):
    return a.ctypes.data_as(F32P)


def gemm_selftest(A, B, trans_a=False, trans_b=False):
    A, B = _f32(A), _f32(B)
    M = A.shape[1] if trans_a else A.shape[0]
    K = A.shape[0] if trans_a else A.shape[1]
    N = B.shape[0] if trans_b else B.shape[1]
    Cm = np.empty((M, N), np.float32)
    _check(lib().amca_mlp_gemm_selftest(M, N, K, int(trans_a), int(trans_b), _p(A), _p(B), _p(Cm)),
           "gemm_selftest")
    return Cm


class MLP:
    def __init__(self, seed=0):
        self.h = C.c_void_p()
        _check(lib().amca_mlp_create(seed, C.byref(self.h)), "mlp_create")
        self.n_params = lib().amca_mlp_param_count()

    def params(self):
        p = np.empty(self.n_params, np.float32)
        _check(lib().amca_mlp_get_params(self.h, _p(p)), "get_params")
        return p

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.n_params
        _check(lib().amca_mlp_set_params(self.h, _p(p)), "set_params")

    def forward(self, X):
        X = _f32(X)
        t = np.empty(len(X), np.float32)
        _check(lib().amca_mlp_forward(self.h, len(X), _p(X), _p(t)), "forward")
        return t

    def gradients(self, X, Y):
        X, Y = _f32(X), _f32(Y)
        g = np.empty(self.n_params, np.float32)
        loss = C.c_float(0.0)
        _check(lib().amca_mlp_gradients(self.h, len(X), _p(X), _p(Y), _p(g), C.byref(loss)),
               "gradients")
        return g, float(loss.value)

    def close(self):
        if self.h:
            lib().amca_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
