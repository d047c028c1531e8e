"""Temperature surrogate (amca_mlp_* C ABI): a 5-128-128-128-128-1 SiLU MLP on
the matrix cores, trained on the GPU."""
import ctypes as C

import numpy as np

from ._lib import AmcaError, load

F32P = C.POINTER(C.c_float)
F64P = C.POINTER(C.c_double)


class SurrogateDesc(C.Structure):
    _fields_ = [("fnx", C.c_int32), ("fny", C.c_int32), ("fnz", C.c_int32),
                ("fx0", C.c_double), ("fy0", C.c_double), ("fz0", C.c_double), ("fdx", C.c_double),
                ("n_frames", C.c_int32), ("frame_dt", C.c_double),
                ("eta", C.c_double), ("k", C.c_double), ("alpha", C.c_double), ("T0", C.c_double),
                ("z_top", C.c_double), ("r_min", C.c_double), ("void_temperature", C.c_double),
                ("xi_scale", C.c_double), ("y_scale", C.c_double), ("z_scale", C.c_double),
                ("P_scale", C.c_double), ("v_scale", C.c_double),
                ("T_ref", C.c_double), ("T_scale", C.c_double), ("T_clip", C.c_double),
                ("log_target", C.c_int32)]


class Scan(C.Structure):
    _fields_ = [("P", C.c_double), ("v", C.c_double), ("x_start", C.c_double), ("y_c", C.c_double)]

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
    "amca_mlp_forward_device_noalloc": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p,
                                                  C.c_void_p]),
    "amca_mlp_forward_fused": (C.c_int, [C.c_void_p, C.c_int32, F32P, F32P]),
    "amca_mlp_enable_timing": (C.c_int, [C.c_void_p, C.c_int32]),
    "amca_mlp_fused_times": (C.c_int, [C.c_void_p, F64P, C.POINTER(C.c_int64),
                                       C.POINTER(C.c_int64)]),
    "amca_mlp_gemm_selftest": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                         F32P, F32P, F32P]),
    "amca_mlp_last_error": (C.c_char_p, []),
    "amca_surrogate_sample": (C.c_int, [C.c_void_p, C.POINTER(SurrogateDesc), C.POINTER(Scan),
                                        C.c_int32, C.c_int64, C.c_double, C.c_uint64]),
    "amca_surrogate_train": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_float, C.c_uint64,
                                       F32P]),
    "amca_surrogate_eval_lattice": (C.c_int, [C.c_void_p, C.POINTER(SurrogateDesc), C.POINTER(Scan),
                                              C.c_int32, C.c_double, C.c_double, C.c_int32,
                                              C.c_double, C.c_double, C.c_int32, C.c_double,
                                              C.c_double, F32P]),
    "amca_surrogate_eval": (C.c_int, [C.c_void_p, C.POINTER(SurrogateDesc), C.POINTER(Scan),
                                      C.c_int32, F64P, F64P, F64P]),
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

    def forward_fused(self, X):
        """The forward pass as one fused kernel (amca_mlp_forward_fused)."""
        X = _f32(X)
        t = np.empty(len(X), np.float32)
        _check(lib().amca_mlp_forward_fused(self.h, len(X), _p(X), _p(t)), "forward_fused")
        return t

    def gradients(self, X, Y):
        X, Y = _f32(X), _f32(Y)
        g = np.empty(self.n_params, np.float32)
        loss = C.c_float(0.0)
        _check(lib().amca_mlp_gradients(self.h, len(X), _p(X), _p(Y), _p(g), C.byref(loss)),
               "gradients")
        return g, float(loss.value)

    def sample(self, desc, scans, n, near_fraction=0.7, seed=1):
        arr = (Scan * len(scans))(*scans)
        _check(lib().amca_surrogate_sample(self.h, C.byref(desc), arr, len(scans), int(n),
                                           float(near_fraction), seed), "sample")

    def train(self, batch=8192, steps=2000, lr=1e-3, seed=2, trace=True):
        tr = np.zeros(max(steps, 1), np.float32)
        _check(lib().amca_surrogate_train(self.h, batch, steps, lr, seed,
                                          _p(tr) if trace else None), "train")
        return tr[:steps]

    def eval_lattice(self, desc, scan, nxi, xi0, dxi, ny, y0, dy, nz, z0, dz):
        """Fused-kernel temperatures (kelvin) on a moving-frame lattice at t = 0,
        shape (nz, ny, nxi); point (i, j, k) is (x_start + xi0 + i dxi, y0 + j dy, z0 + k dz)."""
        out = np.empty(int(nxi) * int(ny) * int(nz), np.float32)
        _check(lib().amca_surrogate_eval_lattice(self.h, C.byref(desc), C.byref(scan), int(nxi),
                                                 float(xi0), float(dxi), int(ny), float(y0),
                                                 float(dy), int(nz), float(z0), float(dz), _p(out)),
               "eval_lattice")
        return out.reshape(int(nz), int(ny), int(nxi))

    def eval(self, desc, scan, xyzt):
        """(surrogate, FE path) temperatures in kelvin at rows (x, y, z, t)."""
        p = np.ascontiguousarray(xyzt, np.float64)
        n = len(p)
        ts = np.empty(n, np.float64)
        tf = np.empty(n, np.float64)
        _check(lib().amca_surrogate_eval(self.h, C.byref(desc), C.byref(scan), n,
                                         p.ctypes.data_as(F64P), ts.ctypes.data_as(F64P),
                                         tf.ctypes.data_as(F64P)), "eval")
        return ts, tf

    def enable_timing(self, on=True):
        """HIP events around the fused kernel's launches (resets the totals)."""
        _check(lib().amca_mlp_enable_timing(self.h, 1 if on else 0), "enable_timing")

    def fused_times(self):
        """(milliseconds, launches, rows) of the fused kernel since enable_timing."""
        ms = C.c_double(0.0)
        launches, rows = C.c_int64(0), C.c_int64(0)
        _check(lib().amca_mlp_fused_times(self.h, C.byref(ms), C.byref(launches), C.byref(rows)),
               "fused_times")
        return ms.value, launches.value, rows.value

    def close(self):
        if self.h:
            lib().amca_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
