"""Builds tests/pow_host_check.cpp (the device velocity's pow restatement
compiled for the host, and an argument generator) into a temporary library."""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def build(dirpath):
    so = os.path.join(str(dirpath), "pow_host_check.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-o", so,
                    os.path.join(ROOT, "tests", "pow_host_check.cpp"), "-lm"], check=True)
    L = C.CDLL(so)
    L.pow_mismatches.restype = C.c_long
    L.pow_mismatches.argtypes = [C.c_long, C.c_double, C.c_int, C.c_int, C.c_uint64, C.c_int,
                                 C.c_double, C.POINTER(C.c_long)]
    L.pow_fast_mismatches.restype = C.c_long
    L.pow_fast_mismatches.argtypes = L.pow_mismatches.argtypes
    L.pow_args.restype = None
    L.pow_args.argtypes = [C.c_long, C.c_double, C.c_int, C.c_int, C.c_uint64, C.c_int,
                           C.c_double, C.POINTER(C.c_double)]

    L.glibc_pow_array.restype = None
    L.glibc_pow_array.argtypes = [C.c_long, C.POINTER(C.c_double), C.c_double,
                                  C.POINTER(C.c_double)]

    def glibc_pow(x, y):
        x = np.ascontiguousarray(x, np.float64)
        out = np.empty_like(x)
        L.glibc_pow_array(x.size, x.ctypes.data_as(C.POINTER(C.c_double)), y,
                          out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def mismatches(n, y, lo_e, hi_e, seed=1, hard=0, band=0.0):
        naive = C.c_long(0)
        bad = L.pow_mismatches(n, y, lo_e, hi_e, seed, hard, band, C.byref(naive))
        return bad, naive.value

    def fast_mismatches(n, y, lo_e, hi_e, seed=1, hard=0, band=0.0):
        sc = C.c_long(0)
        bad = L.pow_fast_mismatches(n, y, lo_e, hi_e, seed, hard, band, C.byref(sc))
        return bad, sc.value
    mismatches.fast = fast_mismatches

    def args(n, y, lo_e, hi_e, seed=1, hard=0, band=0.0):
        out = np.empty(n, np.float64)
        L.pow_args(n, y, lo_e, hi_e, seed, hard, band, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out
    return mismatches, args, glibc_pow
