"""Batches of plain CA domains stepped together (amca_batch_* C ABI), and the
64-domain batch workload: cubes of 64..256 cells of 2e-6 m, each with its own
rand() seed, cooling rate (log-uniform 20..2000 K/s) and bulk site density
(log-uniform 1e12..1e14 m^-3), Gaussian N(3, 1) undercooling."""
import ctypes as C

import numpy as np

from . import decks
from ._lib import AmcaError, load
from .sim import Simulation


class Schedule(C.Structure):
    _fields_ = [("list_blocks", C.c_int32), ("emit_blocks", C.c_int32),
                ("min_blocks", C.c_int32), ("interleave", C.c_int32),
                ("steps_per_graph", C.c_int32)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


SIGNATURES = {
    "amca_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(C.c_void_p)]),
    "amca_batch_destroy": (C.c_int, [C.c_void_p]),
    "amca_batch_set_schedule": (C.c_int, [C.c_void_p, C.POINTER(Schedule)]),
    "amca_batch_get_schedule": (C.c_int, [C.c_void_p, C.POINTER(Schedule)]),
    "amca_batch_run": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_int32),
                                 C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "amca_batch_autotune": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(Schedule)]),
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
        raise AmcaError(f"{where}: status {st}: {lib().amca_last_error().decode()}")


class Batch:
    """A batch over existing domains (amca3d_amd.sim.Domain views); the domains
    stay owned by their simulations."""

    def __init__(self, domains):
        arr = (C.c_void_p * len(domains))(*[d.h.value for d in domains])
        self.h = C.c_void_p()
        self.n = len(domains)
        _check(lib().amca_batch_create(arr, len(domains), C.byref(self.h)), "batch_create")

    def schedule(self):
        s = Schedule()
        _check(lib().amca_batch_get_schedule(self.h, C.byref(s)), "get_schedule")
        return s.as_dict()

    def set_schedule(self, **kw):
        s = Schedule(**{**self.schedule(), **kw})
        _check(lib().amca_batch_set_schedule(self.h, C.byref(s)), "set_schedule")

    def run(self, max_graphs=-1):
        fin, ms, rep = C.c_int32(0), C.c_double(0.0), C.c_int64(0)
        _check(lib().amca_batch_run(self.h, int(max_graphs), C.byref(fin), C.byref(ms),
                                    C.byref(rep)), "batch_run")
        return {"finished": int(fin.value), "ms": float(ms.value), "replays": int(rep.value)}

    def autotune(self, trial_graphs=1, rounds=2):
        s = Schedule()
        _check(lib().amca_batch_autotune(self.h, int(trial_graphs), int(rounds), C.byref(s)),
               "autotune")
        return s.as_dict()

    def close(self):
        if self.h:
            lib().amca_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_specs(n=64, master_seed=20261202, lo=64, hi=256):
    """Per-domain (edge, seed, Tdot, bulk density) of the batch workload."""
    rng = np.random.default_rng(master_seed)
    edges = rng.integers(lo, hi + 1, size=n)
    tdot = -10.0 ** rng.uniform(np.log10(20.0), np.log10(2000.0), size=n)
    rho = 10.0 ** rng.uniform(12.0, 14.0, size=n)
    return [dict(edge=int(edges[d]), seed=d, Tdot=float(tdot[d]), density=float(rho[d]))
            for d in range(n)]


def spec_deck(s, h=2e-6, time_step=1e-3, termination_time=10.0):
    e = s["edge"]
    return decks.plain_deck(e, e, e, h, surface=(0.0, 1.0, 0.5), bulk=(s["density"], 3.0, 1.0),
                            Tdot=s["Tdot"], T0=-0.2, termination_time=termination_time,
                            time_step=time_step, random_seed=s["seed"])


def simulations(specs, **kw):
    """Initialised simulations (host draws done, grids uploaded) for `specs`."""
    return [Simulation(spec_deck(s, **kw)).initialize() for s in specs]
