"""Python handles over the C ABI: a realm-level Simulation and a Domain view."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import AmcaError, Counters, SimInfo, StepRecord, check, load


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class Domain:
    """Non-owning view of an amca_domain handle."""

    def __init__(self, handle, shape):
        self.h = C.c_void_p(handle)
        self.nx, self.ny, self.nz = shape
        self.ng = (self.nx + 2) * (self.ny + 2) * (self.nz + 2)

    def counters(self):
        c = Counters()
        check(load().amca_domain_counters(self.h, C.byref(c)), "counters")
        return c.as_dict()

    def step(self):
        c = Counters()
        check(load().amca_domain_step(self.h, C.byref(c)), "step")
        return c.as_dict()

    def run(self, max_steps=-1):
        c = Counters()
        check(load().amca_domain_run(self.h, int(max_steps), C.byref(c)), "run")
        return c.as_dict()

    def run_fixed(self, n):
        check(load().amca_domain_run_fixed(self.h, int(n)), "run_fixed")

    def enable_log(self, cap):
        check(load().amca_domain_enable_log(self.h, int(cap)), "enable_log")

    def read_log(self, cap=1 << 20):
        n = C.c_int32(0)
        check(load().amca_domain_read_log(self.h, 0, None, C.byref(n)), "read_log")
        m = min(int(n.value), cap)
        rows = (StepRecord * max(m, 1))()
        check(load().amca_domain_read_log(self.h, m, rows, C.byref(n)), "read_log")
        return np.array([(r.time_before, r.dt, r.n_nucleated, r.n_active_before, r.n_captured,
                          r.n_active_after, r.n_liquid, r.n_wave1, r.max_wave, r.n_sites)
                         for r in rows[:m]],
                        dtype=[("time_before", "f8"), ("dt", "f8"), ("n_nucleated", "i4"),
                               ("n_active_before", "i4"), ("n_captured", "i4"),
                               ("n_active_after", "i4"), ("n_liquid", "i8"), ("n_wave1", "i4"),
                               ("max_wave", "i4"), ("n_sites", "i4")])

    def download(self, env=True, site_cap=1 << 22):
        ori = np.empty(self.ng, np.int32)
        gid = np.empty(self.ng, np.int32)
        e = np.empty(4 * self.ng, np.float64) if env else None
        sc = np.empty(site_cap, np.int32)
        sd = np.empty(site_cap, np.float64)
        ns = C.c_int32(0)
        check(load().amca_domain_download(self.h, _ptr(ori, C.c_int32), _ptr(gid, C.c_int32),
                                          _ptr(e, C.c_double) if env else None, site_cap,
                                          _ptr(sc, C.c_int32), _ptr(sd, C.c_double),
                                          C.byref(ns)), "download")
        n = min(int(ns.value), site_cap)
        out = dict(ori=ori, gid=gid, sites_cell=sc[:n].copy(), sites_dt=sd[:n].copy())
        if env:
            out["env"] = e.reshape(-1, 4)
        return out

    def download_grains(self):
        n = C.c_int32(0)
        check(load().amca_domain_download_grains(self.h, 0, None, None, None, C.byref(n)), "grains")
        m = int(n.value)
        env = np.empty((max(m, 1), 4), np.float64)
        cell = np.empty(max(m, 1), np.int32)
        ori = np.empty(max(m, 1), np.int32)
        check(load().amca_domain_download_grains(self.h, m, _ptr(env, C.c_double),
                                                 _ptr(cell, C.c_int32), _ptr(ori, C.c_int32),
                                                 C.byref(n)), "grains")
        return dict(env=env[:m], cell=cell[:m], orient=ori[:m])

    def enable_timing(self, on=True):
        check(load().amca_domain_enable_timing(self.h, int(bool(on))), "enable_timing")

    def kernel_times(self):
        L = load()
        k = L.amca_kernel_count()
        ms = np.zeros(k, np.float64)
        n = np.zeros(k, np.int64)
        check(L.amca_domain_kernel_times(self.h, _ptr(ms, C.c_double), _ptr(n, C.c_int64)),
              "kernel_times")
        return {L.amca_kernel_name(i).decode(): (float(ms[i]), int(n[i])) for i in range(k)}


class OwnedDomain(Domain):
    """An amca_domain created directly from a description (low-level ABI)."""

    def __init__(self, *, nx, ny, nz, h, orient_capacity, x0=0.0, y0=0.0, z0=0.0, T0=-0.2,
                 Tm=0.0, Tdot=-20.0, a1=-0.544e-4, a2=2.03e-4, a3=0.0, trunc=None,
                 time_step_factor=0.2, max_dt=1e-3, final_time=17.81):
        if trunc is None:
            trunc = float(np.sqrt(np.float32(3.0)))
        d = _lib.DomainDesc(nx=nx, ny=ny, nz=nz, h=h, x0=x0, y0=y0, z0=z0, T0=T0, Tm=Tm,
                            Tdot=Tdot, a1=a1, a2=a2, a3=a3, trunc=trunc,
                            time_step_factor=time_step_factor, max_dt=max_dt,
                            final_time=final_time)
        hnd = C.c_void_p()
        check(load().amca_domain_create(C.byref(d), int(orient_capacity), C.byref(hnd)), "create")
        super().__init__(hnd.value, (nx, ny, nz))

    def set_sites(self, cells, dts):
        cells = np.ascontiguousarray(cells, np.int32)
        dts = np.ascontiguousarray(dts, np.float64)
        check(load().amca_domain_set_sites(self.h, len(cells), _ptr(cells, C.c_int32),
                                           _ptr(dts, C.c_double)), "set_sites")

    def push_orientations(self, euler, l2g):
        euler = np.ascontiguousarray(euler, np.float64).reshape(-1)
        l2g = np.ascontiguousarray(l2g, np.float64).reshape(-1)
        check(load().amca_domain_push_orientations(self.h, len(l2g) // 9, _ptr(euler, C.c_double),
                                                   _ptr(l2g, C.c_double)), "push_orientations")

    def close(self):
        if self.h:
            load().amca_domain_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Simulation:
    """A YAML input deck driven through the realm layer of libamca3d."""

    def __init__(self, yaml_text=None, path=None):
        self._h = C.c_void_p()
        if path is not None:
            check(load().amca_sim_create_from_file(str(path).encode(), C.byref(self._h)),
                  "sim_create", sim=True)
        else:
            check(load().amca_sim_create(yaml_text.encode(), C.byref(self._h)), "sim_create",
                  sim=True)
        self._initialized = False

    @classmethod
    def from_file(cls, path):
        return cls(path=path)

    def prepare(self):
        """Host RNG draws only (no GPU needed)."""
        check(load().amca_sim_prepare(self._h), "sim_prepare", sim=True)
        return self

    def initialize(self):
        check(load().amca_sim_initialize(self._h), "sim_initialize", sim=True)
        self._initialized = True
        return self

    def info(self):
        i = SimInfo()
        check(load().amca_sim_get_info(self._h, C.byref(i)), "sim_info", sim=True)
        return {k: getattr(i, k) for k, _ in SimInfo._fields_}

    @property
    def domain(self):
        inf = self.info()
        return Domain(load().amca_sim_domain(self._h), (inf["nx"], inf["ny"], inf["nz"]))

    def run(self, max_steps=-1):
        c = Counters()
        check(load().amca_sim_run(self._h, int(max_steps), C.byref(c)), "sim_run", sim=True)
        return c.as_dict()

    @property
    def remelt(self):
        """Non-owning view of the coupled domain (finite_element_method runs)."""
        from .remelt import RemeltDomain
        inf = self.info()
        h = load().amca_sim_remelt(self._h)
        if not h:
            raise AmcaError("not a coupled simulation, or not initialized")
        return RemeltDomain.view(h, (inf["nx"], inf["ny"], inf["nz"]), inf["fe_nodes"])

    def enable_vtr(self, on=True):
        """The output block's VTR time series during run() and at finalize()
        (off until enabled)."""
        check(load().amca_sim_enable_vtr(self._h, 1 if on else 0), "sim_enable_vtr", sim=True)
        return self

    def finalize(self):
        check(load().amca_sim_finalize(self._h), "sim_finalize", sim=True)

    def initial_sites(self):
        n = C.c_int32(0)
        check(load().amca_sim_initial_sites(self._h, 0, None, None, None, C.byref(n)), "sites",
              sim=True)
        m = int(n.value)
        cell = np.empty(max(m, 1), np.int32)
        dt = np.empty(max(m, 1), np.float64)
        eu = np.empty(3 * max(m, 1), np.float64)
        check(load().amca_sim_initial_sites(self._h, m, _ptr(cell, C.c_int32), _ptr(dt, C.c_double),
                                            _ptr(eu, C.c_double), C.byref(n)), "sites", sim=True)
        return cell[:m], dt[:m], eu[:3 * m].reshape(-1, 3)

    def close(self):
        if self._h:
            load().amca_sim_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_final_grains(vtk_path, result_path, nx, ny, nz, x0, y0, z0, h, orient, melted,
                       euler3):
    """finalGrains.vtk / finalResult.txt of a state (amca_write_final_grains):
    orient and melted ghost-inclusive, melted None = every flag 1, euler3
    (n_orient, 3).  A None path skips that file."""
    ori = np.ascontiguousarray(orient, np.int32).ravel()
    mel = None if melted is None else np.ascontiguousarray(melted, np.int32).ravel()
    eu = np.ascontiguousarray(euler3, np.float64).reshape(-1)
    ng = (nx + 2) * (ny + 2) * (nz + 2)
    if ori.size != ng or (mel is not None and mel.size != ng):
        raise ValueError("orient / melted must hold (nx+2)(ny+2)(nz+2) cells")
    enc = (lambda p: None if p is None else str(p).encode())
    check(load().amca_write_final_grains(enc(vtk_path), enc(result_path), int(nx), int(ny),
                                         int(nz), float(x0), float(y0), float(z0), float(h),
                                         ori.ctypes.data, None if mel is None else mel.ctypes.data,
                                         eu.ctypes.data, eu.size // 3),
          "write_final_grains", sim=True)


def envelope_selftest(env4, l2g9, cell3, h, trunc):
    """Device capture test + child envelope on arrays (for parity checks)."""
    env4 = np.ascontiguousarray(env4, np.float64)
    l2g9 = np.ascontiguousarray(l2g9, np.float64)
    cell3 = np.ascontiguousarray(cell3, np.float64)
    n = env4.shape[0]
    inside = np.empty(n, np.int32)
    child = np.empty((n, 4), np.float64)
    check(load().amca_envelope_selftest(n, _ptr(env4, C.c_double), _ptr(l2g9, C.c_double),
                                        _ptr(cell3, C.c_double), float(h), float(trunc),
                                        _ptr(inside, C.c_int32), _ptr(child, C.c_double)),
          "envelope_selftest")
    return inside, child
