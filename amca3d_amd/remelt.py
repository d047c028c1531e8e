"""Python handle over the coupled CA-FE (remelting) C ABI (amca_remelt_*)."""
import ctypes as C

import numpy as np

from ._lib import AmcaError, load

I32P = C.POINTER(C.c_int32)
F64P = C.POINTER(C.c_double)
U8P = C.POINTER(C.c_uint8)


class RemeltDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("h", C.c_double), ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double),
        ("Tm", C.c_double), ("Ts", C.c_double), ("Tmin_nuc", C.c_double),
        ("a1", C.c_double), ("a2", C.c_double), ("a3", C.c_double),
        ("b1", C.c_double), ("b2", C.c_double), ("b3", C.c_double),
        ("trunc", C.c_double), ("time_step_factor", C.c_double), ("cutoff_factor", C.c_double),
        ("final_time", C.c_double), ("void_temperature", C.c_double),
        ("bin", C.c_int32 * 3), ("tolerance", C.c_double), ("max_iterations", C.c_int32),
        ("initial_melted", C.c_int32),
    ]


class RemeltCounters(C.Structure):
    _fields_ = [("step", C.c_int64), ("time", C.c_double), ("time_fe", C.c_double),
                ("dt", C.c_double), ("max_dt", C.c_double), ("n_grains", C.c_int32),
                ("n_orient", C.c_int32), ("done", C.c_int32), ("frames_exhausted", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class RemeltStepRecord(C.Structure):
    _fields_ = [("time_before", C.c_double), ("dt", C.c_double), ("n_nucleated", C.c_int32),
                ("n_captured", C.c_int32), ("n_fresh", C.c_int32), ("n_flagged", C.c_int32),
                ("n_fallback", C.c_int32), ("n_box", C.c_int32)]


ORIENT_SOURCE = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int32, F64P, F64P)

SIGNATURES = {
    "amca_remelt_create": (C.c_int, [C.POINTER(RemeltDesc), C.c_int32, C.POINTER(C.c_void_p)]),
    "amca_remelt_destroy": (C.c_int, [C.c_void_p]),
    "amca_remelt_set_fe": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, F64P, C.c_int32,
                                     F64P, F64P]),
    "amca_remelt_voxel_map": (C.c_int, [C.c_void_p, I32P, F64P]),
    "amca_fe_map_voxels": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                     C.c_double, C.c_double, I32P, C.c_double, C.c_int32,
                                     C.c_int32, C.c_int32, C.c_int32, F64P, I32P, F64P]),
    "amca_fe_map_voxels_device": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                            C.c_double, C.c_double, C.c_double, I32P, C.c_double,
                                            C.c_int32, C.c_int32, C.c_int32, C.c_int32, F64P,
                                            I32P, F64P]),
    "amca_remelt_set_state": (C.c_int, [C.c_void_p, I32P, I32P, U8P, C.c_int32, C.c_int32, I32P,
                                        I32P]),
    "amca_remelt_push_orientations": (C.c_int, [C.c_void_p, C.c_int32, F64P, F64P]),
    "amca_remelt_set_orientation_source": (C.c_int, [C.c_void_p, ORIENT_SOURCE, C.c_void_p]),
    "amca_remelt_set_sites": (C.c_int, [C.c_void_p, C.c_int32, I32P, F64P]),
    "amca_remelt_run": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(RemeltCounters)]),
    "amca_remelt_get_counters": (C.c_int, [C.c_void_p, C.POINTER(RemeltCounters)]),
    "amca_remelt_enable_log": (C.c_int, [C.c_void_p, C.c_int32]),
    "amca_remelt_read_log": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(RemeltStepRecord), I32P]),
    "amca_remelt_download": (C.c_int, [C.c_void_p, I32P, I32P, F64P, U8P, U8P, C.c_int32, F64P,
                                       I32P, I32P, I32P]),
    "amca_remelt_theta": (C.c_int, [C.c_void_p, F64P]),
    "amca_remelt_draw_sites": (C.c_int, [C.c_void_p, C.c_uint32, C.c_double, C.c_double,
                                         C.c_double, C.c_double, C.c_double, C.c_double, I32P]),
    "amca_remelt_euler": (C.c_int, [C.c_void_p, C.c_int32, F64P, I32P]),
    "amca_remelt_draw_sites_after": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int64, C.c_double,
                                               C.c_double, C.c_double, C.c_double, C.c_double,
                                               C.c_double, I32P, I32P]),
    "amca_remelt_sites": (C.c_int, [C.c_void_p, C.c_int32, I32P, F64P, I32P]),
    "amca_remelt_enable_timing": (C.c_int, [C.c_void_p, C.c_int32]),
    "amca_remelt_set_window": (C.c_int, [C.c_void_p, C.c_int32]),
    "amca_remelt_last_box": (C.c_int, [C.c_void_p, I32P]),
    "amca_remelt_set_graph_steps": (C.c_int, [C.c_void_p, C.c_int32]),
    "amca_remelt_set_row_reserve": (C.c_int, [C.c_void_p, C.c_int32]),
    "amca_remelt_run_many": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32, C.c_int64,
                                       C.POINTER(RemeltCounters)]),
    "amca_remelt_use_surrogate": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_double, C.c_double]),
    "amca_remelt_surrogate_error": (C.c_int, [C.c_void_p, F64P]),
    "amca_remelt_kernel_count": (C.c_int32, []),
    "amca_remelt_kernel_name": (C.c_char_p, [C.c_int32]),
    "amca_remelt_kernel_times": (C.c_int, [C.c_void_p, F64P, C.POINTER(C.c_int64)]),
    "amca_rosenthal_frames": (C.c_int, [C.c_int64, F64P, C.c_int32, C.c_double, C.c_double,
                                        C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_double, F64P]),
    "amca_voronoi_substrate": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                         C.c_double, C.c_double, C.c_int32, C.c_uint64, I32P,
                                         F64P]),
    "amca_write_fe_frames": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                       F64P, C.c_int32, F64P, F64P]),
    "amca_pow_selftest": (C.c_int, [C.c_int64, F64P, C.c_double, F64P]),
    "amca_remelt_last_error": (C.c_char_p, []),
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


def _check(status, where):
    if status != 0:
        raise AmcaError(f"{where}: status {status}: {lib().amca_remelt_last_error().decode()}")


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def run_many(domains, max_steps=-1):
    """Run several RemeltDomain objects at once (amca_remelt_run_many); returns
    each domain's counters."""
    n = len(domains)
    hs = (C.c_void_p * n)(*[d.h.value for d in domains])
    outs = (RemeltCounters * n)()
    _check(lib().amca_remelt_run_many(hs, n, int(max_steps), outs), "run_many")
    return [o.as_dict() for o in outs]


def pow_selftest(x, y):
    """pow(x, y) element-wise by the device code of the remelting velocity."""
    x = np.ascontiguousarray(x, np.float64)
    out = np.empty_like(x)
    _check(lib().amca_pow_selftest(x.size, _p(x, C.c_double), float(y), _p(out, C.c_double)),
           "pow_selftest")
    return out


def map_voxels(grid, fe, bin=(5, 5, 5), tol=1e-16, max_iter=20, device=False):
    """Voxel mapping (element tag and parametric coordinates per cell): on the
    host, or with the per-cell search on the GPU (device=True)."""
    nx, ny, nz = grid["nx"], grid["ny"], grid["nz"]
    ng = (nx + 2) * (ny + 2) * (nz + 2)
    tag = np.empty(ng, np.int32)
    par = np.empty(3 * ng, np.float64)
    b = (C.c_int32 * 3)(*bin)
    xyz = np.ascontiguousarray(fe["xyz"], np.float64)
    fn = lib().amca_fe_map_voxels_device if device else lib().amca_fe_map_voxels
    _check(fn(nx, ny, nz, grid["h"], grid["x0"], grid["y0"], grid["z0"], b, tol, max_iter,
              *fe["shape"], _p(xyz, C.c_double), _p(tag, C.c_int32), _p(par, C.c_double)),
           "map_voxels")
    return tag, par.reshape(-1, 3)


def rosenthal_frames(xyz, times, *, P, eta, v, k, alpha, T0, x_start, y_c, z_top, r_min):
    """Nodal frames [len(times), n_nodes] of a moving point source (host, glibc exp).
    `times` must be evenly spaced."""
    xyz = np.ascontiguousarray(xyz, np.float64)
    times = np.asarray(times, np.float64)
    n = len(xyz)
    dt = float(times[1] - times[0]) if len(times) > 1 else 0.0
    out = np.empty((len(times), n), np.float64)
    _check(lib().amca_rosenthal_frames(n, _p(xyz, C.c_double), len(times), float(times[0]), dt,
                                       P, eta, v, k, alpha, T0, x_start, y_c, z_top, r_min,
                                       _p(out, C.c_double)), "rosenthal_frames")
    return out


def write_fe_frames(path, shape, xyz, times, theta, title_lines=4):
    """Writes frames [n_frames, nodes] on a (fnx, fny, fnz) node mesh as the
    text a finite_element realm reads (coordinates and absolute times as given)."""
    fnx, fny, fnz = (int(v) for v in shape)
    xyz = np.ascontiguousarray(xyz, np.float64)
    times = np.ascontiguousarray(times, np.float64)
    theta = np.ascontiguousarray(theta, np.float64)
    nodes = fnx * fny * fnz
    if xyz.shape != (nodes, 3) or theta.shape != (len(times), nodes):
        raise ValueError("write_fe_frames: xyz must be [nodes, 3] and theta [frames, nodes]")
    _check(lib().amca_write_fe_frames(str(path).encode(), int(title_lines), fnx, fny, fnz,
                                      _p(xyz, C.c_double), len(times), _p(times, C.c_double),
                                      _p(theta, C.c_double)), "write_fe_frames")


def voronoi_substrate(grid, n_seeds, seed):
    """(label[NG] with ghosts -1, euler[n_seeds, 3]) of a Voronoi substrate (host)."""
    nx, ny, nz = grid["nx"], grid["ny"], grid["nz"]
    ng = (nx + 2) * (ny + 2) * (nz + 2)
    lab = np.empty(ng, np.int32)
    eu = np.empty((n_seeds, 3), np.float64)
    _check(lib().amca_voronoi_substrate(nx, ny, nz, grid["h"], grid["x0"], grid["y0"], grid["z0"],
                                        n_seeds, seed, _p(lab, C.c_int32), _p(eu, C.c_double)),
           "voronoi_substrate")
    return lab, eu


class RemeltDomain:
    def __init__(self, desc_kw, orient_capacity):
        d = RemeltDesc()
        for k, v in desc_kw.items():
            if k == "bin":
                for q in range(3):
                    d.bin[q] = v[q]
            else:
                setattr(d, k, v)
        self.grid = (d.nx, d.ny, d.nz)
        self.ng = (d.nx + 2) * (d.ny + 2) * (d.nz + 2)
        self.h = C.c_void_p()
        _check(lib().amca_remelt_create(C.byref(d), int(orient_capacity), C.byref(self.h)), "create")
        self._keep = []

    def set_fe(self, fe):
        xyz = np.ascontiguousarray(fe["xyz"], np.float64)
        tt = np.ascontiguousarray(fe["times"], np.float64)
        th = np.ascontiguousarray(fe["theta"], np.float64)
        self.nodes = int(np.prod(fe["shape"]))
        _check(lib().amca_remelt_set_fe(self.h, *fe["shape"], _p(xyz, C.c_double), len(tt),
                                        _p(tt, C.c_double), _p(th, C.c_double)), "set_fe")

    def voxel_map(self):
        tag = np.empty(self.ng, np.int32)
        par = np.empty(3 * self.ng, np.float64)
        _check(lib().amca_remelt_voxel_map(self.h, _p(tag, C.c_int32), _p(par, C.c_double)), "voxel_map")
        return tag, par.reshape(-1, 3)

    def push_orientations(self, euler, l2g=None):
        """Append orientation rows; without `l2g` the library computes the
        rotations on the host (glibc cos/sin, as the reference does)."""
        e = np.ascontiguousarray(euler, np.float64).reshape(-1)
        r = None if l2g is None else np.ascontiguousarray(l2g, np.float64).reshape(-1)
        _check(lib().amca_remelt_push_orientations(self.h, len(e) // 3, _p(e, C.c_double),
                                                   None if r is None else _p(r, C.c_double)),
               "push_orientations")

    def set_orientation_source(self, fn):
        """fn(n) -> (euler[n,3], l2g[n,9]) continuing the rand() stream."""
        def cb(user, n, eu, r):
            try:
                a, b = fn(int(n))
                C.memmove(eu, np.ascontiguousarray(a, np.float64).ctypes.data, 8 * 3 * n)
                C.memmove(r, np.ascontiguousarray(b, np.float64).ctypes.data, 8 * 9 * n)
                return 0
            except Exception:
                return 1
        f = ORIENT_SOURCE(cb)
        self._keep.append(f)
        _check(lib().amca_remelt_set_orientation_source(self.h, f, None), "set_source")

    def set_state(self, st):
        o = np.ascontiguousarray(st["ori"], np.int32)
        g = np.ascontiguousarray(st["gid"], np.int32)
        m = np.ascontiguousarray(st["melted"], np.uint8)
        gc = np.ascontiguousarray(st["grain_cell"], np.int32)
        go = np.ascontiguousarray(st["grain_orient"], np.int32)
        _check(lib().amca_remelt_set_state(self.h, _p(o, C.c_int32), _p(g, C.c_int32),
                                           _p(m, C.c_uint8), len(st["euler"]), len(gc),
                                           _p(gc, C.c_int32), _p(go, C.c_int32)), "set_state")

    def set_sites(self, cells, dts):
        c = np.ascontiguousarray(cells, np.int32)
        d = np.ascontiguousarray(dts, np.float64)
        _check(lib().amca_remelt_set_sites(self.h, len(c), _p(c, C.c_int32), _p(d, C.c_double)),
               "set_sites")

    def draw_sites(self, seed, surface, bulk):
        """The reference sampler on the host rand() stream: surface = (density,
        mean, sd) on the six faces, bulk = (density, mean, sd); later nuclei take
        their orientations from the same stream."""
        n = C.c_int32(0)
        _check(lib().amca_remelt_draw_sites(self.h, seed, *surface, *bulk, C.byref(n)), "draw_sites")
        return int(n.value)

    def euler(self):
        n = C.c_int32(0)
        _check(lib().amca_remelt_euler(self.h, 0, None, C.byref(n)), "euler")
        out = np.empty((max(int(n.value), 1), 3), np.float64)
        _check(lib().amca_remelt_euler(self.h, int(n.value), _p(out, C.c_double), C.byref(n)),
               "euler")
        return out[:int(n.value)]

    def enable_log(self, cap):
        _check(lib().amca_remelt_enable_log(self.h, int(cap)), "enable_log")

    def read_log(self):
        n = C.c_int32(0)
        _check(lib().amca_remelt_read_log(self.h, 0, None, C.byref(n)), "read_log")
        m = int(n.value)
        rows = (RemeltStepRecord * max(m, 1))()
        _check(lib().amca_remelt_read_log(self.h, m, rows, C.byref(n)), "read_log")
        return np.array([(r.time_before, r.dt, r.n_nucleated, r.n_captured, r.n_fresh, r.n_flagged,
                          r.n_fallback, r.n_box) for r in rows[:m]],
                        dtype=[("time_before", "f8"), ("dt", "f8"), ("n_nucleated", "i4"),
                               ("n_captured", "i4"), ("n_fresh", "i4"), ("n_flagged", "i4"),
                               ("n_fallback", "i4"), ("n_box", "i4")])

    def run(self, max_steps=-1):
        c = RemeltCounters()
        _check(lib().amca_remelt_run(self.h, int(max_steps), C.byref(c)), "run")
        return c.as_dict()

    def counters(self):
        c = RemeltCounters()
        _check(lib().amca_remelt_get_counters(self.h, C.byref(c)), "counters")
        return c.as_dict()

    def download(self, grain_cap=None):
        ng = self.ng
        out = dict(ori=np.empty(ng, np.int32), gid=np.empty(ng, np.int32),
                   T=np.empty(ng, np.float64), melted=np.empty(ng, np.uint8),
                   cooling=np.empty(ng, np.uint8))
        n = C.c_int32(0)
        _check(lib().amca_remelt_download(self.h, None, None, None, None, None, 0, None, None, None,
                                          C.byref(n)), "download")
        m = int(n.value) if grain_cap is None else min(grain_cap, int(n.value))
        env = np.empty(4 * max(m, 1), np.float64)
        gc = np.empty(max(m, 1), np.int32)
        go = np.empty(max(m, 1), np.int32)
        _check(lib().amca_remelt_download(self.h, _p(out["ori"], C.c_int32), _p(out["gid"], C.c_int32),
                                          _p(out["T"], C.c_double), _p(out["melted"], C.c_uint8),
                                          _p(out["cooling"], C.c_uint8), m, _p(env, C.c_double),
                                          _p(gc, C.c_int32), _p(go, C.c_int32), C.byref(n)),
               "download")
        out.update(grain_env=env[:4 * m].reshape(-1, 4), grain_cell=gc[:m], grain_orient=go[:m],
                   n_grains=int(n.value))
        return out

    @classmethod
    def view(cls, handle, shape, nodes):
        """A handle owned elsewhere (a coupled simulation's domain)."""
        d = cls.__new__(cls)
        d.h = C.c_void_p(handle)
        d.grid = tuple(shape)
        d.ng = (shape[0] + 2) * (shape[1] + 2) * (shape[2] + 2)
        d.nodes = int(nodes)
        d._keep = []
        d._owned = False
        return d

    def sites(self):
        n = C.c_int32(0)
        _check(lib().amca_remelt_sites(self.h, 0, None, None, C.byref(n)), "sites")
        m = int(n.value)
        c = np.empty(max(m, 1), np.int32)
        t = np.empty(max(m, 1), np.float64)
        _check(lib().amca_remelt_sites(self.h, m, _p(c, C.c_int32), _p(t, C.c_double), C.byref(n)),
               "sites")
        return c[:m], t[:m]

    def use_surrogate(self, mlp, desc, scan, band=50.0, resid=2.0):
        """Temperatures from the trained network `mlp` (surrogate.MLP) with the
        FE fallback rule; `desc` and `scan` are surrogate.SurrogateDesc / Scan."""
        _check(lib().amca_remelt_use_surrogate(self.h, mlp.h, C.byref(desc), C.byref(scan),
                                               float(band), float(resid)), "use_surrogate")

    def surrogate_error(self):
        o = np.zeros(6, np.float64)
        _check(lib().amca_remelt_surrogate_error(self.h, _p(o, C.c_double)), "surrogate_error")
        return dict(rms_all=o[0], n_all=int(o[1]), rms_served=o[2], n_served=int(o[3]),
                    max_served=o[4], n_fallback=int(o[5]))

    def enable_timing(self, on=True):
        _check(lib().amca_remelt_enable_timing(self.h, 1 if on else 0), "enable_timing")

    def set_window(self, on=True):
        """Active box on (default) or every step over the whole grid."""
        _check(lib().amca_remelt_set_window(self.h, 1 if on else 0), "set_window")

    def set_graph_steps(self, n):
        """Steps per HIP graph in untimed runs (0: plain stream launches)."""
        _check(lib().amca_remelt_set_graph_steps(self.h, int(n)), "set_graph_steps")

    def set_row_reserve(self, n):
        """Orientation rows drawn ahead of the rows in use."""
        _check(lib().amca_remelt_set_row_reserve(self.h, int(n)), "set_row_reserve")

    def last_box(self):
        b = np.zeros(6, np.int32)
        _check(lib().amca_remelt_last_box(self.h, _p(b, C.c_int32)), "last_box")
        return b

    def kernel_times(self):
        """{kernel name: (total ms, launches)} since timing was enabled."""
        L = lib()
        n = L.amca_remelt_kernel_count()
        ms = np.zeros(n, np.float64)
        cnt = np.zeros(n, np.int64)
        _check(L.amca_remelt_kernel_times(self.h, _p(ms, C.c_double), _p(cnt, C.c_int64)),
               "kernel_times")
        return {L.amca_remelt_kernel_name(i).decode(): (float(ms[i]), int(cnt[i]))
                for i in range(n)}

    def theta(self):
        t = np.empty(self.nodes, np.float64)
        _check(lib().amca_remelt_theta(self.h, _p(t, C.c_double)), "theta")
        return t

    def close(self):
        if self.h and getattr(self, "_owned", True):
            lib().amca_remelt_destroy(self.h)
        self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
