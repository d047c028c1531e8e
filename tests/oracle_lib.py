"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure only: the test-suite, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` use it as the checker.  The product never
imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "_build", "liboracle.so")


class PlainParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("h", C.c_double), ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double),
        ("ns", C.c_double), ("dTs_mean", C.c_double), ("dTs_sd", C.c_double),
        ("nv", C.c_double), ("dTv_mean", C.c_double), ("dTv_sd", C.c_double),
        ("surface_mode", C.c_int32), ("seed", C.c_uint32),
        ("T0", C.c_double), ("Tm", C.c_double), ("Tdot", C.c_double),
        ("a1", C.c_double), ("a2", C.c_double), ("a3", C.c_double),
        ("trunc", C.c_double),
        ("tsf", C.c_double), ("max_dt", C.c_double), ("final_time", C.c_double),
        ("max_steps", C.c_int32),
    ]


class StepRecord(C.Structure):
    _fields_ = [
        ("time_before", C.c_double), ("dt", C.c_double),
        ("n_nucleated", C.c_int32), ("n_active_before", C.c_int32),
        ("n_captured", C.c_int32), ("n_active_after", C.c_int32),
        ("n_liquid", C.c_int32), ("max_wave", C.c_int32), ("n_wave1", C.c_int32),
    ]


I32P = C.POINTER(C.c_int32)
F64P = C.POINTER(C.c_double)


class PlainOut(C.Structure):
    _fields_ = [
        ("ori", I32P), ("gid", I32P), ("nuc_dt", F64P),
        ("grain_env", F64P), ("grain_cell", I32P), ("grain_orient", I32P),
        ("grain_active", C.POINTER(C.c_uint8)), ("grain_cap", C.c_int32),
        ("euler", F64P), ("orient_cap", C.c_int32),
        ("site_cell", I32P), ("site_dt", F64P), ("site_cap", C.c_int32),
        ("steps", C.POINTER(StepRecord)), ("steps_cap", C.c_int32),
        ("ng", C.c_int32), ("n_steps", C.c_int32), ("n_records", C.c_int32),
        ("done", C.c_int32), ("n_grains", C.c_int32), ("n_orient", C.c_int32),
        ("n_active", C.c_int32), ("n_sites", C.c_int32),
        ("n_site_draws", C.c_int64), ("time", C.c_double),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.oracle_plain_run.argtypes = [C.POINTER(PlainParams), C.POINTER(PlainOut)]
        L.oracle_plain_run.restype = C.c_int
        L.oracle_plain_run_with.argtypes = [C.POINTER(PlainParams), C.c_int, I32P, F64P, C.c_int,
                                            F64P, C.POINTER(PlainOut)]
        L.oracle_plain_run_with.restype = C.c_int
        L.oracle_rotation.argtypes = [F64P, F64P]
        L.oracle_capture_test.argtypes = [F64P, F64P, F64P]
        L.oracle_capture_test.restype = C.c_int
        L.oracle_child_envelope.argtypes = [F64P, C.c_double, C.c_double, C.c_double, F64P, F64P]
        L.oracle_glibc_rand.argtypes = [C.c_uint32, C.c_int, I32P]
        L.oracle_std_normal.argtypes = [C.c_double, C.c_double, C.c_int, F64P]
        L.oracle_minstd.argtypes = [C.c_int, C.POINTER(C.c_int64)]
        L.oracle_envelope_batch.argtypes = [C.c_int, F64P, F64P, F64P, C.c_double, C.c_double,
                                            I32P, F64P]
        L.oracle_sqrtf.argtypes = [C.c_float]
        L.oracle_sqrtf.restype = C.c_float
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def plain_params(**kw):
    """Parameter block with the reference defaults of a `cellular_automata` realm."""
    d = dict(nx=20, ny=20, nz=20, h=5e-4, x0=0.0, y0=0.0, z0=0.0,
             ns=0.0, dTs_mean=0.5, dTs_sd=0.1, nv=0.0, dTv_mean=6.0, dTv_sd=0.1,
             surface_mode=0, seed=0, T0=-0.2, Tm=0.0, Tdot=-20.0,
             a1=-0.544e-4, a2=2.03e-4, a3=0.0,
             trunc=float(np.sqrt(np.float32(3.0))),
             tsf=0.2, max_dt=1e-3, final_time=17.81, max_steps=-1)
    d.update(kw)
    return PlainParams(**d)


def run_plain(prm, steps_cap=200000, want_grains=True, given=None):
    """Run the oracle to completion; returns a dict of numpy arrays.

    ``given=(site_cells, site_dts, euler_rows)`` replaces the rand() draws by
    explicit sites and orientation angles."""
    L = lib()
    nxG, nyG, nzG = prm.nx + 2, prm.ny + 2, prm.nz + 2
    ng = nxG * nyG * nzG
    ncell = prm.nx * prm.ny * prm.nz
    ori = np.empty(ng, np.int32)
    gid = np.empty(ng, np.int32)
    nuc = np.empty(ng, np.float64)
    gcap = ncell if want_grains else 0
    env = np.empty(4 * max(gcap, 1), np.float64)
    gcell = np.empty(max(gcap, 1), np.int32)
    gori = np.empty(max(gcap, 1), np.int32)
    gact = np.empty(max(gcap, 1), np.uint8)
    euler = np.empty(3 * max(ncell, 1), np.float64)
    scap = ncell
    scell = np.empty(max(scap, 1), np.int32)
    sdt = np.empty(max(scap, 1), np.float64)
    recs = (StepRecord * steps_cap)()
    out = PlainOut(ori=_p(ori, C.c_int32), gid=_p(gid, C.c_int32), nuc_dt=_p(nuc, C.c_double),
                   grain_env=_p(env, C.c_double), grain_cell=_p(gcell, C.c_int32),
                   grain_orient=_p(gori, C.c_int32), grain_active=_p(gact, C.c_uint8),
                   grain_cap=gcap, euler=_p(euler, C.c_double), orient_cap=ncell,
                   site_cell=_p(scell, C.c_int32), site_dt=_p(sdt, C.c_double), site_cap=scap,
                   steps=recs, steps_cap=steps_cap)
    if given is None:
        rc = L.oracle_plain_run(C.byref(prm), C.byref(out))
    else:
        gc = np.ascontiguousarray(given[0], np.int32)
        gd = np.ascontiguousarray(given[1], np.float64)
        ge = np.ascontiguousarray(given[2], np.float64).reshape(-1)
        rc = L.oracle_plain_run_with(C.byref(prm), len(gc), _p(gc, C.c_int32), _p(gd, C.c_double),
                                     len(ge) // 3, _p(ge, C.c_double), C.byref(out))
    assert rc == 0, f"oracle run failed ({rc})"
    ng_ = out.n_grains
    n_rec = out.n_records
    steps = np.array([(r.time_before, r.dt, r.n_nucleated, r.n_active_before, r.n_captured,
                       r.n_active_after, r.n_liquid, r.max_wave, r.n_wave1) for r in recs[:n_rec]],
                     dtype=[("time_before", "f8"), ("dt", "f8"), ("n_nucleated", "i4"),
                            ("n_active_before", "i4"), ("n_captured", "i4"),
                            ("n_active_after", "i4"), ("n_liquid", "i4"), ("max_wave", "i4"),
                            ("n_wave1", "i4")])
    res = dict(ori=ori, gid=gid, nuc_dt=nuc, steps=steps, n_steps=out.n_steps, time=out.time,
               done=out.done, n_grains=ng_, n_orient=out.n_orient, n_active=out.n_active,
               sites_cell=scell[:out.n_sites].copy(), sites_dt=sdt[:out.n_sites].copy(),
               n_site_draws=out.n_site_draws, euler=euler[:3 * out.n_orient].reshape(-1, 3).copy())
    if want_grains:
        res.update(env=env[:4 * ng_].reshape(-1, 4).copy(), grain_cell=gcell[:ng_].copy(),
                   grain_orient=gori[:ng_].copy(), grain_active=gact[:ng_].copy())
    return res


def cell_index(prm, i, j, k):
    """Local-ghost index of cell (i, j, k), 1-based interior coordinates."""
    nxG, nyG = prm.nx + 2, prm.ny + 2
    return (k * nyG + j) * nxG + i


def interior(prm, a):
    """View a ghosted per-cell array as [nz, ny, nx] interior."""
    a = np.asarray(a).reshape(prm.nz + 2, prm.ny + 2, prm.nx + 2, *np.shape(a)[1:])
    return a[1:-1, 1:-1, 1:-1]


def envelope_batch(env4, l2g9, cell3, h, trunc):
    env4 = np.ascontiguousarray(env4, np.float64)
    l2g9 = np.ascontiguousarray(l2g9, np.float64)
    cell3 = np.ascontiguousarray(cell3, np.float64)
    n = env4.shape[0]
    inside = np.empty(n, np.int32)
    child = np.empty((n, 4), np.float64)
    lib().oracle_envelope_batch(n, _p(env4, C.c_double), _p(l2g9, C.c_double), _p(cell3, C.c_double),
                                float(h), float(trunc), _p(inside, C.c_int32), _p(child, C.c_double))
    return inside, child


def rotations(euler):
    """Row-major local-to-global matrices of Euler angle rows (oracle arithmetic)."""
    euler = np.ascontiguousarray(euler, np.float64).reshape(-1, 3)
    out = np.empty((len(euler), 9), np.float64)
    for i in range(len(euler)):
        lib().oracle_rotation(_p(euler[i], C.c_double), _p(out[i], C.c_double))
    return out


# ---- coupled CA-FE (remelting) path ---------------------------------------

U8P = C.POINTER(C.c_uint8)


class RemeltParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("h", C.c_double), ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double),
        ("fnx", C.c_int32), ("fny", C.c_int32), ("fnz", C.c_int32),
        ("node_xyz", F64P), ("n_frames", C.c_int32), ("frame_time", F64P), ("frame_theta", F64P),
        ("bin", C.c_int32 * 3), ("tol", C.c_double), ("max_iter", C.c_int32),
        ("void_temp", C.c_double),
        ("Tm", C.c_double), ("Ts", C.c_double), ("Tmin_nuc", C.c_double),
        ("a1", C.c_double), ("a2", C.c_double), ("a3", C.c_double),
        ("b1", C.c_double), ("b2", C.c_double), ("b3", C.c_double),
        ("trunc", C.c_double), ("tsf", C.c_double), ("cutoff_factor", C.c_double),
        ("final_time", C.c_double), ("max_steps", C.c_int32), ("initial_melted", C.c_int32),
        ("use_sites", C.c_int32), ("ns", C.c_double), ("dTs_mean", C.c_double),
        ("dTs_sd", C.c_double), ("nv", C.c_double), ("dTv_mean", C.c_double),
        ("dTv_sd", C.c_double), ("seed", C.c_uint32),
        ("init_ori", I32P), ("init_gid", I32P), ("init_melted", U8P),
        ("n_init_orient", C.c_int32), ("init_euler", F64P),
        ("n_init_grains", C.c_int32), ("init_grain_cell", I32P), ("init_grain_orient", I32P),
        ("random_micro", C.c_int32), ("init_site_dt", C.POINTER(C.c_double)),
    ]


class RemeltStep(C.Structure):
    _fields_ = [("time_before", C.c_double), ("dt", C.c_double), ("n_nucleated", C.c_int32),
                ("n_captured", C.c_int32), ("n_fresh", C.c_int32), ("n_flagged", C.c_int32),
                ("n_liquid", C.c_int32)]


class RemeltOut(C.Structure):
    _fields_ = [
        ("ori", I32P), ("gid", I32P), ("T", F64P), ("melted", U8P), ("cooling", U8P),
        ("tag", I32P), ("par", F64P), ("theta", F64P), ("grain_env", F64P),
        ("grain_cell", I32P), ("grain_orient", I32P), ("grain_cap", C.c_int32),
        ("euler", F64P), ("orient_cap", C.c_int32),
        ("steps", C.POINTER(RemeltStep)), ("steps_cap", C.c_int32),
        ("site_cell", I32P), ("site_dt", F64P), ("site_cap", C.c_int32),
        ("n_steps", C.c_int32), ("n_records", C.c_int32), ("n_grains", C.c_int32),
        ("n_orient", C.c_int32), ("frames_exhausted", C.c_int32), ("n_sites", C.c_int32),
        ("time", C.c_double), ("time_fe", C.c_double),
    ]


def run_remelt(case, steps_cap=100000, grain_cap=None):
    """Run the remelting oracle on a case dict (see remelt_case)."""
    L = lib()
    if not hasattr(L, "_remelt_bound"):
        L.oracle_remelt_run.argtypes = [C.POINTER(RemeltParams), C.POINTER(RemeltOut)]
        L.oracle_remelt_run.restype = C.c_int
        L._remelt_bound = True
    g = case["grid"]
    nx, ny, nz = g["nx"], g["ny"], g["nz"]
    ng = (nx + 2) * (ny + 2) * (nz + 2)
    keep = []

    def arr(a, t, dt):
        a = np.ascontiguousarray(a, dt)
        keep.append(a)
        return a.ctypes.data_as(C.POINTER(t))

    fe = case["fe"]
    p = RemeltParams()
    for k in ("nx", "ny", "nz", "h", "x0", "y0", "z0"):
        setattr(p, k, g[k])
    p.fnx, p.fny, p.fnz = fe["shape"]
    p.node_xyz = arr(fe["xyz"], C.c_double, np.float64)
    p.n_frames = len(fe["times"])
    p.frame_time = arr(fe["times"], C.c_double, np.float64)
    p.frame_theta = arr(fe["theta"], C.c_double, np.float64)
    for i in range(3):
        p.bin[i] = case["bin"][i]
    p.tol, p.max_iter, p.void_temp = case["tol"], case["max_iter"], case["void_temp"]
    ph = case["physics"]
    for k in ("Tm", "Ts", "Tmin_nuc", "a1", "a2", "a3", "b1", "b2", "b3"):
        setattr(p, k, ph[k])
    for k in ("trunc", "tsf", "cutoff_factor", "final_time", "max_steps", "initial_melted"):
        setattr(p, k, case[k])
    nuc = case.get("nucleation")
    p.use_sites = 1 if nuc else 0
    if nuc:
        p.ns, p.dTs_mean, p.dTs_sd = nuc["surface"]
        p.nv, p.dTv_mean, p.dTv_sd = nuc["bulk"]
    p.seed = case.get("seed", 0)
    p.random_micro = 1 if case.get("random_micro") else 0
    if case.get("init_site_dt") is not None:
        p.init_site_dt = arr(case["init_site_dt"], C.c_double, np.float64)
    st = case.get("state")
    if st is not None:
        p.init_ori = arr(st["ori"], C.c_int32, np.int32)
        p.init_gid = arr(st["gid"], C.c_int32, np.int32)
        p.init_melted = arr(st["melted"], C.c_uint8, np.uint8)
        p.n_init_orient = len(st["euler"])
        p.init_euler = arr(st["euler"], C.c_double, np.float64)
        p.n_init_grains = len(st["grain_cell"])
        p.init_grain_cell = arr(st["grain_cell"], C.c_int32, np.int32)
        p.init_grain_orient = arr(st["grain_orient"], C.c_int32, np.int32)
    if grain_cap is None:
        n_loaded = len(st["grain_cell"]) if st is not None else nx * ny * nz
        grain_cap = n_loaded + nx * ny * nz + 16
    nnodes = int(np.prod(fe["shape"]))
    out_arrays = dict(ori=np.empty(ng, np.int32), gid=np.empty(ng, np.int32),
                      T=np.empty(ng, np.float64), melted=np.empty(ng, np.uint8),
                      cooling=np.empty(ng, np.uint8), tag=np.empty(ng, np.int32),
                      par=np.empty(3 * ng, np.float64), theta=np.empty(nnodes, np.float64),
                      grain_env=np.empty(4 * grain_cap, np.float64),
                      grain_cell=np.empty(grain_cap, np.int32),
                      grain_orient=np.empty(grain_cap, np.int32),
                      euler=np.empty(3 * grain_cap, np.float64),
                      site_cell=np.empty(nx * ny * nz, np.int32),
                      site_dt=np.empty(nx * ny * nz, np.float64))
    recs = (RemeltStep * steps_cap)()
    o = RemeltOut()
    tmap = dict(ori=C.c_int32, gid=C.c_int32, T=C.c_double, melted=C.c_uint8, cooling=C.c_uint8,
                tag=C.c_int32, par=C.c_double, theta=C.c_double, grain_env=C.c_double,
                grain_cell=C.c_int32, grain_orient=C.c_int32, euler=C.c_double,
                site_cell=C.c_int32, site_dt=C.c_double)
    for k, a in out_arrays.items():
        setattr(o, k, a.ctypes.data_as(C.POINTER(tmap[k])))
    o.grain_cap = grain_cap
    o.orient_cap = grain_cap
    o.steps = recs
    o.steps_cap = steps_cap
    o.site_cap = nx * ny * nz
    rc = L.oracle_remelt_run(C.byref(p), C.byref(o))
    assert rc == 0, f"remelt oracle failed ({rc})"
    n = o.n_grains
    assert n <= grain_cap and o.n_orient <= grain_cap, "oracle output buffers too small"
    res = {k: v for k, v in out_arrays.items()}
    res["par"] = res["par"].reshape(-1, 3)
    res["grain_env"] = res["grain_env"][:4 * n].reshape(-1, 4)
    res["grain_cell"] = res["grain_cell"][:n]
    res["grain_orient"] = res["grain_orient"][:n]
    res["euler"] = res["euler"][:3 * o.n_orient].reshape(-1, 3)
    res["site_cell"] = res["site_cell"][:o.n_sites]
    res["site_dt"] = res["site_dt"][:o.n_sites]
    res["steps"] = np.array([(r.time_before, r.dt, r.n_nucleated, r.n_captured, r.n_fresh,
                              r.n_flagged, r.n_liquid) for r in recs[:o.n_records]],
                            dtype=[("time_before", "f8"), ("dt", "f8"), ("n_nucleated", "i4"),
                                   ("n_captured", "i4"), ("n_fresh", "i4"), ("n_flagged", "i4"),
                                   ("n_liquid", "i4")])
    res.update(n_steps=o.n_steps, n_grains=n, n_orient=o.n_orient, time=o.time, time_fe=o.time_fe,
               frames_exhausted=o.frames_exhausted)
    return res
