"""A single-track PBF melt-and-resolidify workload for the coupled CA-FE domain.

Grid of `nx x ny x nz` cells; an FE hex mesh with nodes every `node_spacing`
covering the interior box; nodal frames of a Rosenthal moving point source
(host generator, glibc exp); a Voronoi substrate in which every interior cell
carries its own grain record (id = interior-cell order) with the orientation of
its Voronoi seed.  Default sizes: 1000 x 400 x 200 cells of 2.5e-6 m, nodes every
10e-6 m (251 x 101 x 51), 200 W at 1 m/s, 2000 substrate grains.
"""
import numpy as np

from . import remelt


def box_mesh(grid, node_spacing):
    """Structured nodes from the box origin to its far corner (node x fastest)."""
    cnt = [int(round(grid[n] * grid["h"] / node_spacing)) + 1 for n in ("nx", "ny", "nz")]
    ax = [grid[o] + node_spacing * np.arange(cnt[q]) for q, o in enumerate(("x0", "y0", "z0"))]
    zz, yy, xx = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return dict(shape=tuple(cnt), xyz=np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=1),
                axes=ax)


def substrate(grid, n_grains, seed):
    """Loaded state: per-cell orientation (Voronoi label), one grain record per
    interior cell (ids in cell order), nothing melted."""
    lab, euler = remelt.voronoi_substrate(grid, n_grains, seed)
    cells = np.flatnonzero(lab >= 0).astype(np.int32)
    gid = np.full(lab.shape, -1, np.int32)
    gid[cells] = np.arange(len(cells), dtype=np.int32)
    return dict(ori=lab, gid=gid, melted=np.zeros(lab.shape, np.uint8), euler=euler,
                grain_cell=cells, grain_orient=lab[cells].copy())


def track(nx=1000, ny=400, nz=200, h=2.5e-6, node_spacing=10e-6, n_frames=100, frame_dt=10e-6,
          P=200.0, eta=0.35, v=1.0, k=20.0, alpha=5e-6, T0=300.0, x_start=0.0, n_grains=2000,
          substrate_seed=20261202, nuc_seed=0, bulk=(1e13, 5.0, 1.0), surface=(0.0, 1.0, 0.5),
          Tm=1600.0, Ts=1550.0, void_temperature=200.0, final_time=1.0):
    grid = dict(nx=nx, ny=ny, nz=nz, h=h, x0=0.0, y0=0.0, z0=0.0)
    fe = box_mesh(grid, node_spacing)
    times = frame_dt * np.arange(n_frames)
    fe["times"] = times
    fe["theta"] = remelt.rosenthal_frames(fe["xyz"], times, P=P, eta=eta, v=v, k=k, alpha=alpha,
                                          T0=T0, x_start=x_start, y_c=0.5 * ny * h,
                                          z_top=nz * h, r_min=h)
    desc = dict(nx=nx, ny=ny, nz=nz, h=h, x0=0.0, y0=0.0, z0=0.0, Tm=Tm, Ts=Ts, Tmin_nuc=Ts,
                a1=-0.544e-4, a2=2.03e-4, a3=0.0, b1=1.0, b2=2.0, b3=3.0,
                trunc=float(np.sqrt(np.float32(3.0))), time_step_factor=0.2, cutoff_factor=1e8,
                final_time=final_time, void_temperature=void_temperature, bin=(5, 5, 5),
                tolerance=1e-16, max_iterations=20, initial_melted=0)
    return dict(grid=grid, fe=fe, desc=desc, state=substrate(grid, n_grains, substrate_seed),
                nuc_seed=nuc_seed, bulk=bulk, surface=surface,
                physics=dict(P=P, eta=eta, v=v, k=k, alpha=alpha, T0=T0, x_start=x_start))


def surrogate_desc(w, xi_scale=1e-4, T_clip=2600.0, log_target=1, T_scale=100.0):
    """Surrogate description of workload `w` (mesh box, frames, source physics,
    normalisation)."""
    from .surrogate import SurrogateDesc
    fe, g, ph = w["fe"], w["grid"], w["physics"]
    ax = fe["axes"]
    return SurrogateDesc(fnx=fe["shape"][0], fny=fe["shape"][1], fnz=fe["shape"][2],
                         fx0=ax[0][0], fy0=ax[1][0], fz0=ax[2][0], fdx=ax[0][1] - ax[0][0],
                         n_frames=len(fe["times"]), frame_dt=fe["times"][1] - fe["times"][0],
                         eta=ph["eta"], k=ph["k"], alpha=ph["alpha"], T0=ph["T0"],
                         z_top=g["nz"] * g["h"], r_min=g["h"],
                         void_temperature=w["desc"]["void_temperature"],
                         xi_scale=xi_scale, y_scale=xi_scale, z_scale=xi_scale, P_scale=300.0,
                         v_scale=1.5, T_ref=ph["T0"], T_scale=T_scale, T_clip=T_clip,
                         log_target=log_target)


def scan(w):
    from .surrogate import Scan
    ph = w["physics"]
    return Scan(P=ph["P"], v=ph["v"], x_start=ph["x_start"],
                y_c=ph.get("y_c", 0.5 * w["grid"]["ny"] * w["grid"]["h"]))


def scan_track(w, s):
    """Workload `w` with the frames of scan `s` (power, speed, start, hatch
    position): same grid, mesh, substrate and nucleation."""
    g, ph = w["grid"], w["physics"]
    fe = dict(w["fe"])
    fe["theta"] = remelt.rosenthal_frames(fe["xyz"], fe["times"], P=s.P, eta=ph["eta"], v=s.v,
                                          k=ph["k"], alpha=ph["alpha"], T0=ph["T0"],
                                          x_start=s.x_start, y_c=s.y_c, z_top=g["nz"] * g["h"],
                                          r_min=g["h"])
    out = dict(w, fe=fe)
    out["physics"] = dict(ph, P=s.P, v=s.v, x_start=s.x_start, y_c=s.y_c)
    return out


def train_surrogate(w, n_samples=20_000_000, seed=3, schedule=((8192, 6000, 1e-3),
                                                               (8192, 3000, 2e-4),
                                                               (16384, 2000, 5e-5))):
    """A surrogate trained on the GPU on samples of the FE field over the six
    held-in scans; returns (network, description, per-stage final losses)."""
    from .surrogate import MLP
    desc = surrogate_desc(w)
    m = MLP(seed=seed)
    m.sample(desc, held_in_scans(w), int(n_samples), near_fraction=0.6, seed=seed + 2)
    losses = []
    for q, (batch, steps, lr) in enumerate(schedule):
        tr = m.train(batch=batch, steps=steps, lr=lr, seed=seed + 10 + q)
        losses.append(float(tr[-50:].mean()))
    return m, desc, losses


def held_out_scans(w):
    """The workload's own track and eight variants (power, speed, hatch offset)."""
    from .surrogate import Scan
    g = w["grid"]
    yc = 0.5 * g["ny"] * g["h"]
    ph = w["physics"]
    v = [(ph["P"], ph["v"], 0.0)] + [(P, s, dy) for P, s, dy in
                                     [(160.0, 0.7, 0.0), (200.0, 0.5, 3e-5), (210.0, 1.4, -3e-5),
                                      (240.0, 1.0, 6e-5), (280.0, 0.9, 0.0), (300.0, 1.2, -6e-5),
                                      (175.0, 1.1, 3e-5), (290.0, 0.6, 0.0)]]
    return [Scan(P=P, v=s, x_start=ph["x_start"], y_c=yc + dy) for P, s, dy in v]


def held_in_scans(w):
    """Six training scans around the workload's track (power, speed, hatch offset)."""
    from .surrogate import Scan
    g = w["grid"]
    yc = 0.5 * g["ny"] * g["h"]
    x0 = w["physics"]["x_start"]
    return [Scan(P=P, v=v, x_start=x0, y_c=yc + dy) for P, v, dy in
            [(150.0, 0.6, 0.0), (180.0, 1.3, 2e-5), (230.0, 0.8, -2e-5), (270.0, 1.1, 0.0),
             (300.0, 1.5, 1e-5), (250.0, 0.5, -1e-5)]]


def domain(w, orient_capacity=1 << 20):
    """A device domain loaded with workload `w` (sites drawn by the library)."""
    d = remelt.RemeltDomain(w["desc"], orient_capacity)
    st = w["state"]
    d.push_orientations(st["euler"])
    d.set_fe(w["fe"])
    d.set_state(st)
    d.draw_sites(w["nuc_seed"], w["surface"], w["bulk"])
    return d
