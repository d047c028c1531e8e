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
                nuc_seed=nuc_seed, bulk=bulk, surface=surface)


def domain(w, orient_capacity=1 << 20):
    """A device domain loaded with workload `w` (sites drawn by the library)."""
    d = remelt.RemeltDomain(w["desc"], orient_capacity)
    st = w["state"]
    d.push_orientations(st["euler"])
    d.set_fe(w["fe"])
    d.set_state(st)
    d.draw_sites(w["nuc_seed"], w["surface"], w["bulk"])
    return d
