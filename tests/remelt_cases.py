"""Input generators for the coupled CA-FE (remelting) path: a structured FE hex
mesh over the CA box, Rosenthal moving-point-source frames on its nodes and a
Voronoi substrate in the reference's loaded-microstructure layout.  Used by the
tests to feed the same inputs to the HIP path and to the oracle."""
import numpy as np


def fe_mesh(grid, spacing_cells, pad=1):
    """Nodes every `spacing_cells` CA cells covering the ghosted CA box."""
    h = grid["h"]
    s = spacing_cells * h
    lo = np.array([grid["x0"], grid["y0"], grid["z0"]]) - pad * h - 1e-3 * h
    n = [grid["nx"], grid["ny"], grid["nz"]]
    cnt = [int(np.ceil((ni + 2 * pad) * h / s)) + 1 for ni in n]
    ax = [lo[q] + s * np.arange(cnt[q]) for q in range(3)]
    zz, yy, xx = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    xyz = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=1)
    return dict(shape=tuple(cnt), xyz=xyz, axes=ax)


def rosenthal(xyz, t, *, P=200.0, eta=0.35, v=1.0, k=20.0, alpha=5e-6, T0=300.0, x_start,
              y_c, z_top, r_min):
    xi = xyz[:, 0] - (x_start + v * t)
    R = np.sqrt(xi * xi + (xyz[:, 1] - y_c) ** 2 + (xyz[:, 2] - z_top) ** 2)
    R = np.maximum(R, r_min)
    return T0 + eta * P / (2 * np.pi * k * R) * np.exp(-v * (R + xi) / (2 * alpha))


def voronoi_state(grid, n_seeds, seed):
    """Every interior cell solid with its own grain (id in cell order); the
    orientation of a cell is the index of its nearest Voronoi seed."""
    rng = np.random.default_rng(seed)
    nx, ny, nz, h = grid["nx"], grid["ny"], grid["nz"], grid["h"]
    lo = np.array([grid["x0"], grid["y0"], grid["z0"]])
    pts = lo + rng.uniform(0, 1, size=(n_seeds, 3)) * np.array([nx, ny, nz]) * h
    U, V, W = rng.uniform(0, 1, size=(3, n_seeds))
    euler = np.stack([U * 2 * np.pi, np.arccos(2 * V - 1), W * 0.5 * np.pi], axis=1)
    nxG, nyG, nzG = nx + 2, ny + 2, nz + 2
    ng = nxG * nyG * nzG
    k, j, i = np.meshgrid(np.arange(1, nz + 1), np.arange(1, ny + 1), np.arange(1, nx + 1),
                          indexing="ij")
    cells = ((k * nyG + j) * nxG + i).ravel()
    cen = np.stack([grid["x0"] + h * (i.ravel() - 0.5), grid["y0"] + h * (j.ravel() - 0.5),
                    grid["z0"] + h * (k.ravel() - 0.5)], axis=1)
    lab = np.empty(len(cells), np.int32)
    for a in range(0, len(cells), 65536):
        d = ((cen[a:a + 65536, None, :] - pts[None, :, :]) ** 2).sum(-1)
        lab[a:a + 65536] = d.argmin(1)
    ori = np.full(ng, -1, np.int32)
    gid = np.full(ng, -1, np.int32)
    ori[cells] = lab
    gid[cells] = np.arange(len(cells), dtype=np.int32)
    return dict(ori=ori, gid=gid, melted=np.zeros(ng, np.uint8), euler=euler,
                grain_cell=cells.astype(np.int32), grain_orient=lab.copy())


def track_case(nx=48, ny=20, nz=14, h=2.5e-6, spacing=4, n_frames=12, frame_dt=10e-6,
               P=200.0, v=1.0, n_seeds=40, seed=3, nucleation=None, final_time=1.0,
               max_steps=-1, source_x0=None, z_top=None):
    """A single melt-and-resolidify track over a Voronoi substrate."""
    grid = dict(nx=nx, ny=ny, nz=nz, h=h, x0=0.0, y0=0.0, z0=0.0)
    fe = fe_mesh(grid, spacing)
    if source_x0 is None:
        source_x0 = 0.15 * nx * h
    if z_top is None:
        z_top = nz * h
    times = frame_dt * np.arange(n_frames)
    theta = np.stack([rosenthal(fe["xyz"], t, P=P, v=v, x_start=source_x0, y_c=0.5 * ny * h,
                                z_top=z_top, r_min=h) for t in times])
    fe.update(times=times, theta=theta)
    return dict(grid=grid, fe=fe, bin=(5, 5, 5), tol=1e-16, max_iter=20, void_temp=200.0,
                physics=dict(Tm=1600.0, Ts=1550.0, Tmin_nuc=1550.0, a1=-0.544e-4, a2=2.03e-4,
                             a3=0.0, b1=1.0, b2=2.0, b3=3.0),
                trunc=float(np.sqrt(np.float32(3.0))), tsf=0.2, cutoff_factor=4.0,
                final_time=final_time, max_steps=max_steps, initial_melted=0,
                nucleation=nucleation, seed=seed, state=voronoi_state(grid, n_seeds, seed))


def corner_mesh(grid):
    """FE nodes on the CA cell corners (ghost cells included): every cell is one
    element and its centre sits at the element centre."""
    h = grid["h"]
    cnt = [grid["nx"] + 3, grid["ny"] + 3, grid["nz"] + 3]
    ax = [grid[o] + h * (np.arange(cnt[q]) - 1.0) for q, o in enumerate(("x0", "y0", "z0"))]
    zz, yy, xx = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return dict(shape=tuple(cnt), xyz=np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=1), axes=ax)


def single_cell_case(nx=7, ny=6, nz=5, hot=(4, 3, 3), fresh=False, h=1e-6, n_frames=6,
                     frame_dt=1e-4, undercooling=5.0):
    """One cell melted by a hot corner pattern (its 8 corner nodes at Tm+5, all
    others at Tm-10: the cell averages Tm+5, every neighbour at most Tm-2.5),
    then a uniform field at Tm-undercooling.  Every cell carries its own
    orientation row with zero Euler angles, so the six face neighbours reach the
    molten cell in the same step and the x-major sweep order picks the winner."""
    grid = dict(nx=nx, ny=ny, nz=nz, h=h, x0=0.0, y0=0.0, z0=0.0)
    fe = corner_mesh(grid)
    fnx, fny, fnz = fe["shape"]
    Tm = 1600.0
    hotf = np.full((fnz, fny, fnx), Tm - 10.0)
    i, j, k = hot
    hotf[k:k + 2, j:j + 2, i:i + 2] = Tm + 5.0
    cold = np.full(fnx * fny * fnz, Tm - undercooling)
    theta = np.stack([hotf.ravel(), cold] + [cold] * (n_frames - 2))
    fe.update(times=frame_dt * np.arange(n_frames), theta=theta)
    nxG, nyG, nzG = nx + 2, ny + 2, nz + 2
    ng = nxG * nyG * nzG
    kk, jj, ii = np.meshgrid(np.arange(1, nz + 1), np.arange(1, ny + 1), np.arange(1, nx + 1),
                             indexing="ij")
    cells = ((kk * nyG + jj) * nxG + ii).ravel().astype(np.int32)
    n = len(cells)
    ori = np.full(ng, -1, np.int32)
    gid = np.full(ng, -1, np.int32)
    ori[cells] = np.arange(n, dtype=np.int32)
    gid[cells] = np.arange(n, dtype=np.int32)
    hot_cell = (k * nyG + j) * nxG + i
    if fresh:
        gid[hot_cell] = -1
    state = dict(ori=ori, gid=gid, melted=np.zeros(ng, np.uint8), euler=np.zeros((n, 3)),
                 grain_cell=cells, grain_orient=np.arange(n, dtype=np.int32))
    return dict(grid=grid, fe=fe, bin=(3, 3, 3), tol=1e-16, max_iter=20, void_temp=200.0,
                physics=dict(Tm=Tm, Ts=1550.0, Tmin_nuc=1550.0, a1=-0.544e-4, a2=2.03e-4,
                             a3=0.0, b1=1.0, b2=2.0, b3=3.0),
                trunc=float(np.sqrt(np.float32(3.0))), tsf=0.2, cutoff_factor=4.0,
                final_time=1.0, max_steps=-1, initial_melted=0, nucleation=None, seed=0,
                state=state, hot_cell=hot_cell)
