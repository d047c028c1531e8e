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


def write_frames_text(path, fe, *, unit=1e-3, offset=(0.5, -0.25, 0.0), time_unit=1e-3,
                      t0=5.0, title_lines=4):
    """Frames in the reference's text format (README, "PBF_x"): coordinates in
    `unit` with `offset` removed, times in `time_unit` starting at t0.  Returns
    the node coordinates and relative frame times a reader must recover,
    computed the way the reader does ((c + offset) * unit, (t - t_first) * scale)."""
    fnx, fny, fnz = fe["shape"]
    xyz = np.asarray(fe["xyz"], np.float64)
    raw = xyz / unit - np.asarray(offset)
    t_raw = t0 + np.asarray(fe["times"], np.float64) / time_unit
    theta = np.asarray(fe["theta"], np.float64)
    lines = ["Title:", "frames for a test", "2026/12/02", "End of title"][:title_lines]
    for f in range(len(t_raw)):
        lines += ["" if f else "Subtitle (the second line)", "frames of a structured mesh",
                  "No. step   No. step    time   time   x0    x1   y0   y1    z0   z1",
                  f"{f + 1} {f + 1} {float(t_raw[f])!r} {float(t_raw[f])!r} 0 {fnx - 1} 0 {fny - 1} 0 {fnz - 1}",
                  "x  y  z  tn"]
        lines += [f"{float(a)!r} {float(b)!r} {float(c)!r} {float(v)!r}" for (a, b, c), v in zip(raw, theta[f])]
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    read_xyz = (raw + np.asarray(offset)) * unit
    rel = np.concatenate([[0.0], (t_raw[1:] - t_raw[0]) * time_unit])
    return read_xyz, rel


def cafe_deck(case, *, mesh, unit=1e-3, offset=(0.5, -0.25, 0.0), time_unit=1e-3, micro_load=None,
              micro_out=None, seeds=True, load_seeds=False, title_lines=4):
    """A CafeSolver input deck (finite_element + cellular_automata_remelting)
    for `case` (see track_case)."""
    g, ph = case["grid"], case["physics"]
    nuc = case.get("nucleation") or dict(surface=(0.0, 1.0, 0.5), bulk=(0.0, 5.0, 1.0))
    opts = ["         - phase_state:",
            f"             has_been_melted: {'yes' if case['initial_melted'] else 'no'}",
            "         - cell_grain_length_cutoff_factor:",
            f"             factor_value: {float(case['cutoff_factor'])!r}",
            "         - criterion_to_activate_grain:",
            f"             using_nucleation_seed: {'yes' if seeds else 'no'}",
            "         - load_microstructure_information:",
            f"             file_name: {micro_load or 'random'}"]
    if micro_load and load_seeds:
        opts += ["             load_seeds: yes"]
    if micro_out:
        opts += ["         - output_microstructure_information:",
                 f"             file_name: {micro_out}"]
    return "\n".join([
        "simulations:", "  - name: test", "    time_integrator: ti_1", "",
        "solvers:", "   - cellular_automata", "   - finite_element_method", "",
        "transfers:", "   - name: fe_ca",
        f"     bin_size: [{case['bin'][0]}, {case['bin'][1]}, {case['bin'][2]}]",
        f"     tolerance: {float(case['tol'])!r}", f"     max_iterations: {case['max_iter']}",
        f"     void_temperature: {float(case['void_temp'])!r}", "",
        "realms:", "  - name: realm0", "    type: finite_element", "    dimension: 3",
        f"    mesh: {mesh}", "    solution_options:", "       name: my_options",
        "       options:", "         - load_data_from_file:", "             theta: theta",
        "             for_whole_time: yes", f"             length_scale: {float(unit)!r}",
        f"             time_scale: {float(time_unit)!r}", f"             lines_for_title: {title_lines}",
        "             lines_for_subtitle: 5", f"             x_offset: {float(offset[0])!r}",
        f"             y_offset: {float(offset[1])!r}", f"             z_offset: {float(offset[2])!r}",
        "    output:", "      output_data_base_name: ./Results/FEM", "      output_frequency: 200000",
        "",
        "  - name: realm1", "    type: cellular_automata_remelting", "    dimension: 3",
        f"    random_seed: {case.get('seed', 0)}",
        "    domain:", "      type: cubic",
        f"      original_point: [{float(g['x0'])!r}, {float(g['y0'])!r}, {float(g['z0'])!r}]",
        "    discretization:", f"      cell_size: {float(g['h'])!r}",
        f"      number_cells: [{g['nx']}, {g['ny']}, {g['nz']}]",
        "    nucleation_rules:", "      - surface:", "          type: Gaussian",
        f"          site_density: {float(nuc['surface'][0])!r}", f"          mean: {float(nuc['surface'][1])!r}",
        f"          standard_deviation: {float(nuc['surface'][2])!r}",
        "      - bulk:", "          type: Gaussian", f"          site_density: {float(nuc['bulk'][0])!r}",
        f"          mean: {float(nuc['bulk'][1])!r}", f"          standard_deviation: {float(nuc['bulk'][2])!r}",
        "    problem_physics:", "        type: RappazGandin",
        f"        melting_temperature: {float(ph['Tm'])!r}", f"        solidus_temperature: {float(ph['Ts'])!r}",
        f"        minimum_temperature_for_nucleation: {float(ph['Tmin_nuc'])!r}",
        f"        a1: {float(ph['a1'])!r}", f"        a2: {float(ph['a2'])!r}", f"        a3: {float(ph['a3'])!r}",
        "    solution_options:", "       name: my_options", "       options:"] + opts + [
        "    output:", "      output_data_base_name: ./Results/Grains", "      output_frequency: -1",
        "      output_variables:", "        - grain_orientation", "",
        "time_integrators:", "  - standard_time_integrator:", "      name: ti_1",
        f"      termination_time: {float(case['final_time'])!r}", "      termination_step_count: 10",
        "      time_step: 0.001", f"      time_step_factor_ca: {float(case['tsf'])!r}",
        "      realms:", "        - realm0", "        - realm1", ""])


def read_micro_file(path):
    """Parse a microstructure information file (the reference's text layout)."""
    with open(path) as fh:
        rows = fh.read().split("\n")
    head = rows[1].split()
    n_cells, n_or, n_gr, n_exp = (int(head[0]), int(head[1]), int(head[2]), int(head[3]))
    cells = np.array([r.split() for r in rows[3:3 + n_cells]], dtype=object)
    o0 = 4 + n_cells
    ors = [r.split() for r in rows[o0:o0 + n_or]]
    g0 = o0 + n_or + 1
    grs = [r.split() for r in rows[g0:g0 + n_gr]]
    return dict(header=head, n_expected=n_exp,
                gid=cells[:, 1].astype(np.int32), ori=cells[:, 2].astype(np.int32),
                melted=cells[:, 3].astype(np.uint8), nuc_dt=cells[:, 4].astype(np.float64),
                orient_id=np.array([int(o[0]) for o in ors], np.int32),
                euler=np.array([[float(v) for v in o[1:4]] for o in ors]),
                grain_id=np.array([int(g[0]) for g in grs], np.int32),
                grain_cell=np.array([int(g[1]) for g in grs], np.int32),
                grain_centre=np.array([[float(v) for v in g[2:5]] for g in grs]),
                grain_orient=np.array([int(g[5]) for g in grs], np.int32))


def parse_frames_text(path, title_lines=4, subtitle_lines=5, unit=1e-3, time_unit=1e-3,
                      offset=(0.5, -0.25, 0.0)):
    """Reader of the text frame format: per frame `subtitle_lines` lines whose
    fourth holds "step step time time x0 x1 y0 y1 z0 z1", then one line
    "x y z T" per node (x fastest)."""
    with open(path) as fh:
        rows = fh.read().split("\n")
    r = title_lines
    xyz, times, theta, shape = None, [], [], None
    while r + subtitle_lines <= len(rows):
        head = rows[r + 3].split()
        if len(head) < 10:
            break
        x0, x1, y0, y1, z0, z1 = (int(v) for v in head[4:10])
        shape = (x1 - x0 + 1, y1 - y0 + 1, z1 - z0 + 1)
        n = shape[0] * shape[1] * shape[2]
        body = rows[r + subtitle_lines:r + subtitle_lines + n]
        if len(body) < n or not body[-1].strip():
            break
        a = np.array([[float(v) for v in b.split()[:4]] for b in body])
        if xyz is None:
            xyz = (a[:, :3] + np.asarray(offset)) * unit
        times.append(float(head[2]))
        theta.append(a[:, 3])
        r += subtitle_lines + n
    t = np.asarray(times)
    return dict(shape=shape, xyz=xyz, times=(t - t[0]) * time_unit, theta=np.stack(theta))
