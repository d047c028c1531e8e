"""Input decks (YAML text) for the plain CA realm.

The keys are those of the reference's realm API (CellularAutomataManager_3D::load,
ProblemPhysics_RappazGandin1993::load, VtrManager::load,
CellularAutomataSolver::load).  Two keys are extensions with reference
defaults: ``random_seed`` (srand seed, 0 = the reference's rank-0 value) and the
surface rule's ``faces`` (``all`` = the reference's six faces, ``z_min``).
"""


def _fmt(x):
    return repr(float(x)) if isinstance(x, float) else str(x)


def plain_deck(nx, ny, nz, h, *, origin=(0.0, 0.0, 0.0), lateral=None,
               surface=(0.0, 2.0, 0.5), surface_faces="all", bulk=(0.0, 3.0, 1.0),
               T0=-0.2, Tm=0.0, Tdot=-20.0, a=(-0.544e-4, 2.03e-4, 0.0),
               termination_time=17.81, time_step=1e-3, time_step_factor_ca=None,
               max_half_diagonal_length=None, random_seed=None):
    """A `cellular_automata` deck; surface/bulk are (density, mean, sd)."""
    lines = [
        "simulations:",
        "  - name: test",
        "    time_integrator: ti_1",
        "",
        "solvers:",
        "   - cellular_automata",
        "",
        "realms:",
        "  - name: realm1",
        "    type: cellular_automata",
        "    dimension: 3",
    ]
    if random_seed is not None:
        lines.append(f"    random_seed: {int(random_seed)}")
    lines += ["", "    domain:", "      type: cubic",
              f"      original_point: [{', '.join(_fmt(float(v)) for v in origin)}]"]
    if lateral is not None:
        lines.append(f"      lateral_sizes: [{', '.join(_fmt(float(v)) for v in lateral)}]")
    lines += ["", "    discretization:", f"      cell_size: {_fmt(float(h))}",
              f"      number_cells: [{nx}, {ny}, {nz}]", "",
              "    nucleation_rules:",
              "      - surface:",
              "          type: Gaussian",
              f"          site_density: {_fmt(float(surface[0]))}",
              f"          mean: {_fmt(float(surface[1]))}",
              f"          standard_deviation: {_fmt(float(surface[2]))}"]
    if surface_faces != "all":
        lines.append(f"          faces: {surface_faces}")
    lines += ["",
              "      - bulk:",
              "          type: Gaussian",
              f"          site_density: {_fmt(float(bulk[0]))}",
              f"          mean: {_fmt(float(bulk[1]))}",
              f"          standard_deviation: {_fmt(float(bulk[2]))}", "",
              "    problem_physics:",
              "        type: RappazGandin",
              f"        initial_temperature: {_fmt(float(T0))}",
              f"        melting_temperature: {_fmt(float(Tm))}",
              f"        t_dot: {_fmt(float(Tdot))}",
              f"        a1: {_fmt(float(a[0]))}",
              f"        a2: {_fmt(float(a[1]))}",
              f"        a3: {_fmt(float(a[2]))}", "",
              "    solution_options:",
              "       name: my_options",
              "       options:",
              "         - random_type:",
              "             rejection_sampling: no",
              "         - cell_grain_length_cutoff_factor:",
              "             factor_value: 4.0"]
    if max_half_diagonal_length is not None:
        lines += ["         - envelope:",
                  f"             max_half_diagonal_length: {_fmt(float(max_half_diagonal_length))}"]
    lines += ["", "    output:",
              "      output_data_base_name: ./Results/Grains",
              "      output_frequency: 100",
              "      output_variables:",
              "        - grain_ID", "",
              "time_integrators:",
              "  - standard_time_integrator:",
              "      name: ti_1",
              f"      termination_time: {_fmt(float(termination_time))}",
              "      termination_step_count: 10000",
              f"      time_step: {_fmt(float(time_step))}"]
    if time_step_factor_ca is not None:
        lines.append(f"      time_step_factor_ca: {_fmt(float(time_step_factor_ca))}")
    lines += ["      realms:", "        - realm1", ""]
    return "\n".join(lines)


def nucleation_example_deck():
    """The reference's nucleation example (example/nucleation/Inputfile.i):
    0.03 m cube of 5e-4 m cells (lateral_sizes wins over number_cells -> 60^3),
    bulk density written as 10e6, Gaussian N(3, 1), T0 = -0.2, Tdot = -20."""
    text = plain_deck(800, 720, 252, 0.0005, lateral=(0.03, 0.03, 0.03),
                      surface=(0.0, 2.0, 0.5), bulk=(1e7, 3.0, 1.0),
                      termination_time=17.81, time_step=0.001)
    # keep the example's literal spelling of the bulk density
    return text.replace("site_density: 10000000.0", "site_density: 10e6")


def uniform_cooling_512_deck(n=512, faces="z_min", random_seed=None):
    """512^3 cells of 2e-6 m, Tdot = -200 K/s from T0 = -0.2 K; bulk Gaussian
    nucleation 5e13 m^-3 N(3, 1); surface 1e9 m^-2 N(1, 0.5) on the z = 0 face."""
    return plain_deck(n, n, n, 2e-6, surface=(1e9, 1.0, 0.5), surface_faces=faces,
                      bulk=(5e13, 3.0, 1.0), Tdot=-200.0, termination_time=1.0,
                      time_step=1e-4, random_seed=random_seed)
