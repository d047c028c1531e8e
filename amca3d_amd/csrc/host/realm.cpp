#include "realm.h"

#include "final_grains.h"
#include "micro_io.h"
#include "vtr_series.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <unordered_map>

namespace amca_host {

using namespace yaml;

namespace {

void load_nucleation(const Node& realm, RealmConfig& c) {
  // CellularAutomataManager_3D.cpp:193-236
  const Node* rules = realm.FindValue("nucleation_rules");
  if (!rules) return;
  for (size_t r = 0; r < rules->size(); r++) {
    const Node& rule = (*rules)[r];
    if (const Node* s = rule.FindValue("surface")) {
      std::string type = "Gaussian";
      get_required(*s, "type", type);
      get_required(*s, "site_density", c.surface.density);
      get_required(*s, "mean", c.surface.mean);
      get_required(*s, "standard_deviation", c.surface.sd);
      // extension: which faces carry surface sites (default: the reference's six)
      std::string faces = "all";
      get_if_present(*s, "faces", faces, faces);
      if (faces == "all") c.surface_faces = SurfaceFaces::kReferenceSix;
      else if (faces == "z_min") c.surface_faces = SurfaceFaces::kZMin;
      else throw std::runtime_error("parser error: nucleation_rules-surface-faces must be all or z_min");
    }
    if (const Node* b = rule.FindValue("bulk")) {
      std::string type = "Gaussian";
      get_if_present(*b, "type", type, type);
      get_required(*b, "site_density", c.bulk.density);
      get_required(*b, "mean", c.bulk.mean);
      get_required(*b, "standard_deviation", c.bulk.sd);
    }
  }
}

void load_solution_options(const Node& realm, RealmConfig& c) {
  // CellularAutomataManager_3D.cpp:248-341 and, for remelting realms,
  // CellularAutomataManager_3DRemelting.cpp:56-75
  const Node* so = realm.FindValue("solution_options");
  if (!so) return;
  std::string name;
  get_required(*so, "name", name);
  const Node* opts = expect_sequence(*so, "options", true);
  if (!opts) return;
  for (size_t i = 0; i < opts->size(); i++) {
    const Node& o = (*opts)[i];
    if (expect_map(o, "output_microstructure_information", true)) {
      const Node& f = *o.FindValue("output_microstructure_information");
      get_required(f, "file_name", c.micro_file);
      c.output_micro = true;
    } else if (expect_map(o, "random_type", true)) {
      const Node& p = *o.FindValue("random_type");
      get_if_present(p, "random_microstructure_information", c.micro_random, c.micro_random);
      // shuffled output ids (CellularAutomataManager_3D.cpp:1967-1983) take
      // std::random_shuffle draws from the rand() stream that the sites and
      // nucleus orientations use; output labels are not produced here
      bool shuffle = false;
      get_if_present(p, "shuffer_vtr_grain_ID", shuffle, shuffle);
      if (shuffle)
        throw std::runtime_error("random_type-shuffer_vtr_grain_ID: yes is not supported");
      // the finalGrains.vtk ids shuffle the same way (:277, writeToFile :1424-1431)
      get_if_present(p, "shuffer_vtk_grain_ID", shuffle, shuffle);
      if (shuffle)
        throw std::runtime_error("random_type-shuffer_vtk_grain_ID: yes is not supported");
    } else if (expect_map(o, "phase_state", true)) {
      const Node& p = *o.FindValue("phase_state");
      get_if_present(p, "has_been_melted", c.initial_melted, c.initial_melted);
    } else if (expect_map(o, "cell_grain_length_cutoff_factor", true)) {
      const Node& p = *o.FindValue("cell_grain_length_cutoff_factor");
      get_if_present(p, "factor_value", c.cutoff_factor, c.cutoff_factor);
    } else if (expect_map(o, "parallel_specification", true)) {
      // one device-resident grid: the decomposition axis has no effect
    } else if (expect_map(o, "envelope", true)) {
      const Node& p = *o.FindValue("envelope");
      double f = -1;
      get_if_present(p, "max_half_diagonal_length", f, f);
      if (f > 0) c.trunc = (double)sqrtf(2) * f * 0.5;  // Octahedron.cpp:316-320
    } else if (expect_map(o, "criterion_to_activate_grain", true)) {
      const Node& p = *o.FindValue("criterion_to_activate_grain");
      get_required(p, "using_nucleation_seed", c.using_nucleation_seed);
    } else if (expect_map(o, "load_microstructure_information", true)) {
      if (c.micro_random) continue;
      const Node& p = *o.FindValue("load_microstructure_information");
      std::string f;
      get_required(p, "file_name", f);
      if (f == "random") {
        c.micro_random = true;
        c.micro_load.clear();
      } else {
        c.micro_load = f;
        get_if_present(p, "load_seeds", c.load_seeds, c.load_seeds);
      }
    }
  }
  // neither a file nor random: random (CellularAutomataManager_3D.cpp:343-357)
  if (c.micro_load.empty()) c.micro_random = true;
}

void load_physics(const Node& realm, RealmConfig& c) {
  // ProblemPhysics_RappazGandin1993::load (ProblemPhysics.cpp:57-102)
  const Node* p = realm.FindValue("problem_physics");
  if (!p) throw std::runtime_error("parsing error: realm-problem_physics-RappazGandin");
  get_required(*p, "type", c.physics);
  if (c.physics != "RappazGandin")
    throw std::runtime_error("parsing error: realm-problem_physics-RappazGandin");
  get_if_present(*p, "initial_temperature", c.T0, c.T0);
  get_if_present(*p, "liquidus_temperature", c.Tm, c.Tm);
  get_if_present(*p, "melting_temperature", c.Tm, c.Tm);
  get_if_present(*p, "solidus_temperature", c.Ts, c.Ts);
  c.Tmin_nuc = c.Ts;
  get_if_present(*p, "minimum_temperature_for_nucleation", c.Tmin_nuc, c.Tmin_nuc);
  get_if_present(*p, "t_dot", c.Tdot, c.Tdot);
  get_if_present(*p, "a1", c.a1, c.a1);
  get_if_present(*p, "a2", c.a2, c.a2);
  get_if_present(*p, "a3", c.a3, c.a3);
  get_if_present(*p, "b1", c.b1, c.b1);
  get_if_present(*p, "b2", c.b2, c.b2);
  get_if_present(*p, "b3", c.b3, c.b3);
}

void load_output(const Node& realm, RealmConfig& c) {
  // VtrManager::load (VtrManager.cpp:60-103): the node is required
  const Node* out = realm.FindValue("output");
  if (!out) throw std::runtime_error("parser error: realm-output");
  get_if_present(*out, "output_frequency", c.output_frequency, c.output_frequency);
  get_if_present(*out, "output_data_base_name", c.output_base, c.output_base);
  get_if_present(*out, "output_time_interval", c.output_time_interval, c.output_time_interval);
  if (const Node* vars = out->FindValue("output_variables")) {
    for (size_t i = 0; i < vars->size(); i++) {
      std::string v;
      convert((*vars)[i], v);
      c.output_variables.push_back(v);
    }
  }
  if (c.output_variables.empty()) c.output_variables.push_back("grain_ID");
}

}  // namespace

RealmConfig load_ca_realm(const Node& realm) {
  RealmConfig c;
  c.trunc = (double)sqrtf(3.0);  // Octahedron.cpp:25
  get_if_present(realm, "name", c.name, c.name);
  get_required(realm, "type", c.type);
  get_if_present(realm, "dimension", c.dimension, c.dimension);
  if (c.type != "cellular_automata" && c.type != "cellular_automata_remelting")
    throw std::runtime_error("parser error: realm-type");
  if (c.dimension != 3) throw std::runtime_error("only 3-D realms are supported");
  get_if_present(realm, "random_seed", c.random_seed, c.random_seed);

  // CellularAutomataManager_3D.cpp:79-190
  const Node* domain = realm.FindValue("domain");
  if (!domain) throw std::runtime_error("parser error: realm-domain");
  // domain-load_microstructure_information (:84-131) takes the grid from the
  // file header, whose counts include the ghost layer, so the loaded file
  // never fits the grid it builds; the solution option form is the one used
  std::string domain_file;
  get_if_present(*domain, "load_microstructure_information", domain_file, domain_file);
  if (!domain_file.empty())
    throw std::runtime_error(
        "domain-load_microstructure_information is not supported; use "
        "solution_options-load_microstructure_information-file_name");
  get_required3(*domain, "original_point", c.origin);
  get_if_present_no_default3(*domain, "lateral_sizes", c.lateral);
  const Node* disc = realm.FindValue("discretization");
  if (!disc) throw std::runtime_error("parser error: realm-discretization");
  get_required(*disc, "cell_size", c.h);
  if (c.lateral[0] < 0) {
    get_required3(*disc, "number_cells", c.ncell);
  } else {
    for (int i = 0; i < 3; i++) {
      c.ncell[i] = c.lateral[i] / c.h;  // truncation, as the reference
      c.lateral[i] = c.ncell[i] * c.h;
    }
  }
  if (c.ncell[0] < 1 || c.ncell[1] < 1 || c.ncell[2] < 1)
    throw std::runtime_error("parser error: realm-discretization gives an empty grid");
  load_nucleation(realm, c);
  load_solution_options(realm, c);
  load_output(realm, c);
  load_physics(realm, c);
  return c;
}

SimulationConfig load_simulation(const Node& doc) {
  SimulationConfig s;
  // Simulation::load (Simulation.cpp:36-74)
  const Node* solvers = doc.FindValue("solvers");
  if (!solvers) throw std::runtime_error("parser error: at least one solver should be specified!");
  for (size_t i = 0; i < solvers->size(); i++) {
    std::string name;
    convert((*solvers)[i], name);
    if (name == "cellular_automata") s.cellular_automata = true;
    if (name == "finite_element_method") s.finite_element_method = true;
  }
  // CellularAutomataSolver::load (CellularAutomataSolver.cpp:40-130)
  const Node* realms = doc.FindValue("realms");
  if (!realms) throw std::runtime_error("parser error: realms");
  for (size_t i = 0; i < realms->size(); i++) {
    const Node& r = (*realms)[i];
    std::string type;
    get_required(r, "type", type);
    if (type == "finite_element") {
      // FiniteElementManager::load (FiniteElementManager.cpp:66-135)
      FeRealmConfig f;
      get_if_present(r, "name", f.name, f.name);
      if (const Node* so = r.FindValue("solution_options")) {
        if (const Node* opts = expect_sequence(*so, "options", true)) {
          for (size_t q = 0; q < opts->size(); q++) {
            const Node& o = (*opts)[q];
            if (!expect_map(o, "load_data_from_file", true)) continue;
            const Node& l = *o.FindValue("load_data_from_file");
            get_if_present(l, "for_whole_time", f.whole_time, f.whole_time);
            get_if_present(l, "lines_for_title", f.title_lines, f.title_lines);
            get_if_present(l, "lines_for_subtitle", f.subtitle_lines, f.subtitle_lines);
            get_if_present(l, "length_scale", f.length_scale, f.length_scale);
            get_if_present(l, "time_scale", f.time_scale, f.time_scale);
            const char* keys[3] = {"x_offset", "y_offset", "z_offset"};
            for (int a = 0; a < 3; a++) get_if_present(l, keys[a], f.offset[a], f.offset[a]);
          }
        }
      }
      get_required(r, "mesh", f.mesh);
      s.fe_realms.push_back(f);
      continue;
    }
    s.realms.push_back(load_ca_realm(r));
  }
  // MapVoxelManager::load (MapVoxelManager.cpp:60-90)
  if (const Node* tr = doc.FindValue("transfers")) {
    for (size_t i = 0; i < tr->size(); i++) {
      const Node& t = (*tr)[i];
      if (const Node* b = t.FindValue("bin_size")) {
        for (int a = 0; a < 3 && a < (int)b->size(); a++) convert((*b)[a], s.transfer.bin[a]);
      }
      get_if_present(t, "tolerance", s.transfer.tolerance, s.transfer.tolerance);
      get_if_present(t, "max_iterations", s.transfer.max_iterations, s.transfer.max_iterations);
      get_if_present(t, "void_temperature", s.transfer.void_temperature, s.transfer.void_temperature);
    }
  }
  const Node* tis = doc.FindValue("time_integrators");
  if (!tis) throw std::runtime_error("parser error: tiem_integrations");
  for (size_t i = 0; i < tis->size(); i++) {
    const Node* st = (*tis)[i].FindValue("standard_time_integrator");
    if (!st) continue;
    get_required(*st, "termination_time", s.ti.termination_time);
    get_required(*st, "time_step", s.ti.time_step);
    get_if_present(*st, "termination_step_count", s.ti.termination_step_count,
                   s.ti.termination_step_count);
    get_if_present(*st, "time_step_factor_ca", s.ti.time_step_factor_ca, s.ti.time_step_factor_ca);
    get_if_present(*st, "time_step_factor", s.ti.time_step_factor, s.ti.time_step_factor);
  }
  return s;
}

// ---------------------------------------------------------------------------

PlainRealm::PlainRealm(const RealmConfig& cfg, const TimeIntegratorConfig& ti)
    : cfg_(cfg), ti_(ti) {
  grid_ = Grid{cfg.ncell[0], cfg.ncell[1], cfg.ncell[2]};
}

PlainRealm::~PlainRealm() {
  if (dom_) amca_domain_destroy(dom_);
}

void PlainRealm::prepare() {
  if (prepared_) return;
  // srand(rank*10000) in the constructor (CellularAutomataManager_3D.cpp:40)
  GlibcRand rng(cfg_.random_seed);
  std::unordered_map<int, double> cell_dt;
  std::vector<int> sites;
  // surface first, then bulk (:371-373)
  std::vector<int> surf;
  surface_cells(grid_, cfg_.surface_faces, surf);
  const double h = cfg_.h;
  SiteDraw ds = create_nucleation_sites(
      (int)surf.size(), [&](int i) { return surf[i]; }, h * h, cfg_.surface.density,
      cfg_.surface.mean, cfg_.surface.sd, rng, cell_dt, sites);
  std::vector<int>().swap(surf);
  const long long nb = (long long)grid_.nx * grid_.ny * grid_.nz;
  SiteDraw db = create_nucleation_sites((int)nb, BulkCellAt{grid_}, h * h * h, cfg_.bulk.density,
                                        cfg_.bulk.mean, cfg_.bulk.sd, rng, cell_dt, sites);
  site_draws_ = (long long)ds.n_draws + db.n_draws;
  n_surface_ = ds.n_new;
  n_bulk_ = db.n_new;
  site_cells_ = sites;
  site_dt_.resize(sites.size());
  for (size_t i = 0; i < sites.size(); i++) site_dt_[i] = cell_dt[sites[i]];
  // One orientation row per possible nucleation event: in the plain path every
  // site nucleates at most once, and rows are consumed in event order from the
  // same rand() stream that drew the sites.
  const int nrows = (int)sites.size();
  euler_.resize(3 * (size_t)nrows);
  l2g_.resize(9 * (size_t)nrows);
  for (int o = 0; o < nrows; o++) {
    random_orientation(rng, &euler_[3 * (size_t)o]);
    rotation_from_euler(&euler_[3 * (size_t)o], &l2g_[9 * (size_t)o]);
  }
  prepared_ = true;
}

void PlainRealm::initialize() {
  prepare();
  const int nrows = (int)site_cells_.size();
  const double h = cfg_.h;
  amca_domain_desc d{};
  d.nx = grid_.nx;
  d.ny = grid_.ny;
  d.nz = grid_.nz;
  d.h = h;
  d.x0 = cfg_.origin[0];
  d.y0 = cfg_.origin[1];
  d.z0 = cfg_.origin[2];
  d.T0 = cfg_.T0;
  d.Tm = cfg_.Tm;
  d.Tdot = cfg_.Tdot;
  d.a1 = cfg_.a1;
  d.a2 = cfg_.a2;
  d.a3 = cfg_.a3;
  d.trunc = cfg_.trunc;
  d.time_step_factor = ti_.time_step_factor_ca;
  d.max_dt = ti_.time_step;
  d.final_time = ti_.termination_time;
  if (amca_domain_create(&d, nrows, &dom_) != AMCA_OK)
    throw std::runtime_error(std::string("amca_domain_create: ") + amca_last_error());
  if (amca_domain_set_sites(dom_, nrows, site_cells_.data(), site_dt_.data()) != AMCA_OK)
    throw std::runtime_error(std::string("amca_domain_set_sites: ") + amca_last_error());
  if (amca_domain_push_orientations(dom_, nrows, euler_.data(), l2g_.data()) != AMCA_OK)
    throw std::runtime_error(std::string("amca_domain_push_orientations: ") + amca_last_error());
}

amca_counters PlainRealm::run(long long max_steps) {
  amca_counters c{};
  auto steps = [&](long long n) {
    if (amca_domain_run(dom_, n, &c) != AMCA_OK)
      throw std::runtime_error(std::string("amca_domain_run: ") + amca_last_error());
  };
  if (!vtr_) {
    steps(max_steps);
    return c;
  }
  // VtrManager::execute after every step (CellularAutomataManager_3D.cpp:661):
  // run to the next step that can take a frame, then ask
  if (amca_domain_counters(dom_, &c) != AMCA_OK)
    throw std::runtime_error(std::string("amca_domain_counters: ") + amca_last_error());
  long long taken = 0;
  while (!c.done && (max_steps < 0 || taken < max_steps)) {
    long long chunk = vtr_->steps_to_next(c.step);
    if (chunk == 0) chunk = -1;  // no frame can come during this run
    if (max_steps >= 0) chunk = chunk < 0 ? max_steps - taken : std::min(chunk, max_steps - taken);
    const long long s0 = c.step;
    steps(chunk);
    taken += c.step - s0;
    if (c.step == s0) break;
    if (vtr_->due(c.step, c.time)) vtr_frame(c.time);
  }
  return c;
}

void PlainRealm::write_microstructure(const std::string& base_dir) {
  if (!dom_ || !cfg_.output_micro) return;
  auto fail = [](const char* what) {
    throw std::runtime_error(std::string(what) + ": " + amca_last_error());
  };
  amca_counters c{};
  if (amca_domain_counters(dom_, &c) != AMCA_OK) fail("amca_domain_counters");
  const long long NG = grid_.NG();
  MicroState m;
  m.num_cells = NG;
  // the sites drawn by this realm; a plain realm loads no file
  m.num_global_expected = n_surface_ + n_bulk_;
  m.x0 = cfg_.origin[0];
  m.y0 = cfg_.origin[1];
  m.z0 = cfg_.origin[2];
  m.h = cfg_.h;
  m.nxG = grid_.nxG();
  m.nyG = grid_.nyG();
  m.nzG = grid_.nzG();
  m.gid.resize((size_t)NG);
  m.ori.resize((size_t)NG);
  m.melted.assign((size_t)NG, 0);  // hasBeenMelted_ is empty without remelting
  const int cap = (int)site_cells_.size();
  std::vector<int32_t> sc((size_t)std::max(cap, 1));
  std::vector<double> sd((size_t)std::max(cap, 1));
  int32_t ns = 0;
  if (amca_domain_download(dom_, m.ori.data(), m.gid.data(), nullptr, cap, sc.data(), sd.data(),
                           &ns) != AMCA_OK)
    fail("amca_domain_download");
  // cellNucleationDT_: -1 off the sites; a site the nucleation pass dropped
  // (nucleated or captured) reads 0 (CellularAutomataManager.cpp:286-292); the
  // sites still listed keep their undercooling
  m.nuc_dt.assign((size_t)NG, -1.0);
  for (int c0 : site_cells_) m.nuc_dt[(size_t)c0] = 0.0;
  for (int q = 0; q < std::min(ns, cap); q++) m.nuc_dt[(size_t)sc[q]] = sd[q];
  const int no = c.n_orient;
  m.orient_id.resize((size_t)no);
  m.euler.assign(euler_.begin(), euler_.begin() + 3 * (size_t)no);
  for (int o = 0; o < no; o++) m.orient_id[(size_t)o] = o;
  const int ng = c.n_grains;
  std::vector<double> env(4 * (size_t)std::max(ng, 1));
  m.grain_id.resize((size_t)ng);
  m.grain_cell.resize((size_t)ng);
  m.grain_orient.resize((size_t)ng);
  int32_t got = 0;
  if (amca_domain_download_grains(dom_, ng, env.data(), m.grain_cell.data(),
                                  m.grain_orient.data(), &got) != AMCA_OK)
    fail("amca_domain_download_grains");
  m.grain_centre.resize(3 * (size_t)ng);
  for (int g = 0; g < ng; g++) {
    m.grain_id[(size_t)g] = g;
    for (int a = 0; a < 3; a++) m.grain_centre[3 * (size_t)g + a] = env[4 * (size_t)g + a];
  }
  const std::string name = rank_file_name(cfg_.micro_file);
  write_micro((base_dir.empty() || name[0] == '/') ? name : base_dir + "/" + name, m);
}

void PlainRealm::enable_vtr(const std::string& base_dir, bool on) {
  vtr_.reset();
  if (on) vtr_ = std::make_unique<VtrSeries>(cfg_, base_dir, false);
}

void PlainRealm::vtr_frame(double time) {
  std::vector<int32_t> ori((size_t)grid_.NG());
  if (amca_domain_download(dom_, ori.data(), nullptr, nullptr, 0, nullptr, nullptr, nullptr) !=
      AMCA_OK)
    throw std::runtime_error(std::string("amca_domain_download: ") + amca_last_error());
  VtrFields f;
  f.orient = ori.data();
  f.euler3 = euler_.data();
  f.n_orient = (int32_t)(euler_.size() / 3);
  vtr_->write(time, f);
}

void PlainRealm::vtr_final_frame() {
  if (!vtr_ || !dom_) return;
  amca_counters c{};
  if (amca_domain_counters(dom_, &c) != AMCA_OK)
    throw std::runtime_error(std::string("amca_domain_counters: ") + amca_last_error());
  if (vtr_->due(0, c.time)) vtr_frame(c.time);  // execute(0, time_) (:1990-1993)
}

void PlainRealm::write_final_grains(const std::string& vtk, const std::string& result) {
  if (!dom_) return;
  std::vector<int32_t> ori((size_t)grid_.NG());
  if (amca_domain_download(dom_, ori.data(), nullptr, nullptr, 0, nullptr, nullptr, nullptr) !=
      AMCA_OK)
    throw std::runtime_error(std::string("amca_domain_download: ") + amca_last_error());
  // a realm without remelting writes every melted flag as 1 (hasBeenMelted_ empty)
  amca_host::write_final_grains(vtk, result, grid_.nx, grid_.ny, grid_.nz, cfg_.origin[0],
                                cfg_.origin[1], cfg_.origin[2], cfg_.h, ori.data(), nullptr,
                                euler_.data(), (int32_t)(euler_.size() / 3));
}

}  // namespace amca_host
