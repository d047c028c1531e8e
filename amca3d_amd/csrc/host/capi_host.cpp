// Realm-level C ABI (amca_sim_*): the Simulation / CellularAutomataSolver
// layer of the reference (Simulation.cpp:36-93, CellularAutomataSolver.cpp:40-202)
// over the device domain.
#include <cstring>
#include <filesystem>
#include <memory>
#include <stdexcept>
#include <string>

#include "../../../include/amca3d.h"
#include "cafe.h"
#include "final_grains.h"
#include "realm.h"

struct amca_sim {
  amca_host::SimulationConfig cfg;
  std::string base;  // the deck's folder (empty: the working directory)
  std::unique_ptr<amca_host::PlainRealm> plain;
  std::unique_ptr<amca_host::CafeRealm> cafe;  // finite_element + cellular_automata_remelting
  bool initialized = false;
  bool prepared = false;
};

namespace {
thread_local std::string g_host_error;

amca_status host_fail(amca_status s, const std::string& m) {
  g_host_error = m;
  return s;
}

template <class F>
amca_status guarded(F&& f) {
  try {
    return f();
  } catch (const std::exception& e) {
    return host_fail(AMCA_ERR_IO, e.what());
  }
}

// Simulation::load picks the solver from `solvers` and the realm types
// (CellularAutomataSolver.cpp:53-79, CafeSolver.cpp:54-78).
amca_status create_from_doc(const amca_host::yaml::Node& doc, const std::string& base,
                            amca_sim** out) {
  auto s = std::make_unique<amca_sim>();
  s->cfg = amca_host::load_simulation(doc);
  s->base = base;
  if (s->cfg.realms.empty()) return host_fail(AMCA_ERR_INVALID, "no cellular automata realm");
  const auto& r = s->cfg.realms[0];
  if (s->cfg.finite_element_method) {
    if (r.type != "cellular_automata_remelting")
      return host_fail(AMCA_ERR_INVALID, "a finite_element_method run needs a "
                                         "cellular_automata_remelting realm");
    s->cafe = std::make_unique<amca_host::CafeRealm>(s->cfg, base);
  } else {
    if (r.type != "cellular_automata")
      return host_fail(AMCA_ERR_INVALID, "realm type '" + r.type + "' needs the "
                                         "finite_element_method solver and an FE realm");
    s->plain = std::make_unique<amca_host::PlainRealm>(r, s->cfg.ti);
  }
  *out = s.release();
  return AMCA_OK;
}

std::string dir_of(const std::string& p) {
  const size_t k = p.find_last_of('/');
  return k == std::string::npos ? std::string() : p.substr(0, k);
}
}  // namespace

extern "C" {

const char* amca_sim_last_error(void) { return g_host_error.c_str(); }

amca_status amca_sim_create(const char* yaml_text, amca_sim** out) {
  if (!yaml_text || !out) return host_fail(AMCA_ERR_INVALID, "null argument");
  return guarded([&] { return create_from_doc(amca_host::yaml::parse(yaml_text), "", out); });
}

amca_status amca_sim_create_from_file(const char* path, amca_sim** out) {
  if (!path || !out) return host_fail(AMCA_ERR_INVALID, "null argument");
  return guarded(
      [&] { return create_from_doc(amca_host::yaml::parse_file(path), dir_of(path), out); });
}

amca_status amca_sim_initialize(amca_sim* s) {
  if (!s) return host_fail(AMCA_ERR_INVALID, "null simulation");
  return guarded([&] {
    if (s->cafe)
      s->cafe->initialize();
    else
      s->plain->initialize();
    s->initialized = true;
    s->prepared = true;
    return AMCA_OK;
  });
}

amca_status amca_sim_prepare(amca_sim* s) {
  if (!s) return host_fail(AMCA_ERR_INVALID, "null simulation");
  return guarded([&] {
    if (s->cafe)
      s->cafe->prepare();
    else
      s->plain->prepare();
    s->prepared = true;
    return AMCA_OK;
  });
}

amca_status amca_sim_run(amca_sim* s, int64_t max_steps, amca_counters* out) {
  if (!s || !s->initialized) return host_fail(AMCA_ERR_INVALID, "simulation not initialized");
  return guarded([&] {
    if (s->cafe) {
      const amca_remelt_counters r = s->cafe->run(max_steps);
      if (out) {
        std::memset(out, 0, sizeof(*out));
        out->step = r.step;
        out->time = r.time;
        out->dt = r.dt;
        out->n_grains = r.n_grains;
        out->n_orient = r.n_orient;
        out->done = r.done;
      }
      return AMCA_OK;
    }
    amca_counters c = s->plain->run(max_steps);
    if (out) *out = c;
    return AMCA_OK;
  });
}

amca_status amca_sim_finalize(amca_sim* s) {
  if (!s || !s->initialized) return host_fail(AMCA_ERR_INVALID, "simulation not initialized");
  return guarded([&] {
    if (s->cafe)
      s->cafe->finalize();
    else
      s->plain->write_microstructure(s->base);
    // the last VTR frame (output_last_step_result), then the end-of-run files
    if (s->cafe)
      s->cafe->vtr_final_frame();
    else
      s->plain->vtr_final_frame();
    // finalGrains.vtk next to the deck; finalResult.txt in the folder of the
    // output base name (VtrManager.cpp:156-169 takes everything before the last
    // '/', or the whole name when there is none)
    const auto& r = s->cfg.realms[0];
    auto join = [](const std::string& a, const std::string& b) {
      return (a.empty() || (!b.empty() && b[0] == '/')) ? b : a + "/" + b;
    };
    const size_t k = r.output_base.find_last_of('/');
    const std::string dir = join(s->base, k == std::string::npos ? r.output_base
                                                                 : r.output_base.substr(0, k));
    std::filesystem::create_directories(dir);
    const std::string vtk = join(s->base, "finalGrains.vtk"), res = dir + "/finalResult.txt";
    if (s->cafe)
      s->cafe->write_final_grains(vtk, res);
    else
      s->plain->write_final_grains(vtk, res);
    return AMCA_OK;
  });
}

amca_status amca_sim_enable_vtr(amca_sim* s, int32_t on) {
  if (!s) return host_fail(AMCA_ERR_INVALID, "null simulation");
  return guarded([&] {
    if (s->cafe)
      s->cafe->enable_vtr(s->base, on != 0);
    else
      s->plain->enable_vtr(s->base, on != 0);
    return AMCA_OK;
  });
}

amca_status amca_write_final_grains(const char* vtk_path, const char* result_path, int32_t nx,
                                    int32_t ny, int32_t nz, double x0, double y0, double z0,
                                    double h, const int32_t* orient, const int32_t* melted,
                                    const double* euler3, int32_t n_orient) {
  if (!orient || nx < 1 || ny < 1 || nz < 1 || n_orient < 0)
    return host_fail(AMCA_ERR_INVALID, "bad arguments");
  return guarded([&] {
    amca_host::write_final_grains(vtk_path ? vtk_path : "", result_path ? result_path : "", nx,
                                  ny, nz, x0, y0, z0, h, orient, melted, euler3, n_orient);
    return AMCA_OK;
  });
}

amca_domain* amca_sim_domain(amca_sim* s) { return s && s->plain ? s->plain->domain() : nullptr; }

amca_remelt* amca_sim_remelt(amca_sim* s) { return s && s->cafe ? s->cafe->domain() : nullptr; }

amca_status amca_sim_get_info(amca_sim* s, amca_sim_info* info) {
  if (!s || !info) return host_fail(AMCA_ERR_INVALID, "null argument");
  std::memset(info, 0, sizeof(*info));
  const auto& r = s->cfg.realms[0];
  info->nx = r.ncell[0];
  info->ny = r.ncell[1];
  info->nz = r.ncell[2];
  info->h = r.h;
  info->termination_time = s->cfg.ti.termination_time;
  info->time_step = s->cfg.ti.time_step;
  if (s->cafe && s->prepared) {
    info->n_frames = s->cafe->frames().frames();
    info->fe_nodes = s->cafe->frames().nodes();
    info->n_init_grains = (int32_t)s->cafe->initial_state().grain_id.size();
  }
  if (s->cafe && s->initialized) {
    int32_t ns = 0;
    amca_remelt_sites(s->cafe->domain(), 0, nullptr, nullptr, &ns);
    info->n_sites = ns;
    info->n_surface_sites = s->cafe->n_surface_new();
    info->n_bulk_sites = s->cafe->n_bulk_new();
  }
  info->coupled = s->cafe ? 1 : 0;
  if (s->plain && s->prepared) {
    info->n_sites = (int32_t)s->plain->initial_sites().size();
    info->n_surface_sites = s->plain->n_surface_sites();
    info->n_bulk_sites = s->plain->n_bulk_sites();
    info->n_site_draws = s->plain->site_draws();
  }
  return AMCA_OK;
}

amca_status amca_sim_initial_sites(amca_sim* s, int32_t cap, int32_t* cell, double* dT_crit,
                                   double* euler3, int32_t* n) {
  if (!s || !s->prepared || !n) return host_fail(AMCA_ERR_INVALID, "simulation not prepared");
  if (!s->plain) return host_fail(AMCA_ERR_INVALID, "coupled runs: use amca_remelt_sites");
  const auto& sc = s->plain->initial_sites();
  const auto& sd = s->plain->initial_site_dt();
  const auto& eu = s->plain->euler_rows();
  *n = (int32_t)sc.size();
  int m = std::min<int>(cap, (int)sc.size());
  for (int i = 0; i < m; i++) {
    if (cell) cell[i] = sc[i];
    if (dT_crit) dT_crit[i] = sd[i];
    if (euler3)
      for (int q = 0; q < 3; q++) euler3[3 * i + q] = eu[3 * (size_t)i + q];
  }
  return AMCA_OK;
}

amca_status amca_sim_destroy(amca_sim* s) {
  delete s;
  return AMCA_OK;
}

}  // extern "C"
