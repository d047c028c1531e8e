#include "cafe.h"

#include "final_grains.h"
#include "vtr_series.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "sites.h"

namespace amca_host {

namespace {

void check(amca_status s, const char* what) {
  if (s != AMCA_OK)
    throw std::runtime_error(std::string(what) + ": " + amca_remelt_last_error());
}

}  // namespace

CafeRealm::CafeRealm(const SimulationConfig& cfg, std::string base_dir)
    : cfg_(cfg), base_(std::move(base_dir)) {
  for (const auto& r : cfg.realms)
    if (r.type == "cellular_automata_remelting") ca_ = r;
  if (ca_.type.empty()) throw std::runtime_error("no cellular_automata_remelting realm");
  if (cfg.fe_realms.empty()) throw std::runtime_error("no finite_element realm");
  fe_ = cfg.fe_realms[0];
  if (!fe_.whole_time)
    throw std::runtime_error("the finite_element realm must read its frames (load_data_from_file.for_whole_time)");
}

CafeRealm::~CafeRealm() {
  if (dom_) amca_remelt_destroy(dom_);
}

std::string CafeRealm::path(const std::string& p) const {
  if (p.empty() || p[0] == '/' || base_.empty()) return p;
  return base_ + "/" + p;
}

void CafeRealm::prepare() {
  FeTextOptions o;
  o.title_lines = fe_.title_lines;
  o.subtitle_lines = fe_.subtitle_lines;
  o.length_scale = fe_.length_scale;
  o.time_scale = fe_.time_scale;
  for (int a = 0; a < 3; a++) o.offset[a] = fe_.offset[a];
  frames_ = load_fe_frames_cached(path(fe_.mesh), o);
  if (frames_.frames() < 2) throw std::runtime_error("the temperature file holds fewer than two frames");

  Grid g{ca_.ncell[0], ca_.ncell[1], ca_.ncell[2]};
  const long long NG = g.NG();
  if (ca_.micro_random) {
    // random_microstructure_information (CellularAutomataManager_3D.cpp:1928-1984):
    // one grain per interior cell, in cell order, with a random orientation
    GlibcRand rng(ca_.random_seed);
    init_ = MicroState();
    init_.num_cells = NG;
    init_.gid.assign((size_t)NG, -1);
    init_.ori.assign((size_t)NG, -1);
    init_.melted.assign((size_t)NG, ca_.initial_melted ? 1 : 0);
    init_.nuc_dt.assign((size_t)NG, -1.0);
    int id = 0;
    for (long long c = 0; c < NG; c++) {
      const int k = (int)(c / g.nxyG()), j = (int)((c % g.nxyG()) / g.nxG()),
                i = (int)((c % g.nxyG()) % g.nxG());
      if (i == 0 || j == 0 || k == 0 || i == g.nxG() - 1 || j == g.nyG() - 1 || k == g.nzG() - 1)
        continue;
      double e[3];
      random_orientation(rng, e);
      init_.orient_id.push_back(id);
      init_.euler.insert(init_.euler.end(), e, e + 3);
      init_.grain_id.push_back(id);
      init_.grain_cell.push_back((int32_t)c);
      init_.grain_orient.push_back(id);
      init_.gid[c] = id;
      init_.ori[c] = id;
      id++;
    }
    random_draws_ = 3LL * id;
  } else {
    init_ = read_micro(path(rank_file_name(ca_.micro_load)));
    if (init_.num_cells != NG)
      throw std::runtime_error("Error : microstructure information file is not fit the current domain size");
    for (size_t q = 0; q < init_.orient_id.size(); q++)
      if (init_.orient_id[q] != (int)q)
        throw std::runtime_error("orientation ids of the microstructure file must be 0, 1, 2, ...");
    for (size_t q = 0; q < init_.grain_id.size(); q++)
      if (init_.grain_id[q] != (int)q)
        throw std::runtime_error("grain ids of the microstructure file must be 0, 1, 2, ...");
    random_draws_ = 0;
  }
  prepared_ = true;
}

void CafeRealm::initialize() {
  if (!prepared_) prepare();
  amca_remelt_desc d{};
  d.nx = ca_.ncell[0];
  d.ny = ca_.ncell[1];
  d.nz = ca_.ncell[2];
  d.h = ca_.h;
  d.x0 = ca_.origin[0];
  d.y0 = ca_.origin[1];
  d.z0 = ca_.origin[2];
  d.Tm = ca_.Tm;
  d.Ts = ca_.Ts;
  d.Tmin_nuc = ca_.Tmin_nuc;
  d.a1 = ca_.a1;
  d.a2 = ca_.a2;
  d.a3 = ca_.a3;
  d.b1 = ca_.b1;
  d.b2 = ca_.b2;
  d.b3 = ca_.b3;
  d.trunc = ca_.trunc;
  d.time_step_factor = cfg_.ti.time_step_factor_ca;
  d.cutoff_factor = ca_.cutoff_factor;
  d.final_time = cfg_.ti.termination_time;
  d.void_temperature = cfg_.transfer.void_temperature;
  for (int a = 0; a < 3; a++) d.bin[a] = cfg_.transfer.bin[a];
  d.tolerance = cfg_.transfer.tolerance;
  d.max_iterations = cfg_.transfer.max_iterations;
  d.initial_melted = ca_.initial_melted ? 1 : 0;
  const long long n_init = (long long)init_.orient_id.size();
  const long long cap = n_init + (1LL << 20);
  if (cap > 2147483647LL) throw std::runtime_error("orientation table exceeds int32 rows");
  check(amca_remelt_create(&d, (int32_t)cap, &dom_), "amca_remelt_create");
  check(amca_remelt_push_orientations(dom_, (int32_t)n_init, init_.euler.data(), nullptr),
        "push_orientations");
  check(amca_remelt_set_fe(dom_, frames_.fnx, frames_.fny, frames_.fnz, frames_.xyz.data(),
                           frames_.frames(), frames_.times.data(), frames_.theta.data()),
        "set_fe");
  check(amca_remelt_set_state(dom_, init_.ori.data(), init_.gid.data(), init_.melted.data(),
                              (int32_t)n_init, (int32_t)init_.grain_cell.size(),
                              init_.grain_cell.data(), init_.grain_orient.data()),
        "set_state");
  // Sites.  With load_seeds the cells of the loaded file whose seed undercooling
  // is > -1 form the list, in cell order (CellularAutomataManager_3DRemelting.cpp:244-249);
  // with using_nucleation_seed the surface and bulk draws merge into it
  // (:121-130, CellularAutomataManager.cpp:153-213) and nucleation runs.
  // Without using_nucleation_seed no site nucleates (:660-663); loaded seeds
  // are only written back out.
  loaded_seed_cells_.clear();
  loaded_seed_dts_.clear();
  if (!ca_.micro_random && ca_.load_seeds)
    for (long long c = 0; c < init_.num_cells; c++)
      if (init_.nuc_dt[c] > -1) {
        if (init_.nuc_dt[c] < 0)
          throw std::runtime_error("a loaded nucleation seed has a negative undercooling");
        loaded_seed_cells_.push_back((int32_t)c);
        loaded_seed_dts_.push_back(init_.nuc_dt[c]);
      }
  if (ca_.using_nucleation_seed) {
    check(amca_remelt_set_sites(dom_, (int32_t)loaded_seed_cells_.size(), loaded_seed_cells_.data(),
                                loaded_seed_dts_.data()),
          "set_sites");
    check(amca_remelt_draw_sites_after(dom_, ca_.random_seed, random_draws_, ca_.surface.density,
                                       ca_.surface.mean, ca_.surface.sd, ca_.bulk.density,
                                       ca_.bulk.mean, ca_.bulk.sd, &n_surface_new_, &n_bulk_new_),
          "draw_sites");
  }
}

amca_remelt_counters CafeRealm::run(long long max_steps) {
  if (!dom_) initialize();
  amca_remelt_counters c{};
  if (!vtr_) {
    check(amca_remelt_run(dom_, max_steps, &c), "amca_remelt_run");
    return c;
  }
  // VtrManager::execute after every step (CellularAutomataManager_3DRemelting.cpp:640)
  check(amca_remelt_get_counters(dom_, &c), "counters");
  long long taken = 0;
  while (!c.done && (max_steps < 0 || taken < max_steps)) {
    long long chunk = vtr_->steps_to_next(c.step);
    if (chunk == 0) chunk = -1;  // no frame can come during this run
    if (max_steps >= 0) chunk = chunk < 0 ? max_steps - taken : std::min(chunk, max_steps - taken);
    const long long s0 = c.step;
    check(amca_remelt_run(dom_, chunk, &c), "amca_remelt_run");
    taken += c.step - s0;
    if (c.step == s0) break;
    if (vtr_->due(c.step, c.time)) vtr_frame(c.time);
  }
  return c;
}

void CafeRealm::enable_vtr(const std::string& base_dir, bool on) {
  vtr_.reset();
  if (on) vtr_ = std::make_unique<VtrSeries>(ca_, base_dir, true);
}

void CafeRealm::vtr_frame(double time) {
  Grid g{ca_.ncell[0], ca_.ncell[1], ca_.ncell[2]};
  std::vector<int32_t> ori((size_t)g.NG());
  std::vector<double> T((size_t)g.NG());
  std::vector<uint8_t> melted((size_t)g.NG());
  int32_t ng = 0;
  check(amca_remelt_download(dom_, ori.data(), nullptr, T.data(), melted.data(), nullptr, 0,
                             nullptr, nullptr, nullptr, &ng),
        "download");
  int32_t no = 0;
  check(amca_remelt_euler(dom_, 0, nullptr, &no), "euler");
  std::vector<double> eu(3 * (size_t)std::max(no, 1));
  check(amca_remelt_euler(dom_, no, eu.data(), &no), "euler");
  VtrFields f;
  f.orient = ori.data();
  f.T = T.data();
  f.melted = melted.data();
  f.euler3 = eu.data();
  f.n_orient = no;
  vtr_->write(time, f);
}

void CafeRealm::vtr_final_frame() {
  if (!vtr_ || !dom_) return;
  amca_remelt_counters c{};
  check(amca_remelt_get_counters(dom_, &c), "counters");
  if (vtr_->due(0, c.time)) vtr_frame(c.time);  // output_last_step_result
}

void CafeRealm::write_final_grains(const std::string& vtk, const std::string& result) {
  if (!dom_) return;
  Grid g{ca_.ncell[0], ca_.ncell[1], ca_.ncell[2]};
  std::vector<int32_t> ori((size_t)g.NG()), melted((size_t)g.NG());
  std::vector<uint8_t> m8((size_t)g.NG());
  int32_t ng = 0;
  check(amca_remelt_download(dom_, ori.data(), nullptr, nullptr, m8.data(), nullptr, 0, nullptr,
                             nullptr, nullptr, &ng),
        "download");
  for (size_t c = 0; c < m8.size(); c++) melted[c] = m8[c];
  int32_t no = 0;
  check(amca_remelt_euler(dom_, 0, nullptr, &no), "euler");
  std::vector<double> eu(3 * (size_t)std::max(no, 1));
  check(amca_remelt_euler(dom_, no, eu.data(), &no), "euler");
  amca_host::write_final_grains(vtk, result, g.nx, g.ny, g.nz, ca_.origin[0], ca_.origin[1],
                                ca_.origin[2], ca_.h, ori.data(), melted.data(), eu.data(), no);
}

void CafeRealm::finalize() {
  if (!dom_ || !ca_.output_micro) return;
  Grid g{ca_.ncell[0], ca_.ncell[1], ca_.ncell[2]};
  const long long NG = g.NG();
  MicroState m;
  m.num_cells = NG;
  m.num_global_expected = n_surface_new_ + n_bulk_new_ + (ca_.micro_random ? 0 : init_.num_global_expected);
  m.x0 = ca_.origin[0];
  m.y0 = ca_.origin[1];
  m.z0 = ca_.origin[2];
  m.h = ca_.h;
  m.nxG = g.nxG();
  m.nyG = g.nyG();
  m.nzG = g.nzG();
  m.gid.resize((size_t)NG);
  m.ori.resize((size_t)NG);
  m.melted.resize((size_t)NG);
  int32_t ng = 0;
  check(amca_remelt_download(dom_, nullptr, nullptr, nullptr, nullptr, nullptr, 0, nullptr, nullptr,
                             nullptr, &ng),
        "download");
  std::vector<double> genv(4 * (size_t)std::max(ng, 1));
  std::vector<int32_t> gc((size_t)std::max(ng, 1)), go((size_t)std::max(ng, 1));
  check(amca_remelt_download(dom_, m.ori.data(), m.gid.data(), nullptr, m.melted.data(), nullptr,
                             ng, genv.data(), gc.data(), go.data(), &ng),
        "download");
  // cellNucleationDT_: every site ever listed keeps its undercooling
  m.nuc_dt.assign((size_t)NG, -1.0);
  if (ca_.using_nucleation_seed) {
    int32_t ns = 0;
    check(amca_remelt_sites(dom_, 0, nullptr, nullptr, &ns), "sites");
    std::vector<int32_t> sc((size_t)std::max(ns, 1));
    std::vector<double> sd((size_t)std::max(ns, 1));
    check(amca_remelt_sites(dom_, ns, sc.data(), sd.data(), &ns), "sites");
    for (int q = 0; q < ns; q++) m.nuc_dt[sc[q]] = sd[q];
  } else {
    for (size_t q = 0; q < loaded_seed_cells_.size(); q++)
      m.nuc_dt[loaded_seed_cells_[q]] = loaded_seed_dts_[q];
  }
  int32_t no = 0;
  check(amca_remelt_euler(dom_, 0, nullptr, &no), "euler");
  m.euler.resize(3 * (size_t)std::max(no, 1));
  check(amca_remelt_euler(dom_, no, m.euler.data(), &no), "euler");
  m.euler.resize(3 * (size_t)no);
  for (int o = 0; o < no; o++) m.orient_id.push_back(o);
  for (int q = 0; q < ng; q++) {
    m.grain_id.push_back(q);
    m.grain_cell.push_back(gc[q]);
    m.grain_centre.insert(m.grain_centre.end(), &genv[4 * (size_t)q], &genv[4 * (size_t)q] + 3);
    m.grain_orient.push_back(go[q]);
  }
  write_micro(path(rank_file_name(ca_.micro_file)), m);
}

}  // namespace amca_host
