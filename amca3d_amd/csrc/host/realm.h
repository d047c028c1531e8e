// Host mirror of the reference's realm layer: YAML -> realm configuration ->
// device domain.  Class and key names follow CellularAutomataManager_3D
// (CellularAutomataManager_3D.cpp:70-379), ProblemPhysics_RappazGandin1993
// (ProblemPhysics.cpp:57-102), VtrManager::load (VtrManager.cpp:60-103) and
// CellularAutomataSolver (CellularAutomataSolver.cpp:40-202).
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "../../../include/amca3d.h"
#include "rng.h"
#include "sites.h"
#include "yaml_lite.h"

namespace amca_host {

struct NucleationRule {
  double density = 0.0;
  double mean = 0.0;
  double sd = 0.0;
};

// Everything a `cellular_automata` realm reads from its YAML node.
struct RealmConfig {
  std::string name;
  std::string type;          // cellular_automata | cellular_automata_remelting
  int dimension = 3;
  double origin[3] = {-1, -1, -1};
  double lateral[3] = {-1, -1, -1};
  double h = 0.0;
  int ncell[3] = {-1, -1, -1};
  // nucleation_rules (defaults of CellularAutomataManager.cpp:48-53)
  NucleationRule surface{0.0, 0.5, 0.1};
  NucleationRule bulk{0.0, 6.0, 0.1};
  SurfaceFaces surface_faces = SurfaceFaces::kReferenceSix;  // extension key `faces`
  unsigned random_seed = 0;  // extension key; 0 reproduces srand(0) of rank 0
  // problem_physics (ProblemPhysics.cpp:33-46 defaults)
  std::string physics = "RappazGandin";
  double T0 = 300.0, Tm = 0.0, Ts = -1.0e8, Tmin_nuc = -1.0e8, Tdot = -2.3;
  double a1 = 0, a2 = 0, a3 = 0, b1 = 1, b2 = 2, b3 = 3;
  // solution_options
  double trunc = 0.0;        // set from sqrtf(3) or envelope.max_half_diagonal_length
  double cutoff_factor = 1.0e8;
  bool output_micro = false;
  std::string micro_file;
  bool initial_melted = true;
  bool using_nucleation_seed = false;  // remelting: criterion_to_activate_grain
  // initial microstructure (CellularAutomataManager_3D.cpp:273-357): a file, or
  // random (one grain per interior cell, the default when nothing is named)
  std::string micro_load;
  bool micro_random = false;
  bool load_seeds = false;
  // output (VtrManager::load)
  int output_frequency = 1;
  std::string output_base = "./Results/Grains";
  double output_time_interval = -1.0;
  std::vector<std::string> output_variables;
};

// time_integrators[].standard_time_integrator
struct TimeIntegratorConfig {
  double termination_time = 0.0;
  double time_step = 1.0;
  int termination_step_count = 1000;
  double time_step_factor_ca = 0.2;
  double time_step_factor = -1.0;
};

// A `finite_element` realm that reads its temperatures from a text file
// (FiniteElementManager::load, FiniteElementManager.cpp:66-135).
struct FeRealmConfig {
  std::string name;
  std::string mesh;          // the text file of frames
  bool whole_time = false;   // load_data_from_file.for_whole_time
  int title_lines = 4, subtitle_lines = 5;
  double length_scale = 1.0, time_scale = 1.0;
  double offset[3] = {0, 0, 0};
};

// transfers[] (MapVoxelManager::load, MapVoxelManager.cpp:60-90)
struct TransferConfig {
  int bin[3] = {10, 10, 10};
  double tolerance = 1.0e-16;
  int max_iterations = 20;
  double void_temperature = -1.0e8;
};

struct SimulationConfig {
  bool cellular_automata = false;
  bool finite_element_method = false;
  std::vector<RealmConfig> realms;
  std::vector<FeRealmConfig> fe_realms;
  TransferConfig transfer;
  TimeIntegratorConfig ti;
};

// Parses a whole input deck (Simulation::load + CellularAutomataSolver::load).
SimulationConfig load_simulation(const yaml::Node& doc);
RealmConfig load_ca_realm(const yaml::Node& realm);

// A plain `cellular_automata` realm on the GPU.
class VtrSeries;

class PlainRealm {
 public:
  explicit PlainRealm(const RealmConfig& cfg, const TimeIntegratorConfig& ti);
  ~PlainRealm();
  // CellularAutomataManager_3D::initialize (:366-379), split in two: the host
  // draws (sites surface then bulk, orientation rows) and the device upload.
  void prepare();
  void initialize();
  // CellularAutomataManager::time_integration (CellularAutomataManager.cpp:134-147)
  amca_counters run(long long max_steps);
  // finalGrains.vtk and finalResult.txt of the current state (output_orientations,
  // CellularAutomataManager.cpp:462-465); an empty path skips that file
  void write_final_grains(const std::string& vtk, const std::string& result);
  // the VTR time series of the output block (off until enabled): frames after
  // the steps it names during run(), and the last frame (output_last_step_result)
  void enable_vtr(const std::string& base_dir, bool on);
  void vtr_final_frame();
  // output_microstructure_information (CellularAutomataManager_3D.cpp:1835-1913)
  // at the end of time_integration (CellularAutomataManager.cpp:144-145), when
  // the deck asks for it; relative names resolve against base_dir
  void write_microstructure(const std::string& base_dir);
  amca_domain* domain() const { return dom_; }
  const Grid& grid() const { return grid_; }
  const std::vector<int>& initial_sites() const { return site_cells_; }
  const std::vector<double>& initial_site_dt() const { return site_dt_; }
  const std::vector<double>& euler_rows() const { return euler_; }
  long long site_draws() const { return site_draws_; }
  int n_surface_sites() const { return n_surface_; }
  int n_bulk_sites() const { return n_bulk_; }

 private:
  RealmConfig cfg_;
  TimeIntegratorConfig ti_;
  Grid grid_{0, 0, 0};
  amca_domain* dom_ = nullptr;
  std::unique_ptr<VtrSeries> vtr_;
  void vtr_frame(double time);
  std::vector<int> site_cells_;
  std::vector<double> site_dt_;
  std::vector<double> euler_;
  std::vector<double> l2g_;
  long long site_draws_ = 0;
  int n_surface_ = 0, n_bulk_ = 0;
  bool prepared_ = false;
};

}  // namespace amca_host
