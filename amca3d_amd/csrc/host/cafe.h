// The coupled CA-FE simulation of the reference's CafeSolver (CafeSolver.cpp:
// 40-219): a `finite_element` realm that reads temperature frames from a text
// file, a `cellular_automata_remelting` realm, and the transfer settings,
// assembled from an input deck and run on the device (amca_remelt_*).
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "../../../include/amca3d.h"
#include "fe_text.h"
#include "micro_io.h"
#include "realm.h"

namespace amca_host {

class CafeRealm {
 public:
  CafeRealm(const SimulationConfig& cfg, std::string base_dir);
  ~CafeRealm();
  // Host part of the realms' load/initialize: frames, the initial
  // microstructure (file or random, CellularAutomataManager_3DRemelting.cpp:
  // 109-120), no GPU.
  void prepare();
  // Device domain: voxel map, state, sites (:122-131), orientation stream.
  void initialize();
  // CafeSolver::time_integration_weak_coupling (:178-219)
  amca_remelt_counters run(long long max_steps);
  // finalize: output_microstructure_information when the deck asks for it
  void finalize();
  // finalGrains.vtk and finalResult.txt of the current state (output_orientations,
  // CellularAutomataManager.cpp:462-465, with hasBeenMelted_); empty paths skip
  void write_final_grains(const std::string& vtk, const std::string& result);
  // the VTR time series of the CA realm's output block (off until enabled)
  void enable_vtr(const std::string& base_dir, bool on);
  void vtr_final_frame();
  amca_remelt* domain() const { return dom_; }
  const RealmConfig& ca() const { return ca_; }
  const FeFrames& frames() const { return frames_; }
  const MicroState& initial_state() const { return init_; }
  int n_surface_new() const { return n_surface_new_; }
  int n_bulk_new() const { return n_bulk_new_; }

 private:
  std::string path(const std::string& p) const;
  SimulationConfig cfg_;
  RealmConfig ca_;
  FeRealmConfig fe_;
  std::string base_;
  FeFrames frames_;
  MicroState init_;
  long long random_draws_ = 0;
  std::vector<int32_t> loaded_seed_cells_;  // load_seeds: sites of the loaded file
  std::vector<double> loaded_seed_dts_;
  int n_surface_new_ = 0, n_bulk_new_ = 0;
  amca_remelt* dom_ = nullptr;
  std::unique_ptr<VtrSeries> vtr_;
  void vtr_frame(double time);
  bool prepared_ = false;
};

}  // namespace amca_host
