// The reference's VTR time series (VtrManager / VtkManager) for one rank.
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "realm.h"

namespace amca_host {

// The fields one frame writes (ghost-inclusive, x fastest); T and melted may
// be null (a realm without remelting has neither).
struct VtrFields {
  const int32_t* orient = nullptr;  // cellOrientationID_
  const double* T = nullptr;        // cellTemperature_
  const uint8_t* melted = nullptr;  // hasBeenMelted_
  const double* euler3 = nullptr;   // Euler angles by orientation id
  int32_t n_orient = 0;
};

// VtrManager::load / initialize / execute (VtrManager.cpp:41-182, 504-561) and
// the VtkManager helpers (VtkManager.cpp:67-260): a frame is due after step n
// when n % output_frequency == 0, or, with output_frequency < 1, when the time
// passes a threshold that output_time_interval advances.  A frame writes
// <dir>/grainsTimeStates/timeStep_NNNNN/<name>00000.vtr and <name>NNNNN.pvtr,
// and rewrites <dir>/<name>.pvd with every time stamp so far.  <dir> and
// <name> come from output_data_base_name (everything before / after its last
// '/').  Supported variables: temperature (remelting realms; a realm without
// a temperature field writes only its header line in the .pvtr, as the
// reference), grain_ID, grain_orientation, melted_flag (remelting realms);
// the others the reference lists write nothing, or (debug, debugMC) .pvtr
// declarations only.
class VtrSeries {
 public:
  VtrSeries(const RealmConfig& cfg, const std::string& base_dir, bool remelting);
  // VtkManager::output_result; call once after every step, in order
  bool due(long long step, double time);
  // steps to run before the next call of due() can be true (frequency mode),
  // or 1 (time mode); 0 when no frame can come during a run
  long long steps_to_next(long long step) const;
  void write(double time, const VtrFields& f);
  int frames() const { return n_out_; }

 private:
  int nx_, ny_, nz_;
  double x0_, y0_, z0_, h_;
  int freq_;
  bool by_freq_;
  double interval_, threshold_;
  std::vector<int> vars_;
  std::string dir_, name_;
  int n_out_ = 0;
  std::vector<std::pair<int, double>> stamps_;
};

}  // namespace amca_host
