// Microstructure information files: the text the reference writes at the end
// of a run (output_microstructure_information, CellularAutomataManager_3D.cpp:
// 1835-1913) and reads to start the next track (load_microstructure_information,
// CellularAutomataManager_3DRemelting.cpp:165-307).  The rank number is
// appended to the stem: "grainInfo.txt" -> "grainInfo0.txt" at one rank.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace amca_host {

struct MicroState {
  // header
  long long num_cells = 0;
  int num_global_expected = 0;
  double x0 = 0, y0 = 0, z0 = 0, h = 0;
  int nxG = 0, nyG = 0, nzG = 0;
  // per cell
  std::vector<int32_t> gid, ori;
  std::vector<uint8_t> melted;
  std::vector<double> nuc_dt;  // -1 where no seed
  // orientations: id + Euler angles
  std::vector<int32_t> orient_id;
  std::vector<double> euler;   // 3 per row
  // grains
  std::vector<int32_t> grain_id, grain_cell, grain_orient;
  std::vector<double> grain_centre;  // 3 per grain (as written)
};

std::string rank_file_name(const std::string& name, int rank = 0);
MicroState read_micro(const std::string& path);
// Writes with the reference's stream formatting (default precision, two-space
// separators) so that files compare byte for byte.
void write_micro(const std::string& path, const MicroState& m);

}  // namespace amca_host
