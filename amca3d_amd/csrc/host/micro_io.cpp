#include "micro_io.h"

#include <fstream>
#include <sstream>
#include <stdexcept>

namespace amca_host {

std::string rank_file_name(const std::string& name, int rank) {
  const size_t dot = name.find_last_of(".");
  return name.substr(0, dot) + std::to_string(rank) + ".txt";
}

MicroState read_micro(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("fail to open the specified file");
  MicroState m;
  std::string line;
  int num_orient = 0, num_grains = 0;
  double nx = 0, ny = 0, nz = 0;
  std::getline(in, line);  // description
  std::getline(in, line);
  {
    std::istringstream iss(line);
    iss >> m.num_cells >> num_orient >> num_grains >> m.num_global_expected >> m.x0 >> m.y0 >> m.z0 >>
        nx >> ny >> nz >> m.h;
  }
  m.nxG = (int)nx;
  m.nyG = (int)ny;
  m.nzG = (int)nz;
  if (m.num_cells < 1) throw std::runtime_error("microstructure information file has no cells");
  const size_t nc = (size_t)m.num_cells;
  m.gid.assign(nc, -1);
  m.ori.assign(nc, -1);
  m.melted.assign(nc, 0);
  m.nuc_dt.assign(nc, -1.0);
  std::getline(in, line);  // description
  long long count = 0;
  while (count < m.num_cells) {
    if (!std::getline(in, line)) throw std::runtime_error("microstructure file ends in the cell block");
    if (line.empty()) continue;
    std::istringstream iss(line);
    long long id;
    int g, o;
    bool melt;
    double dt = -1.0;
    iss >> id >> g >> o >> melt >> dt;
    if (id < 0 || id >= m.num_cells) throw std::runtime_error("cell id out of range in microstructure file");
    m.gid[id] = g;
    m.ori[id] = o;
    m.melted[id] = melt ? 1 : 0;
    m.nuc_dt[id] = dt;
    count++;
  }
  std::getline(in, line);  // description
  count = 0;
  while (count < num_orient) {
    if (!std::getline(in, line)) throw std::runtime_error("microstructure file ends in the orientation block");
    if (line.empty()) continue;
    std::istringstream iss(line);
    int id;
    double e[3];
    iss >> id >> e[0] >> e[1] >> e[2];
    m.orient_id.push_back(id);
    m.euler.insert(m.euler.end(), e, e + 3);
    count++;
  }
  std::getline(in, line);  // description
  count = 0;
  while (count < num_grains) {
    if (!std::getline(in, line)) throw std::runtime_error("microstructure file ends in the grain block");
    if (line.empty()) continue;
    std::istringstream iss(line);
    int id, cell, o;
    double c[3];
    iss >> id >> cell >> c[0] >> c[1] >> c[2] >> o;
    m.grain_id.push_back(id);
    m.grain_cell.push_back(cell);
    m.grain_centre.insert(m.grain_centre.end(), c, c + 3);
    m.grain_orient.push_back(o);
    count++;
  }
  return m;
}

void write_micro(const std::string& path, const MicroState& m) {
  std::ofstream f(path);
  if (!f) throw std::runtime_error("fail to open the output file");
  const size_t no = m.orient_id.size(), ng = m.grain_id.size();
  f << "numCell; numOrientation; numLocalGrain; numGlobalExpectedGrain; start Point X; Y; Z; "
       "number Point nx; ny; nz; dCell"
    << "\n";
  f << m.num_cells << "  " << no << "  " << ng << "  " << m.num_global_expected << "  " << m.x0
    << "  " << m.y0 << "  " << m.z0 << "  " << m.nxG << "  " << m.nyG << "  " << m.nzG << "  "
    << m.h << std::endl;
  f << "Cell informatin: ID; grainID; orientationID; hasBeenMelted; undercooling of seed"
    << std::endl;
  for (long long c = 0; c < m.num_cells; c++) {
    f << c << "  " << m.gid[c] << "  " << m.ori[c] << "  " << (bool)m.melted[c] << "  "
      << m.nuc_dt[c];
    f << std::endl;
  }
  f << "Orientation informatin: ID; EularAngle" << std::endl;
  for (size_t o = 0; o < no; o++) {
    f << m.orient_id[o] << "  ";
    for (int i = 0; i < 3; i++) f << m.euler[3 * o + i] << "  ";
    f << std::endl;
  }
  f << "Grain information: ID; cellID; grainCenter; orientationID" << std::endl;
  for (size_t g = 0; g < ng; g++) {
    f << m.grain_id[g] << "  ";
    f << m.grain_cell[g] << "  ";
    f << m.grain_centre[3 * g] << "  ";
    f << m.grain_centre[3 * g + 1] << "  ";
    f << m.grain_centre[3 * g + 2] << "  ";
    f << m.grain_orient[g] << "   " << std::endl;
  }
  if (!f) throw std::runtime_error("write failed");
}

}  // namespace amca_host
