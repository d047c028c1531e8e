#include "vtr_series.h"

#include <algorithm>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <stdexcept>

namespace amca_host {
namespace {

// the output variable names of the reference, by id (CafeMacro.h:16-24)
const char* const kVarNames[] = {"temperature",    "temperature_rate",  "grain_ID",
                                 "grain_orientation", "grain_velocity", "element_birth",
                                 "melted_flag",    "phase_ID",          "alpha_phase",
                                 "phase_orientation", "velocity",       "debug",
                                 "debugMC"};
constexpr int kTemperature = 0, kGrainId = 2, kOrientation = 3, kMelted = 6, kDebug = 11,
              kDebugMC = 12;

std::string pad5(int v) {
  std::string s = std::to_string(v);
  return s.size() >= 5 ? s : std::string(5 - s.size(), '0') + s;
}

const char* const kW2 = "  ";
const char* const kW3 = "    ";
const char* const kW4 = "      ";
const char* const kW5 = "        ";
const char* const kW6 = "          ";

}  // namespace

VtrSeries::VtrSeries(const RealmConfig& cfg, const std::string& base_dir, bool remelting)
    : nx_(cfg.ncell[0]), ny_(cfg.ncell[1]), nz_(cfg.ncell[2]), x0_(cfg.origin[0]),
      y0_(cfg.origin[1]), z0_(cfg.origin[2]), h_(cfg.h), freq_(cfg.output_frequency),
      by_freq_(cfg.output_frequency >= 1), interval_(cfg.output_time_interval),
      threshold_(cfg.output_time_interval > 0 ? 0.0 : 1.0e8) {
  for (const std::string& v : cfg.output_variables) {
    for (int id = 0; id < (int)(sizeof(kVarNames) / sizeof(kVarNames[0])); id++)
      if (v == kVarNames[id]) vars_.push_back(id);  // unknown names are skipped
  }
  if (vars_.empty()) vars_.push_back(kGrainId);
  if (!remelting && std::find(vars_.begin(), vars_.end(), kMelted) != vars_.end())
    throw std::runtime_error("output_variables: melted_flag needs a cellular_automata_remelting "
                             "realm");
  // VtrManager::initialize (:156-169): the folder is what precedes the last
  // '/', the name what follows it (cut at the last '.' as the reference cuts)
  const std::string& b = cfg.output_base;
  const size_t slash = b.find_last_of('/'), dot = b.find_last_of('.');
  const std::string dir = b.substr(0, slash);
  name_ = dot ? b.substr(slash + 1, dot) : b.substr(slash + 1);
  dir_ = (base_dir.empty() || (!dir.empty() && dir[0] == '/')) ? dir : base_dir + "/" + dir;
  if (freq_ != 0) std::filesystem::create_directories(dir_ + "/grainsTimeStates");
}

bool VtrSeries::due(long long step, double time) {
  if (by_freq_) return step % freq_ == 0;
  if (time >= threshold_) {
    threshold_ += interval_;
    threshold_ = std::max((double)(int)(time / interval_ + 1) * interval_, threshold_);
    return true;
  }
  return false;
}

long long VtrSeries::steps_to_next(long long step) const {
  if (by_freq_) return freq_ - step % freq_;
  return threshold_ < 1.0e8 ? 1 : 0;
}

void VtrSeries::write(double time, const VtrFields& f) {
  const int n = n_out_;
  const std::string frame = dir_ + "/grainsTimeStates/timeStep_" + pad5(n);
  std::filesystem::create_directories(frame);
  stamps_.emplace_back(n, time);
  const long long sx = nx_ + 2, sxy = (long long)(nx_ + 2) * (ny_ + 2);
  auto at = [&](int i, int j, int k) { return (long long)k * sxy + (long long)j * sx + i; };
  const std::string extent = "0 " + std::to_string(nx_) + " 0 " + std::to_string(ny_) + " 0 " +
                             std::to_string(nz_);
  std::vector<char> buf(1 << 20);
  {
    const std::string path = frame + "/" + name_ + pad5(0) + ".vtr";
    std::ofstream o;
    o.rdbuf()->pubsetbuf(buf.data(), (std::streamsize)buf.size());
    o.open(path);
    if (!o) throw std::runtime_error("fail to open " + path);
    o << "<?xml version=\"1.0\"?>\n"
      << "<VTKFile type=\"RectilinearGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
      << kW2 << "<RectilinearGrid WholeExtent=\"" << extent << "\">\n"
      << kW3 << "<Piece Extent=\"" << extent << "\">\n"
      << kW4 << "<CellData Scalars=\"grain_id\">\n";
    for (int v : vars_) {
      if (v == kTemperature && f.T) {
        o << kW5 << "<DataArray type=\"Float64\" Name=\"cell_temperature\" format=\"ascii\">\n";
        for (int k = 1; k <= nz_; k++)
          for (int j = 1; j <= ny_; j++) {
            o << kW6;
            for (int i = 1; i <= nx_; i++) o << (int)f.T[at(i, j, k)] << " ";  // as int
            o << '\n';
          }
      } else if (v == kGrainId) {
        o << kW5 << "<DataArray type=\"Int64\" Name=\"grain_id\" format=\"ascii\">\n";
        for (int k = 1; k <= nz_; k++)
          for (int j = 1; j <= ny_; j++) {
            o << kW6;
            for (int i = 1; i <= nx_; i++) o << f.orient[at(i, j, k)] << " ";
            o << '\n';
          }
      } else if (v == kOrientation) {
        o << kW5 << "<DataArray type=\"Float64\" Name=\"grain_orientation\" "
                    "NumberOfComponents=\"3\" format=\"ascii\">\n";
        for (int k = 1; k <= nz_; k++)
          for (int j = 1; j <= ny_; j++) {
            o << kW6;
            for (int i = 1; i <= nx_; i++) {
              const int id = f.orient[at(i, j, k)];
              if (id >= 0) {
                if (!f.euler3 || id >= f.n_orient)
                  throw std::runtime_error("vtr: an orientation id has no Euler angles");
                const double* e = f.euler3 + 3 * (size_t)id;
                o << e[0] << " " << e[1] << " " << e[2] << " ";
              } else {
                o << "0 0 0 ";
              }
            }
            o << '\n';
          }
      } else if (v == kMelted && f.melted) {
        o << kW5 << "<DataArray type=\"Int8\" Name=\"melted_flag\" format=\"ascii\">\n";
        for (int k = 1; k <= nz_; k++)
          for (int j = 1; j <= ny_; j++) {
            o << kW6;
            for (int i = 1; i <= nx_; i++) o << (int)f.melted[at(i, j, k)] << " ";
            o << '\n';
          }
      } else {
        continue;  // nothing written for this variable
      }
      o << "\n" << kW5 << "</DataArray>\n";
    }
    o << kW4 << "</CellData>\n" << kW4 << "<Coordinates>\n";
    const char* axis[3] = {"X", "Y", "Z"};
    const int cnt[3] = {nx_, ny_, nz_};
    const double org[3] = {x0_, y0_, z0_};
    for (int a = 0; a < 3; a++) {
      o << kW5 << "<DataArray type=\"Float64\" name=\"" << axis[a]
        << "_Coordinates\" format=\"ascii\">\n"
        << kW6;
      for (int q = 0; q <= cnt[a]; q++) o << org[a] + q * h_ << " ";
      o << kW5 << "</DataArray>\n";  // on the values' line
    }
    o << kW4 << "</Coordinates>\n\n"
      << kW3 << "</Piece>\n"
      << kW2 << "</RectilinearGrid>\n"
      << "</VTKFile>\n";
    if (!o) throw std::runtime_error("fail to write " + path);
  }
  {
    const std::string path = frame + "/" + name_ + pad5(n) + ".pvtr";
    std::ofstream o(path);
    if (!o) throw std::runtime_error("fail to open " + path);
    o << "<?xml version=\"1.0\"?>\n"
      << "<VTKFile type=\"PRectilinearGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
      << kW2 << "<PRectilinearGrid  WholeExtent=\"" << extent << "\" \n"
      << "                  GhostLevel = \"0\">\n"
      << kW3 << "<PCellData Scalars=\"grain_id\">\n";
    for (int v : vars_) {
      if (v == kTemperature)
        o << kW4 << "<PDataArray type=\"Float64\" Name=\"cell_temperature\"  format=\"ascii\"/>\n";
      else if (v == kGrainId)
        o << kW4 << "<PDataArray type=\"Int64\" Name=\"grain_id\" format=\"ascii\"/>\n";
      else if (v == kOrientation)
        o << kW4 << "<PDataArray type=\"Float64\" Name=\"grain_orientation\" "
                    "NumberOfComponents=\"3\" format=\"ascii\"/>\n";
      else if (v == kMelted)
        o << kW4 << "<PDataArray type=\"Int8\" Name=\"melted_flag\" format=\"ascii\"/>\n";
      else if (v == kDebug)
        o << kW4 << "<PDataArray type=\"Int8\" Name=\"growth_flag\"  format=\"ascii\"/>\n"
          << kW4 << "<PDataArray type=\"Float64\" Name=\"cooling\"  format=\"ascii\"/>\n";
      else if (v == kDebugMC)
        for (const char* nm : {"p", "tMCS", "vf_a", "vf_b", "Hardness"})
          o << kW4 << "<PDataArray type=\"Float64\" Name=\"" << nm << "\"  format=\"ascii\"/>\n";
    }
    o << kW3 << "</PCellData>\n" << kW3 << "<PCoordinates>\n";
    for (int a = 0; a < 3; a++) o << kW4 << "<PDataArray type=\"Float64\" format=\"ascii\"/>\n";
    o << kW3 << "</PCoordinates>\n"
      << kW3 << "<Piece Extent=\"" << extent << "\"\n"
      << kW3 << "      Source=\"" << name_ << pad5(0) << ".vtr\" />\n"
      << kW2 << "</PRectilinearGrid>\n"
      << "</VTKFile>\n";
    if (!o) throw std::runtime_error("fail to write " + path);
  }
  {
    const std::string path = dir_ + "/" + name_ + ".pvd";
    std::ofstream o(path);
    if (!o) throw std::runtime_error("fail to open " + path);
    o << "<?xml version=\"1.0\"?>\n"
      << "<VTKFile type=\"Collection\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
      << kW2 << "<Collection>\n";
    o << std::fixed << std::setprecision(8);
    for (size_t c = 0; c < stamps_.size(); c++)
      o << kW3 << "<DataSet timestep=\"" << stamps_[c].second
        << "\" group=\"\" part=\"0\" file=\"./grainsTimeStates/timeStep_"
        << pad5(stamps_[c].first) << "/" << name_ << pad5((int)c) << ".pvtr\"/>\n";
    o << kW2 << "</Collection>\n" << "</VTKFile>\n";
    if (!o) throw std::runtime_error("fail to write " + path);
  }
  n_out_++;
}

}  // namespace amca_host
