#include "fe_text.h"

#include <algorithm>
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>

#include <unistd.h>

namespace amca_host {

FeFrames load_fe_text(const std::string& path, const FeTextOptions& o) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("fail to open the specified file");
  if (o.subtitle_lines < 4) throw std::runtime_error("lines_for_subtitle must be at least 4");
  FeFrames f;
  std::string line;
  int dummy[2];
  double time[2] = {0, 0};
  int range[6] = {0, 0, 0, 0, 0, 0};
  // title and first subtitle (:38-53): the 4th subtitle line holds the range
  const int skip = o.title_lines + o.subtitle_lines;
  for (int count = 0; count < skip; count++) {
    if (!std::getline(in, line)) throw std::runtime_error("fail to read temperature file, please check the file type and the lines of title!!");
    if (count == o.title_lines + 3) {
      std::istringstream iss(line);
      iss >> dummy[0] >> dummy[1] >> time[0] >> time[1];
      for (int i = 0; i < 6; i++) iss >> range[i];
    }
  }
  const double t_first = time[0];
  f.fnx = range[1] - range[0] + 1;
  f.fny = range[3] - range[2] + 1;
  f.fnz = range[5] - range[4] + 1;
  const long long nodes = (long long)f.fnx * f.fny * f.fnz;
  if (nodes <= 1 || f.fnx < 2 || f.fny < 2 || f.fnz < 2)
    throw std::runtime_error("fail to read temperature file, please check the file type and the lines of title!!");
  f.xyz.resize(3 * (size_t)nodes);
  f.theta.resize((size_t)nodes);
  f.times.push_back(0.0);  // inTimeSeries_[0] = 0 for the first frame
  for (long long n = 0; n < nodes; n++) {
    if (!std::getline(in, line)) throw std::runtime_error("temperature file ends inside the first frame");
    std::istringstream iss(line);
    for (int q = 0; q < 3; q++) {
      double c = 0;
      iss >> c;
      f.xyz[3 * n + q] = (c + o.offset[q]) * o.length_scale;
    }
    iss >> f.theta[n];
  }
  // later frames (:142-191)
  while (true) {
    std::vector<double> th((size_t)nodes);
    double tnow = 0;
    int count = 0;
    bool complete = true;
    while (count < o.subtitle_lines + nodes) {
      if (!std::getline(in, line)) {
        complete = false;
        break;
      }
      if (count == 0 && line != "\r" && !line.empty())
        throw std::runtime_error("fail to read next step temperature file, please check title of line!!");
      if (count == 3) {
        std::istringstream iss(line);
        iss >> dummy[0] >> dummy[1] >> time[0] >> time[1];
        tnow = (time[0] - t_first) * o.time_scale;
      }
      if (count >= o.subtitle_lines) {
        std::istringstream iss(line);
        double c[3];
        iss >> c[0] >> c[1] >> c[2] >> th[count - o.subtitle_lines];
      }
      count++;
    }
    if (!complete) break;  // end of file: the reference stops at the last whole frame
    f.times.push_back(tnow);
    f.theta.insert(f.theta.end(), th.begin(), th.end());
  }
  return f;
}

namespace {

// cache layout: magic, key (source bytes, source mtime, options), mesh shape,
// frame count, then xyz, times, theta as raw doubles
constexpr char kCacheMagic[8] = {'A', 'M', 'C', 'A', 'F', 'R', '0', '1'};

struct CacheKey {
  int64_t src_bytes, src_mtime;
  int32_t title_lines, subtitle_lines;
  double length_scale, time_scale, offset[3];
};

bool same_key(const CacheKey& a, const CacheKey& b) {
  return a.src_bytes == b.src_bytes && a.src_mtime == b.src_mtime &&
         a.title_lines == b.title_lines && a.subtitle_lines == b.subtitle_lines &&
         a.length_scale == b.length_scale && a.time_scale == b.time_scale &&
         a.offset[0] == b.offset[0] && a.offset[1] == b.offset[1] && a.offset[2] == b.offset[2];
}

}  // namespace

FeFrames load_fe_frames_cached(const std::string& path, const FeTextOptions& o) {
  namespace fs = std::filesystem;
  std::error_code ec;
  CacheKey key{};
  key.src_bytes = (int64_t)fs::file_size(path, ec);
  if (ec) return load_fe_text(path, o);  // reports the missing file as the loader does
  key.src_mtime = (int64_t)fs::last_write_time(path, ec).time_since_epoch().count();
  if (ec) return load_fe_text(path, o);
  key.title_lines = o.title_lines;
  key.subtitle_lines = o.subtitle_lines;
  key.length_scale = o.length_scale;
  key.time_scale = o.time_scale;
  for (int q = 0; q < 3; q++) key.offset[q] = o.offset[q];
  const std::string cache = path + ".amcacache";
  {
    std::ifstream in(cache, std::ios::binary);
    char magic[8];
    CacheKey k{};
    int32_t shape[4];
    if (in.read(magic, 8) && std::equal(magic, magic + 8, kCacheMagic) &&
        in.read((char*)&k, sizeof(k)) && same_key(k, key) &&
        in.read((char*)shape, sizeof(shape)) && shape[0] > 1 && shape[1] > 1 && shape[2] > 1 &&
        shape[3] > 0) {
      FeFrames f;
      f.fnx = shape[0];
      f.fny = shape[1];
      f.fnz = shape[2];
      const size_t n = (size_t)f.fnx * f.fny * f.fnz;
      f.xyz.resize(3 * n);
      f.times.resize((size_t)shape[3]);
      f.theta.resize(n * (size_t)shape[3]);
      if (in.read((char*)f.xyz.data(), sizeof(double) * f.xyz.size()) &&
          in.read((char*)f.times.data(), sizeof(double) * f.times.size()) &&
          in.read((char*)f.theta.data(), sizeof(double) * f.theta.size()))
        return f;
    }
  }
  FeFrames f = load_fe_text(path, o);
  // a name of this process's own, so that concurrent loads cannot interleave
  const std::string tmp = cache + "." + std::to_string((long long)getpid()) + ".tmp";
  {
    std::ofstream out(tmp, std::ios::binary);
    const int32_t shape[4] = {f.fnx, f.fny, f.fnz, f.frames()};
    out.write(kCacheMagic, 8);
    out.write((const char*)&key, sizeof(key));
    out.write((const char*)shape, sizeof(shape));
    out.write((const char*)f.xyz.data(), sizeof(double) * f.xyz.size());
    out.write((const char*)f.times.data(), sizeof(double) * f.times.size());
    out.write((const char*)f.theta.data(), sizeof(double) * f.theta.size());
    if (!out) {
      out.close();
      fs::remove(tmp, ec);
      return f;
    }
  }
  fs::rename(tmp, cache, ec);
  if (ec) fs::remove(tmp, ec);
  return f;
}

void write_fe_text(const std::string& path, int title_lines, int fnx, int fny, int fnz,
                   const double* xyz, int n_frames, const double* times, const double* theta) {
  std::ofstream out(path);
  if (!out) throw std::runtime_error("fail to open the output file");
  out << std::setprecision(17);
  for (int i = 0; i < title_lines; i++) out << (i == 0 ? "Title:" : "Temperature frames") << "\n";
  const long long nodes = (long long)fnx * fny * fnz;
  for (int f = 0; f < n_frames; f++) {
    if (f > 0) out << "\n";
    else out << "Subtitle\n";
    out << "frames of a structured hex mesh\n";
    out << "No. step   No. step    time   time   x0    x1   y0   y1    z0   z1\n";
    out << f + 1 << " " << f + 1 << " " << times[f] << " " << times[f] << " 0 " << fnx - 1
        << " 0 " << fny - 1 << " 0 " << fnz - 1 << "\n";
    out << "x  y  z  tn\n";
    for (long long n = 0; n < nodes; n++)
      out << xyz[3 * n] << " " << xyz[3 * n + 1] << " " << xyz[3 * n + 2] << " "
          << theta[(size_t)f * nodes + n] << "\n";
  }
  if (!out) throw std::runtime_error("write failed");
}

}  // namespace amca_host
