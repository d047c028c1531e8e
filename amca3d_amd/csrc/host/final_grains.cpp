#include "final_grains.h"

#include <fstream>
#include <stdexcept>
#include <vector>

namespace amca_host {
namespace {

// Values go ten to a line.  The line counter runs on across sections: a
// section ends its last partial line but keeps the count, so the next
// section's first break comes early.  Only the z coordinates reset it.
struct Rows {
  std::ostream& os;
  int count = 0;
  void value_written() {
    if (++count % 10 == 0) {
      os << '\n';
      count = 0;
    }
  }
  void end_section(bool reset) {
    if (count != 0) {
      os << '\n';
      if (reset) count = 0;
    }
  }
};

}  // namespace

void write_final_grains(const std::string& vtk_path, const std::string& result_path, int nx,
                        int ny, int nz, double x0, double y0, double z0, double h,
                        const int32_t* orient, const int32_t* melted, const double* euler3,
                        int32_t n_orient) {
  if (nx < 1 || ny < 1 || nz < 1 || !orient) throw std::runtime_error("final grains: bad grid");
  const long long sx = nx + 2, sxy = (long long)(nx + 2) * (ny + 2);
  auto at = [&](int i, int j, int k) { return (long long)k * sxy + (long long)j * sx + i; };
  // cells per orientation id over the interior cells
  int max_id = -1;
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        const int v = orient[at(i, j, k)];
        if (v > max_id) max_id = v;
      }
  if (max_id >= 0 && (!euler3 || max_id >= n_orient))
    throw std::runtime_error("final grains: an orientation id has no Euler angles");
  std::vector<long long> cells((size_t)(max_id + 1), 0);
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        const int v = orient[at(i, j, k)];
        if (v >= 0) cells[(size_t)v]++;
      }
  auto melt = [&](long long c) { return melted ? melted[c] : 1; };
  std::vector<char> buf(1 << 20);

  if (!vtk_path.empty()) {
    std::ofstream f;
    f.rdbuf()->pubsetbuf(buf.data(), (std::streamsize)buf.size());
    f.open(vtk_path);
    if (!f) throw std::runtime_error("fail to open " + vtk_path);
    f << "# vtk DataFile Version 2.0\nTest Data\nASCII\nDATASET RECTILINEAR_GRID\n";
    f << "DIMENSIONS " << nx + 1 << " " << ny + 1 << " " << nz + 1 << '\n';
    Rows r{f};
    const char* axis[3] = {"X", "Y", "Z"};
    const int n[3] = {nx, ny, nz};
    const double o[3] = {x0, y0, z0};
    for (int a = 0; a < 3; a++) {
      f << axis[a] << "_COORDINATES " << n[a] + 1 << " double\n";
      f << o[a] << " ";  // the origin is not counted
      for (int q = 0; q < n[a]; q++) {
        f << o[a] + (q + 1) * h << " ";
        r.value_written();
      }
      r.end_section(a == 2);
      f << '\n';
    }
    f << "CELL_DATA " << (long long)nx * ny * nz << '\n';
    f << "SCALARS grain_id double 1\nLOOKUP_TABLE default\n";
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
          f << orient[at(i, j, k)] << " ";
          r.value_written();
        }
    r.end_section(false);
    f << "SCALARS number_cells int 1\nLOOKUP_TABLE default\n";
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
          const int v = orient[at(i, j, k)];
          if (v >= 0)
            f << cells[(size_t)v] << " ";
          else
            f << v << " ";
          r.value_written();
        }
    r.end_section(false);
    f << "SCALARS melted_flag int 1\nLOOKUP_TABLE default\n";
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
          f << melt(at(i, j, k)) << " ";
          r.value_written();
        }
    r.end_section(false);
    f << "VECTORS EulerAngle float\n";
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
          const int v = orient[at(i, j, k)];
          if (v >= 0) {
            const double* e = euler3 + 3 * (size_t)v;
            f << e[0] << " " << e[1] << " " << e[2] << " ";
            r.value_written();
          } else {
            f << 0.0 << " " << 0.0 << " " << 0.0 << " ";  // not counted
          }
        }
    f << '\n';
    if (!f) throw std::runtime_error("fail to write " + vtk_path);
  }

  if (!result_path.empty()) {
    std::ofstream f;
    f.rdbuf()->pubsetbuf(buf.data(), (std::streamsize)buf.size());
    f.open(result_path);
    if (!f) throw std::runtime_error("fail to open " + result_path);
    f << "X0  Y0  Z0   Dcell  nx_cell  ny_cell  nz_cell\n";
    f << x0 << " " << y0 << " " << z0 << " " << h << " " << nx << " " << ny << " " << nz << '\n';
    f << "grain_ID  number_cells  melted_flag  EulerAngle1 EulerAngle2 EulerAngle3\n";
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
          const long long c = at(i, j, k);
          const int v = orient[c];
          f << v << " ";
          if (v >= 0)
            f << cells[(size_t)v] << " ";
          else
            f << v << " ";
          f << melt(c) << " ";
          if (v >= 0) {
            const double* e = euler3 + 3 * (size_t)v;
            f << e[0] << " " << e[1] << " " << e[2] << " ";
          } else {
            f << 0.0 << " " << 0.0 << " " << 0.0 << " ";
          }
          f << '\n';
        }
    if (!f) throw std::runtime_error("fail to write " + result_path);
  }
}

}  // namespace amca_host
