#include "sites.h"

#include <cmath>

namespace amca_host {

namespace {
const double kPi = 3.14159265358979323846;  // CellularAutomataManager_3D.cpp:24
}

void bulk_cells(const Grid& g, std::vector<int>& ids) {
  ids.clear();
  ids.reserve((size_t)g.nx * g.ny * g.nz);
  for (int k = 1; k <= g.nz; ++k)
    for (int j = 1; j <= g.ny; ++j)
      for (int i = 1; i <= g.nx; ++i) ids.push_back(g.cell(i, j, k));
}

void surface_cells(const Grid& g, SurfaceFaces faces, std::vector<int>& ids) {
  ids.clear();
  if (faces == SurfaceFaces::kZMin) {
    for (int j = 1; j <= g.ny; ++j)
      for (int i = 1; i <= g.nx; ++i) ids.push_back(g.cell(i, j, 1));
    return;
  }
  // Interior cells in index order; each face a cell lies on adds one entry.
  // The row bound is compared with the x extent (nx+1), as the reference does.
  const int nxG = g.nxG(), nzG = g.nzG();
  for (int k = 1; k < nzG - 1; ++k) {
    for (int j = 1; j < g.nyG() - 1; ++j) {
      if (j == nxG - 1) continue;
      for (int i = 1; i < nxG - 1; ++i) {
        const int c = g.cell(i, j, k);
        if (k == 1 || k == g.nz) ids.push_back(c);
        if (i == 1 || i == g.nx) ids.push_back(c);
        if (j == 1 || j == g.ny) ids.push_back(c);
      }
    }
  }
}

void random_orientation(GlibcRand& rng, double* a) {
  const double U = static_cast<double>(rng.next()) / GlibcRand::kRandMax;
  const double V = static_cast<double>(rng.next()) / GlibcRand::kRandMax;
  a[0] = U * 2 * kPi;
  a[1] = acos(2 * V - 1);
  a[2] = static_cast<double>(rng.next()) / GlibcRand::kRandMax * 0.5 * kPi;
}

void rotation_from_euler(const double* e, double* m) {
  const double ca = cos(e[0]), sa = sin(e[0]);
  const double cb = cos(e[1]), sb = sin(e[1]);
  const double cg = cos(e[2]), sg = sin(e[2]);
  m[0] = ca * cg - sa * cb * sg;
  m[1] = -ca * sg - sa * cb * cg;
  m[2] = sa * sb;
  m[3] = sa * cg + ca * cb * sg;
  m[4] = -sa * sg + ca * cb * cg;
  m[5] = -ca * sb;
  m[6] = sb * sg;
  m[7] = sb * cg;
  m[8] = cb;
}

}  // namespace amca_host
