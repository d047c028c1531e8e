// Host-side draws that the reference makes with its CPU RNG: nucleation sites
// and orientation rows.  They stay on the host because they consume the single
// glibc rand() stream in a fixed order; the device consumes their results.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <unordered_map>
#include <vector>

#include "rng.h"

namespace amca_host {

// Ghosted grid geometry of one rank (np = 1: the whole domain).
struct Grid {
  int nx, ny, nz;
  int nxG() const { return nx + 2; }
  int nyG() const { return ny + 2; }
  int nzG() const { return nz + 2; }
  int nxyG() const { return nxG() * nyG(); }
  long long NG() const { return (long long)nxyG() * nzG(); }
  int cell(int i, int j, int k) const { return k * nxyG() + j * nxG() + i; }
};

enum class SurfaceFaces { kReferenceSix = 0, kZMin = 1 };

// Candidate cells for one nucleation population:
// get_bulk_cells (CellularAutomataManager_3D.cpp:934-953) and get_surface_cells
// (:958-1000, six faces with edge cells twice and corners three times, and
// its row test against the x extent kept as written), or the z-min face only.
void bulk_cells(const Grid& g, std::vector<int>& ids);
void surface_cells(const Grid& g, SurfaceFaces faces, std::vector<int>& ids);

// CellularAutomataManager::create_nucleation_sites (CellularAutomataManager.cpp:153-213):
// draws round(cellArea*density*numCells)+0.5 sites with replacement, keeps the
// smallest critical undercooling per cell, merges into `sites` (cell -> dT)
// and returns the number of new sites.  `cell_dt` is the per-cell table
// (-1 = none) shared across populations.
struct SiteDraw {
  int n_draws = 0;
  int n_new = 0;
};
// `cell_at(i)` returns the i-th candidate cell (i < num_cells).
template <class CellAt>
SiteDraw create_nucleation_sites(int num_cells, CellAt cell_at, double area_or_volume,
                                 double density, double mean, double sd, GlibcRand& rng,
                                 std::unordered_map<int, double>& cell_dt,
                                 std::vector<int>& site_list) {
  SiteDraw r;
  const int numSites = round(area_or_volume * density * num_cells) + 0.5;
  SiteNormal normal(mean, sd);
  for (int s = 0; s < numSites; ++s) {
    const int c = cell_at(rng.next() % num_cells);
    r.n_draws++;
    double dT = normal.next();
    dT = (dT < 0) ? 0 : dT;
    auto it = cell_dt.find(c);
    if (it == cell_dt.end()) {
      site_list.push_back(c);
      cell_dt.emplace(c, dT);
      r.n_new++;
    } else if (dT < it->second) {
      it->second = dT;
    }
  }
  // std::list::sort of the merged list: ascending cell, unique cells
  std::sort(site_list.begin(), site_list.end());
  return r;
}

// i-th cell of get_bulk_cells without building the list
struct BulkCellAt {
  Grid g;
  int operator()(int idx) const {
    const int i = idx % g.nx;
    const int j = (idx / g.nx) % g.ny;
    const int k = idx / (g.nx * g.ny);
    return g.cell(i + 1, j + 1, k + 1);
  }
};

// get_random_orientation_angle (CellularAutomataManager_3D.cpp:838-850) and the
// rotation of Octahedron::set_coordinate_transformation_matrix (Octahedron.cpp:281-310).
void random_orientation(GlibcRand& rng, double* euler3);
void rotation_from_euler(const double* euler3, double* l2g9);

}  // namespace amca_host
