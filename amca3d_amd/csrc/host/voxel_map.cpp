#include "voxel_map.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <thread>

namespace amca_host {

void FeMesh::element_nodes(int e, int* nid) const {
  amca_vox::element_nodes(view(), e, nid);
}

bool plan_voxels(const FeMesh& M, const VoxelMapParams& p, VoxelPlan& plan) {
  amca_vox::VoxelGrid& G = plan.grid;
  const amca_vox::MeshView mv = M.view();
  G.nxG = p.nx + 2;
  G.nyG = p.ny + 2;
  G.nzG = p.nz + 2;
  G.h = p.h;
  const double h = p.h;
  // get_CA_domain_xyz_min_max (:198-256) at one rank
  double box[6];
  box[0] = (0 - 1) * h + p.x0;
  box[2] = (0 - 1) * h + p.y0;
  box[4] = (0 - 1) * h + p.z0;
  box[1] = box[0] + G.nxG * h;
  box[3] = box[2] + G.nyG * h;
  box[5] = box[4] + G.nzG * h;
  double grow[6];
  std::copy(box, box + 6, grow);
  const int ne = M.elements();
  for (int e = 0; e < ne; e++) {
    int nid[8];
    amca_vox::element_nodes(mv, e, nid);
    bool overlap = false;
    for (int m = 0; m < 8 && !overlap; m++) {
      const double* c = &M.xyz[3 * (size_t)nid[m]];
      overlap = c[0] >= box[0] && c[0] <= box[1] && c[1] >= box[2] && c[1] <= box[3] &&
                c[2] >= box[4] && c[2] <= box[5];
    }
    if (!overlap) continue;
    double r[6];
    amca_vox::element_box(mv, e, r);
    for (int q = 0; q < 3; q++) {
      grow[2 * q] = (grow[2 * q] < r[2 * q]) ? grow[2 * q] : r[2 * q];
      grow[2 * q + 1] = (grow[2 * q + 1] > r[2 * q + 1]) ? grow[2 * q + 1] : r[2 * q + 1];
    }
  }
  std::copy(grow, grow + 6, G.box);
  for (int q = 0; q < 3; q++) {
    G.dBin[q] = p.bin[q] * h;
    G.nBin[q] = (int)std::ceil((G.box[2 * q + 1] - G.box[2 * q]) / G.dBin[q]);
  }
  G.org[0] = (0 - 0.5) * h + p.x0;
  G.org[1] = (0 - 0.5) * h + p.y0;
  G.org[2] = (0 - 0.5) * h + p.z0;
  // attach_elements (:294-352): elements wholly inside the box, per bin, in
  // ascending element id (a count pass, then a fill pass)
  const long long nb = (long long)G.nBin[0] * G.nBin[1] * G.nBin[2];
  plan.bin_start.assign((size_t)nb + 1, 0);
  for (int pass = 0; pass < 2; pass++) {
    std::vector<int> fill;
    if (pass == 1) {
      for (long long b = 0; b < nb; b++) plan.bin_start[(size_t)b + 1] += plan.bin_start[(size_t)b];
      plan.bin_elems.assign((size_t)plan.bin_start[(size_t)nb], 0);
      fill.assign(plan.bin_start.begin(), plan.bin_start.end() - 1);
    }
    for (int e = 0; e < ne; e++) {
      double r[6];
      amca_vox::element_box(mv, e, r);
      bool inside = true;
      for (int q = 0; q < 3 && inside; q++)
        inside = !(r[2 * q] < G.box[2 * q] || r[2 * q + 1] > G.box[2 * q + 1]);
      if (!inside) continue;
      const int imin = (int)std::floor((r[0] - G.box[0]) / G.dBin[0]);
      const int imax = (int)std::ceil((r[1] - G.box[0]) / G.dBin[0]);
      const int jmin = (int)std::floor((r[2] - G.box[2]) / G.dBin[1]);
      const int jmax = (int)std::ceil((r[3] - G.box[2]) / G.dBin[1]);
      const int kmin = (int)std::floor((r[4] - G.box[4]) / G.dBin[2]);
      const int kmax = (int)std::ceil((r[5] - G.box[4]) / G.dBin[2]);
      for (int kk = kmin; kk < kmax; kk++)
        for (int jj = jmin; jj < jmax; jj++)
          for (int ii = imin; ii < imax; ii++) {
            if (ii < 0 || jj < 0 || kk < 0 || ii >= G.nBin[0] || jj >= G.nBin[1] || kk >= G.nBin[2])
              continue;
            const size_t b = ii + jj * (size_t)G.nBin[0] + kk * (size_t)G.nBin[0] * G.nBin[1];
            if (pass == 0)
              plan.bin_start[b + 1]++;
            else
              plan.bin_elems[(size_t)fill[b]++] = e;
          }
    }
  }
  return true;
}

bool map_voxels(const FeMesh& M, const VoxelMapParams& p, std::vector<int32_t>& tag,
                std::vector<double>& par, int threads) {
  VoxelPlan plan;
  plan_voxels(M, p, plan);
  const amca_vox::VoxelGrid& G = plan.grid;
  const amca_vox::MeshView mv = M.view();
  const long long NG = (long long)G.nxG * G.nyG * G.nzG;
  tag.assign((size_t)NG, -1);
  par.assign(3 * (size_t)NG, -2.0);
  // tag_voxels (:357-482), voxels split over threads
  auto work = [&](long long v0, long long v1) {
    for (long long v = v0; v < v1; v++) {
      double pc[3];
      const int e = amca_vox::tag_voxel(mv, G, plan.bin_start.data(), plan.bin_elems.data(), v,
                                        p.tol, p.max_iter, pc);
      if (e < 0) continue;
      tag[(size_t)v] = e;
      for (int q = 0; q < 3; q++) par[3 * (size_t)v + q] = pc[q];
    }
  };
  int nt = threads > 0 ? threads : (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (NG < 100000) nt = 1;
  std::vector<std::thread> pool;
  const long long chunk = (NG + nt - 1) / nt;
  for (int t = 0; t < nt; t++) {
    const long long a = t * chunk, b = std::min(NG, a + chunk);
    if (a < b) pool.emplace_back(work, a, b);
  }
  for (auto& th : pool) th.join();
  for (long long v = 0; v < NG; v++)
    if (tag[(size_t)v] >= 0) return true;
  return false;
}

}  // namespace amca_host
