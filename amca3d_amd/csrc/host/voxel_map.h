// FE-to-CA voxel mapping, run once per FE mesh (MapVoxelManager.cpp:140-699):
// each CA cell (ghosts included) gets the FE element holding its centre and the
// parametric coordinates found by the reference's Newton iteration, so the
// device interpolation reproduces the reference's temperatures bit for bit.
#pragma once
#include <stdint.h>

#include <vector>

#include "../voxel_math.h"

namespace amca_host {

struct FeMesh {
  int fnx = 0, fny = 0, fnz = 0;   // nodes per axis; node id (k*fny + j)*fnx + i
  std::vector<double> xyz;         // 3 per node
  int nodes() const { return fnx * fny * fnz; }
  int elements() const { return (fnx - 1) * (fny - 1) * (fnz - 1); }
  void element_nodes(int e, int* nid) const;
  amca_vox::MeshView view() const { return amca_vox::MeshView{xyz.data(), fnx, fny, fnz}; }
};

struct VoxelMapParams {
  int nx, ny, nz;                  // CA interior cells
  double h, x0, y0, z0;
  int bin[3] = {10, 10, 10};       // transfers.bin_size
  double tol = 1.0e-16;            // transfers.tolerance
  int max_iter = 20;               // transfers.max_iterations
};

// The search box and its bins (CSR: bin b's elements, ascending, are
// bin_elems[bin_start[b] .. bin_start[b+1]]), as the host and device mappers share them.
struct VoxelPlan {
  amca_vox::VoxelGrid grid;
  std::vector<int> bin_start, bin_elems;
};
bool plan_voxels(const FeMesh& mesh, const VoxelMapParams& p, VoxelPlan& plan);

// tag[NG] (element or -1) and par[3*NG] (-2 where untagged); returns false when
// no cell found an element (the reference throws "a proc has no corresponding
// FEM mesh").
bool map_voxels(const FeMesh& mesh, const VoxelMapParams& p, std::vector<int32_t>& tag,
                std::vector<double>& par, int threads = 0);

}  // namespace amca_host
