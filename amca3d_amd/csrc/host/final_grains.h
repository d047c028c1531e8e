// The end-of-run grain files of the reference: finalGrains.vtk (legacy ASCII
// rectilinear grid) and finalResult.txt (one text row per cell).
#pragma once

#include <cstdint>
#include <string>

namespace amca_host {

// CellularAutomataManager_3D::writeToFile (CellularAutomataManager_3D.cpp:1381-1639),
// called through output_orientations (CellularAutomataManager.cpp:462-465) with the
// cell orientation ids.  `orient` and `melted` are ghost-inclusive
// ((nx+2)(ny+2)(nz+2), x fastest); melted may be null (every flag 1, as for a
// realm without remelting); euler3 holds 3 angles per orientation id.  Either
// path may be empty, which skips that file.  Ids are written unshuffled.
void write_final_grains(const std::string& vtk_path, const std::string& result_path, int nx,
                        int ny, int nz, double x0, double y0, double z0, double h,
                        const int32_t* orient, const int32_t* melted, const double* euler3,
                        int32_t n_orient);

}  // namespace amca_host
