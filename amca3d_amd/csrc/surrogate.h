// Lattice evaluation of the temperature surrogate, shared by the MLP and the
// coupled domain: T[k][j][i] = mlp(xi0 + i*dxi, y0 + j*dy, z0 + k*dz; scan),
// de-normalised to kelvin, written to device memory.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/amca3d.h"

struct SurLattice {
  double xi0, dxi;
  int nxi;
  double y0, dy;
  int ny;
  double z0, dz;
  int nz;
};

amca_status surrogate_eval_lattice(amca_mlp* m, const amca_surrogate_desc& d,
                                   const amca_scan& s, const SurLattice& g, float* out);
const char* surrogate_error();
