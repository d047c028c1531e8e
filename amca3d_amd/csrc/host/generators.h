// Synthetic inputs of the coupled CA-FE path, produced on the host in the
// forms the reference reads: nodal temperature frames of a moving point heat
// source, and a Voronoi substrate (per-cell orientation label).
#pragma once
#include <stdint.h>

namespace amca_host {

struct RosenthalParams {
  double P, eta, v, k, alpha, T0;   // power, absorptivity, speed, conductivity, diffusivity, ambient
  double x_start, y_c, z_top;       // source position at t = 0 and track line
  double r_min;                     // distance floor (one cell)
};

// T(x, t) = T0 + eta*P/(2*pi*k*R) * exp(-v*(R + xi)/(2*alpha)), xi = x - (x_start + v*t),
// R = max(|(xi, y - y_c, z - z_top)|, r_min), for n_frames times t0 + f*dt.
void rosenthal_frames(int64_t n_nodes, const double* xyz, int n_frames, double t0, double dt,
                      const RosenthalParams& p, double* out, int threads = 0);

// Voronoi labels of the interior cells of a ghosted grid (ghosts get -1):
// n_seeds points uniform in the box and Euler angles uniform on SO(3)
// (alpha = 2*pi*U, beta = acos(2V - 1), gamma = W*pi/2), all from a
// splitmix64 stream of `seed`.  Nearest seed by squared distance, ties to the
// lower index.
void voronoi_labels(int nx, int ny, int nz, double h, double x0, double y0, double z0,
                    int n_seeds, uint64_t seed, int32_t* label, double* euler3, int threads = 0);

}  // namespace amca_host
