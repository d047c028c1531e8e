// Decentred-octahedron envelope arithmetic shared by every CA kernel.
//
// FP64 with the reference's operation order, compiled with -ffp-contract=off so
// no product is fused into an add: these values decide integer grain IDs and
// must match the reference bit for bit.  Two square roots of the reference go
// through single precision (`sqrtf` on a double argument, Octahedron.cpp:238-239);
// they are evaluated as the correctly rounded float square root of the
// float-rounded argument.
#pragma once
#include <hip/hip_runtime.h>

#define AMCA_HD __host__ __device__ __forceinline__

namespace amca {

// Octahedron.cpp:13
constexpr double kTheodorus = 1.732050807568877;

struct Vec3 {
  double x, y, z;
};

// Local (grain-frame) coordinates of a global offset `g`: local = R^T g with
// R the row-major local-to-global rotation, each component accumulated over
// the global index in order from 0.0 (Octahedron.cpp:113-121).
AMCA_HD Vec3 to_local(const double* __restrict__ r, double g0, double g1, double g2) {
  double cx = 0.0, cy = 0.0, cz = 0.0;
  cx += g0 * r[0];
  cy += g0 * r[1];
  cz += g0 * r[2];
  cx += g1 * r[3];
  cy += g1 * r[4];
  cz += g1 * r[5];
  cx += g2 * r[6];
  cy += g2 * r[7];
  cz += g2 * r[8];
  return {cx, cy, cz};
}

// Cell-centre coordinate along one axis: origin + h*(index - 0.5)
// (CellularAutomataManager_3D.cpp:885-887).
AMCA_HD double centre_coord(double origin, double h, int idx) {
  return origin + h * ((double)idx - 0.5);
}

// Octahedron::cell_is_captured (Octahedron.cpp:98-128): an envelope of
// half-diagonal L captures a cell centre when the L1 norm of its local
// coordinates does not exceed L; L == 0 never captures.
AMCA_HD bool inside_envelope(const double* __restrict__ r, double L, double gx, double gy,
                             double gz) {
  if (L == 0) return false;
  Vec3 c = to_local(r, gx, gy, gz);
  return fabs(c.x) + fabs(c.y) + fabs(c.z) <= 1.0 * L;
}

AMCA_HD double min_ref(double a, double b) { return (b < a) ? b : a; }  // std::min
AMCA_HD double max_ref(double a, double b) { return (a < b) ? b : a; }  // std::max

AMCA_HD double sqrt_via_float(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (double)__builtin_sqrtf((float)x);
#else
  return (double)sqrtf((float)x);
#endif
}

// Octahedron::capture_cell (Octahedron.cpp:133-275).  Given the captured cell
// centre relative to the parent centre (g), the parent half-diagonal L, the
// cell size h and the truncation coefficient, returns the shift of the child
// centre (out[0..2], to be added to the parent centre) and the child
// half-diagonal (out[3]).
AMCA_HD void child_envelope(const double* __restrict__ r, double L, double h, double trunc,
                            double sqrt2f, double gx, double gy, double gz, double* out) {
  Vec3 c = to_local(r, gx, gy, gz);
  double inter = fabs(c.x) + fabs(c.y) + fabs(c.z);
  // Face <111> holding the centre (Octahedron.cpp:162-176): faces 0..3
  // counter-clockwise from +x+y in the upper half-space, 4..7 below; the
  // comparisons and the default of 0 are kept as written.
  int q = 0;
  if (c.z >= 0) {
    if (c.x >= 0 && c.y >= 0) q = 0;
    else if (c.x >= 0 && c.y < 0) q = 3;
    else if (c.x < 0 && c.y >= 0) q = 1;
    else if (c.x < 0 && c.y < 0) q = 2;
  } else {
    if (c.x >= 0 && c.y >= 0) q = 4;
    else if (c.x >= 0 && c.y < 0) q = 7;
    else if (c.x < 0 && c.y >= 0) q = 5;
    else if (c.x < 0 && c.y < 0) q = 6;
  }
  // Face normal signs (Octahedron.cpp:27-47): n = (sx, sy, sz).
  const double sx = (q == 1 || q == 2 || q == 5 || q == 6) ? -1.0 : 1.0;
  const double sy = (q == 2 || q == 3 || q == 6 || q == 7) ? -1.0 : 1.0;
  const double sz = (q >= 4) ? -1.0 : 1.0;
  // Projection A of the centre onto the face (Octahedron.cpp:180-187).
  double perp = L - inter;
  double pa0 = c.x + perp * sx / 3;
  double pa1 = c.y + perp * sy / 3;
  double pa2 = c.z + perp * sz / 3;
  // The face's three vertices, in the order (Octahedron.cpp:61-84):
  // first two on the equator (0:+x, 1:+y, 2:-x, 3:-y) then the pole (4:+z, 5:-z).
  int fv[3];
  fv[0] = q & 3;
  fv[1] = (q & 3) == 3 ? 0 : (q & 3) + 1;
  fv[2] = (q < 4) ? 4 : 5;
  // vertex coordinates scaled by L
  double vc[3][3];
  for (int m = 0; m < 3; m++) {
    int id = fv[m];
    vc[m][0] = (id == 0) ? L : (id == 2) ? -L : 0.0;
    vc[m][1] = (id == 1) ? L : (id == 3) ? -L : 0.0;
    vc[m][2] = (id == 4) ? L : (id == 5) ? -L : 0.0;
    // 0.0*L in the reference is +0.0 for L > 0 and -0.0 never arises: the
    // table entries are +1, -1 or 0.0 and L > 0 here.
  }
  double dist[3];
  for (int m = 0; m < 3; m++) {
    double d = 0.0;
    double e0 = pa0 - vc[m][0], e1 = pa1 - vc[m][1], e2 = pa2 - vc[m][2];
    d += e0 * e0;
    d += e1 * e1;
    d += e2 * e2;
    dist[m] = d;
  }
  // Vertex ordering with the reference's swap rule (Octahedron.cpp:201-215):
  // slot 0 ends as the nearest vertex S1; a later nearer vertex moves the
  // previous slot-0 vertex into its own slot.
  int cid[3];
  double dmin = dist[0];
  cid[0] = 0;
  for (int m = 1; m < 3; m++) {
    if (dmin > dist[m]) {
      dmin = dist[m];
      cid[m] = cid[0];
      cid[0] = m;
    } else {
      cid[m] = m;
    }
  }
  const double* s1 = vc[cid[0]];
  const double* s2 = vc[cid[1]];
  const double* s3 = vc[cid[2]];
  double n12 = 0.0, n13 = 0.0, d12 = 0.0, d13 = 0.0;
  {
    double a0 = pa0 - s1[0], a1 = pa1 - s1[1], a2 = pa2 - s1[2];
    double e0 = s2[0] - s1[0], e1 = s2[1] - s1[1], e2 = s2[2] - s1[2];
    double f0 = s3[0] - s1[0], f1 = s3[1] - s1[1], f2 = s3[2] - s1[2];
    n12 += e0 * e0; d12 += a0 * e0; n13 += f0 * f0; d13 += a0 * f0;
    n12 += e1 * e1; d12 += a1 * e1; n13 += f1 * f1; d13 += a1 * f1;
    n12 += e2 * e2; d12 += a2 * e2; n13 += f2 * f2; d13 += a2 * f2;
  }
  n12 = sqrt_via_float(n12);
  n13 = sqrt_via_float(n13);
  double i0 = d12 / n12;
  double i1 = n12 - i0;
  double j0 = d13 / n13;
  double j1 = n13 - j0;
  // Truncation (Octahedron.cpp:248-253).
  double tv = trunc * h;
  if (i0 < tv && i1 < tv) {
    i0 = n12 / 2;
    i1 = n13 / 2;
  }
  double L1 = 0.5 * (min_ref(i0, tv) + min_ref(i1, tv));
  double L2 = 0.5 * (min_ref(j0, tv) + min_ref(j1, tv));
  double Lnew = sqrt2f * (max_ref(L1, L2));
  // Shift along S1, back to the global frame (Octahedron.cpp:266-274).
  double mag = (L - Lnew);
  double o0 = 0.0, o1 = 0.0, o2 = 0.0;
  for (int i = 0; i < 3; i++) {
    double comp = mag * s1[i] / (L * kTheodorus);
    o0 += comp * r[i + 0];
    o1 += comp * r[i + 3];
    o2 += comp * r[i + 6];
  }
  out[0] = o0;
  out[1] = o1;
  out[2] = o2;
  out[3] = Lnew;
}

}  // namespace amca
