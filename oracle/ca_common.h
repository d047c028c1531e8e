// ORACLE -- test infrastructure only.  Envelope arithmetic shared by the
// plain and the remelting restatements (Octahedron.cpp, Grain.h).
#pragma once
#include <cmath>
#include <algorithm>

namespace oracle_ca {

const double kPi = 3.14159265358979323846;       // CellularAutomataManager_3D.cpp:24
const double kTheodorus = 1.732050807568877;      // Octahedron.cpp:13

struct Grain {
  int id;
  int cell;
  double xc, yc, zc;
  int orient;
  double length;
  bool active;
  int wave;  // capture-cascade depth within the step that created it
};

struct Orient {
  double euler[3];
  double l2g[9];
};

// Octahedron::set_coordinate_transformation_matrix (Octahedron.cpp:281-310).
// The Z-X-Z rotation, written row-major; its transpose maps global to local.
inline void rotation_from_euler(const double* e, double* l2g) {
  double ca = cos(e[0]), sa = sin(e[0]);
  double cb = cos(e[1]), sb = sin(e[1]);
  double cg = cos(e[2]), sg = sin(e[2]);
  l2g[0] = ca * cg - sa * cb * sg;
  l2g[1] = -ca * sg - sa * cb * cg;
  l2g[2] = sa * sb;
  l2g[3] = sa * cg + ca * cb * sg;
  l2g[4] = -sa * sg + ca * cb * cg;
  l2g[5] = -ca * sb;
  l2g[6] = sb * sg;
  l2g[7] = sb * cg;
  l2g[8] = cb;
}

// Octahedron::cell_is_captured (Octahedron.cpp:98-128).
inline bool capture_test(const Grain& g, const double* l2g, const double* cc) {
  double L = g.length;
  if (L == 0) return false;
  double gc0 = cc[0] - g.xc, gc1 = cc[1] - g.yc, gc2 = cc[2] - g.zc;
  double cx = 0, cy = 0, cz = 0;
  // local = R^T * gc, accumulated term by term in index order
  cx += gc0 * l2g[0]; cy += gc0 * l2g[1]; cz += gc0 * l2g[2];
  cx += gc1 * l2g[3]; cy += gc1 * l2g[4]; cz += gc1 * l2g[5];
  cx += gc2 * l2g[6]; cy += gc2 * l2g[7]; cz += gc2 * l2g[8];
  return fabs(cx) + fabs(cy) + fabs(cz) <= 1.0 * L;
}

// Octahedron::capture_cell (Octahedron.cpp:133-275): the decentred child
// octahedron.  `gc` is the cell centre relative to the parent centre; out[0..2]
// is the centre shift, out[3] the new half-diagonal.
inline void child_envelope(const double* gc, double L, double h, double trunc,
                    const double* l2g, double* out) {
  // unit normals of the eight <111> faces, and the six vertex directions
  static const double nrm[8][3] = {{1, 1, 1},  {-1, 1, 1},  {-1, -1, 1},  {1, -1, 1},
                                   {1, 1, -1}, {-1, 1, -1}, {-1, -1, -1}, {1, -1, -1}};
  static const double vtx[6][3] = {{1, 0, 0}, {0, 1, 0}, {-1, 0, 0},
                                   {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
  static const int face[8][3] = {{0, 1, 4}, {1, 2, 4}, {2, 3, 4}, {3, 0, 4},
                                 {0, 1, 5}, {1, 2, 5}, {2, 3, 5}, {3, 0, 5}};
  out[0] = out[1] = out[2] = out[3] = 0.0;
  double cx = 0, cy = 0, cz = 0;
  cx += gc[0] * l2g[0]; cy += gc[0] * l2g[1]; cz += gc[0] * l2g[2];
  cx += gc[1] * l2g[3]; cy += gc[1] * l2g[4]; cz += gc[1] * l2g[5];
  cx += gc[2] * l2g[6]; cy += gc[2] * l2g[7]; cz += gc[2] * l2g[8];
  double inter = fabs(cx) + fabs(cy) + fabs(cz);
  double corner[6][3];
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 3; j++) corner[i][j] = vtx[i][j] * L;

  // octant of the cell centre in the grain frame (Octahedron.cpp:162-176)
  int q = 0;
  if (cz >= 0) {
    if (cx >= 0 && cy >= 0) q = 0;
    else if (cx >= 0 && cy < 0) q = 3;
    else if (cx < 0 && cy >= 0) q = 1;
    else if (cx < 0 && cy < 0) q = 2;
  } else {
    if (cx >= 0 && cy >= 0) q = 4;
    else if (cx >= 0 && cy < 0) q = 7;
    else if (cx < 0 && cy >= 0) q = 5;
    else if (cx < 0 && cy < 0) q = 6;
  }
  // projection A of the centre onto face q (Octahedron.cpp:180-187)
  double pc[3] = {cx, cy, cz};
  double perp = L - inter;
  double pa[3];
  for (int i = 0; i < 3; i++) pa[i] = pc[i] + perp * nrm[q][i] / 3;

  // squared distances from A to the face's three vertices (Octahedron.cpp:191-200)
  double dist[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < 3; i++) {
    int id = face[q][i];
    for (int j = 0; j < 3; j++) dist[i] += (pa[j] - corner[id][j]) * (pa[j] - corner[id][j]);
  }
  // vertex ordering with the reference's swap rule (Octahedron.cpp:201-215)
  int cid[3];
  double dmin = dist[0];
  cid[0] = face[q][0];
  for (int i = 1; i < 3; i++) {
    if (dmin > dist[i]) {
      dmin = dist[i];
      cid[i] = cid[0];
      cid[0] = face[q][i];
    } else {
      cid[i] = face[q][i];
    }
  }
  int s1 = cid[0], s2 = cid[1], s3 = cid[2];
  // projections on S1S2 and S1S3 (Octahedron.cpp:222-245); the norms go
  // through single precision
  double n12 = 0.0, n13 = 0.0, d12 = 0, d13 = 0.0;
  for (int i = 0; i < 3; i++) {
    double s1a = pa[i] - corner[s1][i];
    double e12 = corner[s2][i] - corner[s1][i];
    n12 += e12 * e12;
    d12 += s1a * e12;
    double e13 = corner[s3][i] - corner[s1][i];
    n13 += e13 * e13;
    d13 += s1a * e13;
  }
  n12 = sqrtf((float)n12);
  n13 = sqrtf((float)n13);
  double is1[2], js1[2];
  is1[0] = d12 / n12;
  is1[1] = n12 - is1[0];
  js1[0] = d13 / n13;
  js1[1] = n13 - js1[0];
  // truncation (Octahedron.cpp:248-253)
  double tv = trunc * h;
  if (is1[0] < tv && is1[1] < tv) {
    is1[0] = n12 / 2;
    is1[1] = n13 / 2;
  }
  double L1 = 0.5 * (std::min(is1[0], tv) + std::min(is1[1], tv));
  double L2 = 0.5 * (std::min(js1[0], tv) + std::min(js1[1], tv));
  double Lnew = (double)sqrtf(2.0f) * (std::max(L1, L2));
  // centre shift along S1, back to the global frame (Octahedron.cpp:266-274)
  double mag = (L - Lnew);
  for (int i = 0; i < 3; i++) {
    double comp = mag * corner[s1][i] / (L * kTheodorus);
    for (int j = 0; j < 3; j++) out[j] += comp * l2g[i + j * 3];
  }
  out[3] = Lnew;
}


}  // namespace oracle_ca
