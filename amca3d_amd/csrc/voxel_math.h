// FE-to-CA voxel mapping arithmetic shared by the host mapper (host/voxel_map.cpp)
// and the device kernel (remelt_domain.hip): one definition, compiled by both,
// so that the element tags and parametric coordinates agree bit for bit.
// MapVoxelManager.cpp:263-289 (element box), :357-482 (tag_voxels),
// :487-699 (is_in_element).
#pragma once

#include <float.h>
#include <math.h>
#include <stdint.h>

#ifdef __HIPCC__
#define AMCA_HD __host__ __device__
#else
#define AMCA_HD
#endif

namespace amca_vox {

// A structured hex mesh: node id (k*fny + j)*fnx + i, 3 coordinates per node.
struct MeshView {
  const double* xyz;
  int fnx, fny, fnz;
};

// The CA grid (ghosts included), the mesh-covered search box and its bins
// (get_CA_domain_xyz_min_max :198-256, attach_elements :294-352).
struct VoxelGrid {
  int nxG, nyG, nzG;
  double org[3];   // centre of cell (0, 0, 0)
  double h;
  double box[6];   // x0 x1 y0 y1 z0 z1
  double dBin[3];
  int nBin[3];
};

// LoadDataFromFile.cpp:110-115 node order
AMCA_HD inline void element_nodes(const MeshView& M, int e, int* nid) {
  const int ncx = M.fnx - 1, ncy = M.fny - 1;
  const int ei = e % ncx, ej = (e / ncx) % ncy, ek = e / (ncx * ncy);
  const int nxy = M.fnx * M.fny;
  nid[0] = nxy * ek + M.fnx * ej + ei;
  nid[1] = nid[0] + 1;
  nid[2] = nid[1] + M.fnx;
  nid[3] = nid[2] - 1;
  for (int m = 4; m < 8; m++) nid[m] = nid[m - 4] + nxy;
}

// bounding box of an element: min/max per axis, with a < b ? a : b and a > b ? a : b
AMCA_HD inline void element_box(const MeshView& M, int e, double* r) {
  int nid[8];
  element_nodes(M, e, nid);
  for (int q = 0; q < 3; q++) {
    r[2 * q] = 1.0e8;
    r[2 * q + 1] = -1.0e8;
  }
  for (int m = 0; m < 8; m++) {
    const double* c = &M.xyz[3 * (size_t)nid[m]];
    for (int q = 0; q < 3; q++) {
      r[2 * q] = (r[2 * q] < c[q]) ? r[2 * q] : c[q];
      r[2 * q + 1] = (r[2 * q + 1] > c[q]) ? r[2 * q + 1] : c[q];
    }
  }
}

AMCA_HD inline double fac(int s, double u) { return s > 0 ? (1.0 + u) : (1.0 - u); }

// Jacobian columns of the reference's Newton step (MapVoxelManager.cpp:538-617):
// entry (axis a, derivative d) = sum over nodes m = 1..7 of
// sign * (f1 * f2) * coord_a[m], accumulated in node order, where f1, f2 are
// the factors (1 +- s) of the two other parametric coordinates:
// d = 0: factors (eta, zeta); d = 1: (xi, zeta); d = 2: (eta, xi).
AMCA_HD inline double jac_entry(int d, const double* coord, double xi, double eta, double zeta) {
  // (sign, s1, s2) per node m = 1..7
  const signed char kJ[3][7][3] = {
      {{-1, -1, -1}, {-1, +1, -1}, {+1, +1, -1}, {+1, -1, +1}, {-1, -1, +1}, {-1, +1, +1},
       {+1, +1, +1}},
      {{+1, +1, -1}, {-1, +1, -1}, {-1, -1, -1}, {+1, -1, +1}, {+1, +1, +1}, {-1, +1, +1},
       {-1, -1, +1}},
      {{+1, -1, +1}, {+1, +1, +1}, {+1, +1, -1}, {-1, -1, -1}, {-1, -1, +1}, {-1, +1, +1},
       {-1, +1, -1}},
  };
  double u1, u2;
  if (d == 0) {
    u1 = eta;
    u2 = zeta;
  } else if (d == 1) {
    u1 = xi;
    u2 = zeta;
  } else {
    u1 = eta;
    u2 = xi;
  }
  double acc = 0.0;
  for (int m = 0; m < 7; m++) {
    const signed char* t = kJ[d][m];
    const double term = (fac(t[1], u1) * fac(t[2], u2)) * coord[m + 1];
    if (m == 0)
      acc = t[0] > 0 ? term : -term;
    else
      acc = t[0] > 0 ? acc + term : acc - term;
  }
  return acc;
}

// MapVoxelManager::is_in_element (:487-699)
AMCA_HD inline bool in_element(const MeshView& M, int e, const double* pt, double tol,
                               int max_iter, double* pc) {
  int nid[8];
  element_nodes(M, e, nid);
  double c[3][8];
  const double* o = &M.xyz[3 * (size_t)nid[0]];
  for (int q = 0; q < 3; q++) c[q][0] = 0.0;
  for (int m = 1; m < 8; m++) {
    const double* p = &M.xyz[3 * (size_t)nid[m]];
    for (int q = 0; q < 3; q++) c[q][m] = 0.125 * (p[q] - o[q]);
  }
  const double rp[3] = {pt[0] - o[0], pt[1] - o[1], pt[2] - o[2]};
  double nw[3] = {0.5, 0.5, 0.5};
  double cur[3] = {0.5, 0.5, 0.5};
  double diff[3] = {1.0, 1.0, 1.0};
  int it = 0;
  do {
    double j[9];
    for (int a = 0; a < 3; a++)
      for (int d = 0; d < 3; d++) j[3 * a + d] = jac_entry(d, c[a], cur[0], cur[1], cur[2]);
    const double jdet = -(j[2] * j[4] * j[6]) + j[1] * j[5] * j[6] + j[2] * j[3] * j[7] -
                        j[0] * j[5] * j[7] - j[1] * j[3] * j[8] + j[0] * j[4] * j[8];
    if (!jdet) {
      it = max_iter;
      break;
    }
    const double xi = cur[0], eta = cur[1], zeta = cur[2];
    // shape functions times 8, factors in the order (eta, xi, zeta)
    const signed char se[8] = {-1, -1, +1, +1, -1, -1, +1, +1};
    const signed char sx[8] = {-1, +1, +1, -1, -1, +1, +1, -1};
    const signed char sz[8] = {-1, -1, -1, -1, +1, +1, +1, +1};
    double sf[8];
    for (int m = 0; m < 8; m++) sf[m] = fac(se[m], eta) * fac(sx[m], xi) * fac(sz[m], zeta);
    double f[3];
    for (int a = 0; a < 3; a++) {
      double v = rp[a];
      for (int m = 1; m < 8; m++) v = v - sf[m] * c[a][m];
      f[a] = v;
    }
    nw[0] = (jdet * xi + f[2] * (j[2] * j[4] - j[1] * j[5]) + f[1] * (j[1] * j[8] - j[2] * j[7]) +
             f[0] * (j[5] * j[7] - j[4] * j[8])) / jdet;
    nw[1] = (eta * jdet + f[2] * (j[0] * j[5] - j[2] * j[3]) + f[1] * (j[2] * j[6] - j[0] * j[8]) +
             f[0] * (j[3] * j[8] - j[5] * j[6])) / jdet;
    nw[2] = (jdet * zeta + f[2] * (j[1] * j[3] - j[0] * j[4]) + f[1] * (j[0] * j[7] - j[1] * j[6]) +
             f[0] * (j[4] * j[6] - j[3] * j[7])) / jdet;
    for (int q = 0; q < 3; q++) {
      diff[q] = nw[q] - cur[q];
      cur[q] = nw[q];
    }
  } while (!(fabs(0.0 + diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2]) < tol) &&
           ++it < max_iter);
  double dist = DBL_MAX;
  for (int q = 0; q < 3; q++) pc[q] = DBL_MAX;
  if (it < max_iter) {
    double dd = 0;
    for (int q = 0; q < 3; q++) {
      double v = nw[q];
      if (v > 1.0 && v < 1.000001) v = 1.0;
      if (v < -1.0 && v > -1.000001) v = -1.0;
      pc[q] = v;
      if (dd < fabs(v)) dd = fabs(v);
    }
    dist = dd;
  }
  return !(dist > 1.0);
}

// tag_voxels (:357-482) for voxel v (x fastest over the ghost-inclusive grid):
// the first element, in bin order and within a bin in ascending id, whose box
// holds the cell centre and whose Newton iteration lands inside; -1 if none.
// bin_start[b]..bin_start[b+1] index bin b's elements in bin_elems.
AMCA_HD inline int tag_voxel(const MeshView& M, const VoxelGrid& G, const int* bin_start,
                             const int* bin_elems, long long v, double tol, int max_iter,
                             double* pc) {
  const long long nxy = (long long)G.nxG * G.nyG;
  const int idx[3] = {(int)((v % nxy) % G.nxG), (int)((v % nxy) / G.nxG), (int)(v / nxy)};
  double cv[3];
  int mm[6];
  for (int q = 0; q < 3; q++) {
    cv[q] = G.org[q] + idx[q] * G.h;
    if (cv[q] < G.box[2 * q] || cv[q] > G.box[2 * q + 1]) return -1;
    const double dist = cv[q] - G.box[2 * q];
    mm[2 * q] = (int)floor(dist / G.dBin[q]);
    mm[2 * q + 1] = (int)ceil(dist / G.dBin[q]);
  }
  for (int q = 0; q < 3; q++) {
    int& a = mm[2 * q];
    int& b = mm[2 * q + 1];
    if (a == b && a > 0 && b < G.nBin[q]) {
      a -= 1;
      b += 1;
    } else if (b == 0) {
      b = 1;
    } else if (b == G.nBin[q]) {
      if (a > 0) a -= 1;
    }
  }
  for (int kk = mm[4]; kk < mm[5]; kk++)
    for (int jj = mm[2]; jj < mm[3]; jj++)
      for (int ii = mm[0]; ii < mm[1]; ii++) {
        if (ii < 0 || jj < 0 || kk < 0 || ii >= G.nBin[0] || jj >= G.nBin[1] || kk >= G.nBin[2])
          continue;
        const long long b = ii + jj * (long long)G.nBin[0] + kk * (long long)G.nBin[0] * G.nBin[1];
        for (int q = bin_start[b]; q < bin_start[b + 1]; q++) {
          const int e = bin_elems[q];
          double r[6];
          element_box(M, e, r);
          if (cv[0] < r[0] || cv[0] > r[1] || cv[1] < r[2] || cv[1] > r[3] || cv[2] < r[4] ||
              cv[2] > r[5])
            continue;
          if (in_element(M, e, cv, tol, max_iter, pc)) return e;
        }
      }
  return -1;
}

}  // namespace amca_vox
