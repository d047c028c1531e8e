#include "voxel_map.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <thread>

namespace amca_host {

void FeMesh::element_nodes(int e, int* nid) const {
  const int ncx = fnx - 1, ncy = fny - 1;
  const int ei = e % ncx, ej = (e / ncx) % ncy, ek = e / (ncx * ncy);
  const int nxy = fnx * fny;
  // LoadDataFromFile.cpp:110-115 node order
  nid[0] = nxy * ek + fnx * ej + ei;
  nid[1] = nid[0] + 1;
  nid[2] = nid[1] + fnx;
  nid[3] = nid[2] - 1;
  for (int m = 4; m < 8; m++) nid[m] = nid[m - 4] + nxy;
}

namespace {

// (MapVoxelManager.cpp:263-289) bounding box of an element: min/max per axis,
// with a < b ? a : b and a > b ? a : b
void element_box(const FeMesh& M, int e, double* r) {
  int nid[8];
  M.element_nodes(e, nid);
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

// Jacobian columns of the reference's Newton step (MapVoxelManager.cpp:538-617):
// entry (axis a, derivative d) = sum over nodes m = 1..7 of
// sign * (f1 * f2) * coord_a[m], accumulated in node order, where f1, f2 are
// the factors (1 +- s) of the two other parametric coordinates.
struct JTerm {
  signed char sign;
  signed char s1, s2;  // +1: (1 + u), -1: (1 - u)
};
// d = 0: factors (eta, zeta); d = 1: (xi, zeta); d = 2: (eta, xi)
const JTerm kJ[3][7] = {
    {{-1, -1, -1}, {-1, +1, -1}, {+1, +1, -1}, {+1, -1, +1}, {-1, -1, +1}, {-1, +1, +1}, {+1, +1, +1}},
    {{+1, +1, -1}, {-1, +1, -1}, {-1, -1, -1}, {+1, -1, +1}, {+1, +1, +1}, {-1, +1, +1}, {-1, -1, +1}},
    {{+1, -1, +1}, {+1, +1, +1}, {+1, +1, -1}, {-1, -1, -1}, {-1, -1, +1}, {-1, +1, +1}, {-1, +1, -1}},
};

inline double fac(int s, double u) { return s > 0 ? (1.0 + u) : (1.0 - u); }

double jac_entry(int d, const double* coord, double xi, double eta, double zeta) {
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
    const JTerm& t = kJ[d][m];
    const double term = (fac(t.s1, u1) * fac(t.s2, u2)) * coord[m + 1];
    if (m == 0) acc = t.sign > 0 ? term : -term;
    else acc = t.sign > 0 ? acc + term : acc - term;
  }
  return acc;
}

// MapVoxelManager::is_in_element (:487-699)
bool in_element(const FeMesh& M, int e, const double* pt, double tol, int max_iter, double* pc) {
  int nid[8];
  M.element_nodes(e, nid);
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
    double sf[8];
    const signed char se[8] = {-1, -1, +1, +1, -1, -1, +1, +1};
    const signed char sx[8] = {-1, +1, +1, -1, -1, +1, +1, -1};
    const signed char sz[8] = {-1, -1, -1, -1, +1, +1, +1, +1};
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
  } while (!(std::fabs(0.0 + diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2]) < tol) &&
           ++it < max_iter);
  double dist = std::numeric_limits<double>::max();
  for (int q = 0; q < 3; q++) pc[q] = std::numeric_limits<double>::max();
  if (it < max_iter) {
    double d = 0;
    for (int q = 0; q < 3; q++) {
      double v = nw[q];
      if (v > 1.0 && v < 1.000001) v = 1.0;
      if (v < -1.0 && v > -1.000001) v = -1.0;
      pc[q] = v;
      if (d < std::fabs(v)) d = std::fabs(v);
    }
    dist = d;
  }
  return !(dist > 1.0);
}

}  // namespace

bool map_voxels(const FeMesh& M, const VoxelMapParams& p, std::vector<int32_t>& tag,
                std::vector<double>& par, int threads) {
  const int nxG = p.nx + 2, nyG = p.ny + 2, nzG = p.nz + 2;
  const long long NG = (long long)nxG * nyG * nzG;
  tag.assign((size_t)NG, -1);
  par.assign(3 * (size_t)NG, -2.0);
  const double h = p.h;
  // get_CA_domain_xyz_min_max (:198-256) at one rank
  double box[6];
  box[0] = (0 - 1) * h + p.x0;
  box[2] = (0 - 1) * h + p.y0;
  box[4] = (0 - 1) * h + p.z0;
  box[1] = box[0] + nxG * h;
  box[3] = box[2] + nyG * h;
  box[5] = box[4] + nzG * h;
  double grow[6];
  std::copy(box, box + 6, grow);
  const int ne = M.elements();
  for (int e = 0; e < ne; e++) {
    int nid[8];
    M.element_nodes(e, nid);
    bool overlap = false;
    for (int m = 0; m < 8 && !overlap; m++) {
      const double* c = &M.xyz[3 * (size_t)nid[m]];
      overlap = c[0] >= box[0] && c[0] <= box[1] && c[1] >= box[2] && c[1] <= box[3] &&
                c[2] >= box[4] && c[2] <= box[5];
    }
    if (!overlap) continue;
    double r[6];
    element_box(M, e, r);
    for (int q = 0; q < 3; q++) {
      grow[2 * q] = (grow[2 * q] < r[2 * q]) ? grow[2 * q] : r[2 * q];
      grow[2 * q + 1] = (grow[2 * q + 1] > r[2 * q + 1]) ? grow[2 * q + 1] : r[2 * q + 1];
    }
  }
  std::copy(grow, grow + 6, box);
  double dBin[3];
  int nBin[3];
  for (int q = 0; q < 3; q++) {
    dBin[q] = p.bin[q] * h;
    nBin[q] = (int)std::ceil((box[2 * q + 1] - box[2 * q]) / dBin[q]);
  }
  // attach_elements (:294-352): elements wholly inside the box, per bin
  std::vector<std::vector<int>> bins((size_t)nBin[0] * nBin[1] * nBin[2]);
  for (int e = 0; e < ne; e++) {
    double r[6];
    element_box(M, e, r);
    bool inside = true;
    for (int q = 0; q < 3 && inside; q++)
      inside = !(r[2 * q] < box[2 * q] || r[2 * q + 1] > box[2 * q + 1]);
    if (!inside) continue;
    const int imin = (int)std::floor((r[0] - box[0]) / dBin[0]);
    const int imax = (int)std::ceil((r[1] - box[0]) / dBin[0]);
    const int jmin = (int)std::floor((r[2] - box[2]) / dBin[1]);
    const int jmax = (int)std::ceil((r[3] - box[2]) / dBin[1]);
    const int kmin = (int)std::floor((r[4] - box[4]) / dBin[2]);
    const int kmax = (int)std::ceil((r[5] - box[4]) / dBin[2]);
    for (int kk = kmin; kk < kmax; kk++)
      for (int jj = jmin; jj < jmax; jj++)
        for (int ii = imin; ii < imax; ii++) {
          if (ii < 0 || jj < 0 || kk < 0 || ii >= nBin[0] || jj >= nBin[1] || kk >= nBin[2]) continue;
          bins[ii + jj * (size_t)nBin[0] + kk * (size_t)nBin[0] * nBin[1]].push_back(e);
        }
  }
  // tag_voxels (:357-482), voxels split over threads
  const double org[3] = {(0 - 0.5) * h + p.x0, (0 - 0.5) * h + p.y0, (0 - 0.5) * h + p.z0};
  auto work = [&](long long v0, long long v1) {
    for (long long v = v0; v < v1; v++) {
      const int nxy = nxG * nyG;
      const int idx[3] = {(int)((v % nxy) % nxG), (int)((v % nxy) / nxG), (int)(v / nxy)};
      double cv[3];
      int mm[6];
      bool out = false;
      for (int q = 0; q < 3 && !out; q++) {
        cv[q] = org[q] + idx[q] * h;
        if (cv[q] < box[2 * q] || cv[q] > box[2 * q + 1]) {
          out = true;
          break;
        }
        const double dist = cv[q] - box[2 * q];
        mm[2 * q] = (int)std::floor(dist / dBin[q]);
        mm[2 * q + 1] = (int)std::ceil(dist / dBin[q]);
      }
      if (out) continue;
      for (int q = 0; q < 3; q++) {
        int& a = mm[2 * q];
        int& b = mm[2 * q + 1];
        if (a == b && a > 0 && b < nBin[q]) {
          a -= 1;
          b += 1;
        } else if (b == 0) {
          b = 1;
        } else if (b == nBin[q]) {
          if (a > 0) a -= 1;
        }
      }
      bool got = false;
      for (int kk = mm[4]; kk < mm[5] && !got; kk++)
        for (int jj = mm[2]; jj < mm[3] && !got; jj++)
          for (int ii = mm[0]; ii < mm[1] && !got; ii++) {
            if (ii < 0 || jj < 0 || kk < 0 || ii >= nBin[0] || jj >= nBin[1] || kk >= nBin[2]) continue;
            const auto& bl = bins[ii + jj * (size_t)nBin[0] + kk * (size_t)nBin[0] * nBin[1]];
            for (int e : bl) {
              double r[6];
              element_box(M, e, r);
              if (cv[0] < r[0] || cv[0] > r[1] || cv[1] < r[2] || cv[1] > r[3] || cv[2] < r[4] ||
                  cv[2] > r[5])
                continue;
              double pc[3];
              if (in_element(M, e, cv, p.tol, p.max_iter, pc)) {
                tag[(size_t)v] = e;
                for (int q = 0; q < 3; q++) par[3 * (size_t)v + q] = pc[q];
                got = true;
                break;
              }
            }
          }
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
