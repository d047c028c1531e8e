// ORACLE -- test infrastructure only.
//
// Sequential CPU restatement of AMCA3D's coupled CA-FE path at one MPI rank:
// CafeSolver::time_integration_weak_coupling (CafeSolver.cpp:178-219) driving a
// `cellular_automata_remelting` realm (CellularAutomataManager_3DRemelting.cpp)
// with temperatures interpolated from FE nodal frames (MapVoxelManager.cpp,
// LoadDataFromFile.cpp).  Frames are given in memory (the data of the
// reference's text frames after its offset/scale conversions).  Only the
// test-suite and the benchmark's CPU-baseline leg use it.
//
// Pinning: the FE parts are pinned by closed forms (hex8 interpolation of an
// affine field is exact, Newton on an axis-aligned hex gives the affine
// parametric coordinates, the nodal time interpolation is linear); the remelt
// sweep by hand cases and invariants in tests/test_oracle_remelt.py.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <list>
#include <random>
#include <stdexcept>
#include <vector>

#include "ca_common.h"
#include "oracle.h"

using namespace oracle_ca;

namespace {

struct Bin {
  bool has = false;
  std::vector<int> elements;
};

struct RemeltCA {
  oracle_remelt_params p;
  int nxG, nyG, nzG, nxyG, NG;
  std::vector<int> ori, gid;
  std::vector<double> nucDT, T;
  std::vector<char> cooling, melted, flag;
  std::vector<Grain> grains;
  std::vector<Orient> orients;
  std::list<int> sites;
  // FE provider (FiniteElementManager + LoadDataFromFile)
  int nNodes = 0, nEle = 0;
  std::vector<int> eleNodes;        // 8 per element
  std::vector<double> theta0, theta1, theta;  // node.theta_[0], [1], thetaArray_
  double inT[2] = {0.0, 0.0};
  double timeFE = 0.0, dTFE = 0.0;
  int nextFrame = 0;
  // MapVoxelManager
  std::vector<int> tag;
  std::vector<double> par;
  double xyz[6];
  // CA state
  double time = 0.0, dT = 0.0, maxDt = 0.0;
  int step = 0;
  double cutoff = 0.0;

  int cell(int i, int j, int k) const { return k * nxyG + j * nxG + i; }
  void ijk(int c, int& i, int& j, int& k) const {
    k = c / nxyG;
    i = (c % nxyG) % nxG;
    j = (c % nxyG) / nxG;
  }
  bool ghost(int c) const {
    int i, j, k;
    ijk(c, i, j, k);
    return i == 0 || i == nxG - 1 || j == 0 || j == nyG - 1 || k == 0 || k == nzG - 1;
  }
  void centre(int c, double* x) const {
    int i, j, k;
    ijk(c, i, j, k);
    x[0] = p.x0 + p.h * (i - 0.5);
    x[1] = p.y0 + p.h * (j - 0.5);
    x[2] = p.z0 + p.h * (k - 0.5);
  }
  void neighbours(int c, int* nb) const {
    int n = 0;
    for (int k = -1; k < 2; k++)
      for (int j = -1; j < 2; j++)
        for (int i = -1; i < 2; i++) {
          if (k == 0 && j == 0 && i == 0) continue;
          nb[n++] = c + k * nxyG + j * nxG + i;
        }
  }
  const double* node_xyz(int n) const { return p.node_xyz + 3 * (size_t)n; }

  // ProblemPhysics_RappazGandin1993::compute_growth_velocity_with_temperature_input
  // (ProblemPhysics.cpp:142-158)
  double velocity(double Tc) const {
    double dTu = p.Tm - Tc;
    double v = 0.0;
    if (dTu > 0) {
      v = p.a1 * pow(dTu, p.b1) + p.a2 * pow(dTu, p.b2) + p.a3 * pow(dTu, p.b3);
      if (v < 0.0) v = 0.0;
    }
    return v;
  }

  // ---- FE provider -------------------------------------------------------
  // load_mesh_and_first_two_steps_results_from_txt_file (LoadDataFromFile.cpp:21-137)
  // on in-memory frames, then FiniteElementManager::initialize (:147-189)
  void fe_setup() {
    nNodes = p.fnx * p.fny * p.fnz;
    int ncx = p.fnx - 1, ncy = p.fny - 1, ncz = p.fnz - 1;
    nEle = ncx * ncy * ncz;
    eleNodes.resize(8 * (size_t)nEle);
    int nxy = p.fnx * p.fny, e = 0;
    for (int k = 0; k < ncz; k++)
      for (int j = 0; j < ncy; j++)
        for (int i = 0; i < ncx; i++) {
          int* n = &eleNodes[8 * (size_t)e++];
          n[0] = nxy * k + p.fnx * j + i;
          n[1] = n[0] + 1;
          n[2] = n[1] + p.fnx;
          n[3] = n[2] - 1;
          for (int L = 4; L < 8; L++) n[L] = n[L - 4] + nxy;
        }
    theta0.assign(p.frame_theta, p.frame_theta + nNodes);
    theta1.assign(nNodes, 0.0);
    inT[0] = 0.0;
    nextFrame = 1;
    load_next();
    theta = theta1;  // thetaArray_ = theta_[1]
    dTFE = inT[0] - inT[1];  // get_output_time_step_from_txt_file
  }

  // load_next_step_nodal_data (LoadDataFromFile.cpp:142-191)
  bool load_next() {
    if (nextFrame >= p.n_frames) return false;
    inT[1] = inT[0];
    inT[0] = p.frame_time[nextFrame] - p.frame_time[0];
    const double* th = p.frame_theta + (size_t)nextFrame * nNodes;
    for (int n = 0; n < nNodes; n++) {
      theta1[n] = theta0[n];
      theta0[n] = th[n];
    }
    nextFrame++;
    return true;
  }

  // pseudo_single_integration_step (LoadDataFromFile.cpp:207-313); every node
  // is updated (the reference's grain-around subset holds every node the CA
  // cells read)
  bool fe_advance(double dt) {
    if (timeFE == 0) timeFE = inT[1];
    timeFE += dt;
    while (timeFE >= inT[0])
      if (!load_next()) return false;
    double r = (timeFE - inT[1]) / (inT[0] - inT[1]);
    if (r < 0.0) throw std::runtime_error("time ratio error");
    for (int n = 0; n < nNodes; n++) {
      if (theta0[n] <= p.void_temp) theta[n] = p.void_temp;
      else if (theta1[n] <= p.void_temp) theta[n] = theta0[n];
      else theta[n] = r * (theta0[n] - theta1[n]) + theta1[n];
    }
    return true;
  }

  // ---- MapVoxelManager ------------------------------------------------------
  void ele_range(int e, double* r) const {
    for (int q = 0; q < 3; q++) {
      r[2 * q] = 1.0e8;
      r[2 * q + 1] = -1.0e8;
    }
    for (int m = 0; m < 8; m++) {
      const double* c = node_xyz(eleNodes[8 * (size_t)e + m]);
      for (int q = 0; q < 3; q++) {
        r[2 * q] = ((r[2 * q]) < (c[q]) ? (r[2 * q]) : (c[q]));
        r[2 * q + 1] = ((r[2 * q + 1]) > (c[q]) ? (r[2 * q + 1]) : (c[q]));
      }
    }
  }

  // is_in_element (MapVoxelManager.cpp:487-699): Newton iteration for the
  // parametric coordinates of a point in a hex8 element.
  bool is_in_element(int e, const double* pt, double* pc) const {
    const int* nid = &eleNodes[8 * (size_t)e];
    double x[8], y[8], z[8];
    x[0] = y[0] = z[0] = 0.0;
    const double* o = node_xyz(nid[0]);
    for (int m = 1; m < 8; m++) {
      const double* c = node_xyz(nid[m]);
      x[m] = 0.125 * (c[0] - o[0]);
      y[m] = 0.125 * (c[1] - o[1]);
      z[m] = 0.125 * (c[2] - o[2]);
    }
    double xp = pt[0] - o[0], yp = pt[1] - o[1], zp = pt[2] - o[2];
    double j[9], f[3], sf[8];
    double xinew = 0.5, etanew = 0.5, zetanew = 0.5;
    double xi = 0.5, eta = 0.5, zeta = 0.5;
    double d[3] = {1.0, 1.0, 1.0};
    int it = 0;
    do {
      j[0] = -(1.0 - eta) * (1.0 - zeta) * x[1] - (1.0 + eta) * (1.0 - zeta) * x[2] +
             (1.0 + eta) * (1.0 - zeta) * x[3] + (1.0 - eta) * (1.0 + zeta) * x[4] -
             (1.0 - eta) * (1.0 + zeta) * x[5] - (1.0 + eta) * (1.0 + zeta) * x[6] +
             (1.0 + eta) * (1.0 + zeta) * x[7];
      j[1] = (1.0 + xi) * (1.0 - zeta) * x[1] - (1.0 + xi) * (1.0 - zeta) * x[2] -
             (1.0 - xi) * (1.0 - zeta) * x[3] + (1.0 - xi) * (1.0 + zeta) * x[4] +
             (1.0 + xi) * (1.0 + zeta) * x[5] - (1.0 + xi) * (1.0 + zeta) * x[6] -
             (1.0 - xi) * (1.0 + zeta) * x[7];
      j[2] = (1.0 - eta) * (1.0 + xi) * x[1] + (1.0 + eta) * (1.0 + xi) * x[2] +
             (1.0 + eta) * (1.0 - xi) * x[3] - (1.0 - eta) * (1.0 - xi) * x[4] -
             (1.0 - eta) * (1.0 + xi) * x[5] - (1.0 + eta) * (1.0 + xi) * x[6] -
             (1.0 + eta) * (1.0 - xi) * x[7];
      j[3] = -(1.0 - eta) * (1.0 - zeta) * y[1] - (1.0 + eta) * (1.0 - zeta) * y[2] +
             (1.0 + eta) * (1.0 - zeta) * y[3] + (1.0 - eta) * (1.0 + zeta) * y[4] -
             (1.0 - eta) * (1.0 + zeta) * y[5] - (1.0 + eta) * (1.0 + zeta) * y[6] +
             (1.0 + eta) * (1.0 + zeta) * y[7];
      j[4] = (1.0 + xi) * (1.0 - zeta) * y[1] - (1.0 + xi) * (1.0 - zeta) * y[2] -
             (1.0 - xi) * (1.0 - zeta) * y[3] + (1.0 - xi) * (1.0 + zeta) * y[4] +
             (1.0 + xi) * (1.0 + zeta) * y[5] - (1.0 + xi) * (1.0 + zeta) * y[6] -
             (1.0 - xi) * (1.0 + zeta) * y[7];
      j[5] = (1.0 - eta) * (1.0 + xi) * y[1] + (1.0 + eta) * (1.0 + xi) * y[2] +
             (1.0 + eta) * (1.0 - xi) * y[3] - (1.0 - eta) * (1.0 - xi) * y[4] -
             (1.0 - eta) * (1.0 + xi) * y[5] - (1.0 + eta) * (1.0 + xi) * y[6] -
             (1.0 + eta) * (1.0 - xi) * y[7];
      j[6] = -(1.0 - eta) * (1.0 - zeta) * z[1] - (1.0 + eta) * (1.0 - zeta) * z[2] +
             (1.0 + eta) * (1.0 - zeta) * z[3] + (1.0 - eta) * (1.0 + zeta) * z[4] -
             (1.0 - eta) * (1.0 + zeta) * z[5] - (1.0 + eta) * (1.0 + zeta) * z[6] +
             (1.0 + eta) * (1.0 + zeta) * z[7];
      j[7] = (1.0 + xi) * (1.0 - zeta) * z[1] - (1.0 + xi) * (1.0 - zeta) * z[2] -
             (1.0 - xi) * (1.0 - zeta) * z[3] + (1.0 - xi) * (1.0 + zeta) * z[4] +
             (1.0 + xi) * (1.0 + zeta) * z[5] - (1.0 + xi) * (1.0 + zeta) * z[6] -
             (1.0 - xi) * (1.0 + zeta) * z[7];
      j[8] = (1.0 - eta) * (1.0 + xi) * z[1] + (1.0 + eta) * (1.0 + xi) * z[2] +
             (1.0 + eta) * (1.0 - xi) * z[3] - (1.0 - eta) * (1.0 - xi) * z[4] -
             (1.0 - eta) * (1.0 + xi) * z[5] - (1.0 + eta) * (1.0 + xi) * z[6] -
             (1.0 + eta) * (1.0 - xi) * z[7];
      double jdet = -(j[2] * j[4] * j[6]) + j[1] * j[5] * j[6] + j[2] * j[3] * j[7] -
                    j[0] * j[5] * j[7] - j[1] * j[3] * j[8] + j[0] * j[4] * j[8];
      if (!jdet) {
        it = p.max_iter;
        break;
      }
      sf[0] = (1.0 - eta) * (1.0 - xi) * (1.0 - zeta);
      sf[1] = (1.0 - eta) * (1.0 + xi) * (1.0 - zeta);
      sf[2] = (1.0 + eta) * (1.0 + xi) * (1.0 - zeta);
      sf[3] = (1.0 + eta) * (1.0 - xi) * (1.0 - zeta);
      sf[4] = (1.0 - eta) * (1.0 - xi) * (1.0 + zeta);
      sf[5] = (1.0 - eta) * (1.0 + xi) * (1.0 + zeta);
      sf[6] = (1.0 + eta) * (1.0 + xi) * (1.0 + zeta);
      sf[7] = (1.0 + eta) * (1.0 - xi) * (1.0 + zeta);
      f[0] = xp - sf[1] * x[1] - sf[2] * x[2] - sf[3] * x[3] - sf[4] * x[4] - sf[5] * x[5] -
             sf[6] * x[6] - sf[7] * x[7];
      f[1] = yp - sf[1] * y[1] - sf[2] * y[2] - sf[3] * y[3] - sf[4] * y[4] - sf[5] * y[5] -
             sf[6] * y[6] - sf[7] * y[7];
      f[2] = zp - sf[1] * z[1] - sf[2] * z[2] - sf[3] * z[3] - sf[4] * z[4] - sf[5] * z[5] -
             sf[6] * z[6] - sf[7] * z[7];
      xinew = (jdet * xi + f[2] * (j[2] * j[4] - j[1] * j[5]) + f[1] * (j[1] * j[8] - j[2] * j[7]) +
               f[0] * (j[5] * j[7] - j[4] * j[8])) / jdet;
      etanew = (eta * jdet + f[2] * (j[0] * j[5] - j[2] * j[3]) + f[1] * (j[2] * j[6] - j[0] * j[8]) +
                f[0] * (j[3] * j[8] - j[5] * j[6])) / jdet;
      zetanew = (jdet * zeta + f[2] * (j[1] * j[3] - j[0] * j[4]) + f[1] * (j[0] * j[7] - j[1] * j[6]) +
                 f[0] * (j[4] * j[6] - j[3] * j[7])) / jdet;
      d[0] = xinew - xi;
      d[1] = etanew - eta;
      d[2] = zetanew - zeta;
      xi = xinew;
      eta = etanew;
      zeta = zetanew;
    } while (!(fabs(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) < p.tol) && ++it < p.max_iter);
    pc[0] = pc[1] = pc[2] = std::numeric_limits<double>::max();
    double dist = std::numeric_limits<double>::max();
    if (it < p.max_iter) {
      pc[0] = xinew;
      pc[1] = etanew;
      pc[2] = zetanew;
      for (int q = 0; q < 3; q++) {
        if (pc[q] > 1.0 && pc[q] < 1.000001) pc[q] = 1.0;
        if (pc[q] < -1.0 && pc[q] > -1.000001) pc[q] = -1.0;
      }
      double dd = 0;
      for (int q = 0; q < 3; q++)
        if (dd < fabs(pc[q])) dd = fabs(pc[q]);
      dist = dd;
    }
    return !(dist > 1.0);
  }

  // execute_with_focus_on_CA_region (MapVoxelManager.cpp:140-164)
  void map_voxels() {
    tag.assign(NG, -1);
    par.assign(3 * (size_t)NG, -2.0);
    const double h = p.h;
    // get_CA_domain_xyz_min_max (:198-256), xStart = 0
    xyz[0] = (0 - 1) * h + p.x0;
    xyz[2] = (0 - 1) * h + p.y0;
    xyz[4] = (0 - 1) * h + p.z0;
    xyz[1] = xyz[0] + nxG * h;
    xyz[3] = xyz[2] + nyG * h;
    xyz[5] = xyz[4] + nzG * h;
    double cur[6];
    for (int q = 0; q < 6; q++) cur[q] = xyz[q];
    for (int e = 0; e < nEle; e++) {
      bool overlap = false;
      for (int m = 0; m < 8; m++) {
        const double* c = node_xyz(eleNodes[8 * (size_t)e + m]);
        if (c[0] >= xyz[0] && c[0] <= xyz[1] && c[1] >= xyz[2] && c[1] <= xyz[3] &&
            c[2] >= xyz[4] && c[2] <= xyz[5]) {
          overlap = true;
          break;
        }
      }
      if (overlap) {
        double r[6];
        ele_range(e, r);
        for (int q = 0; q < 3; q++) {
          cur[2 * q] = ((cur[2 * q]) < (r[2 * q]) ? (cur[2 * q]) : (r[2 * q]));
          cur[2 * q + 1] = ((cur[2 * q + 1]) > (r[2 * q + 1]) ? (cur[2 * q + 1]) : (r[2 * q + 1]));
        }
      }
    }
    for (int q = 0; q < 6; q++) xyz[q] = cur[q];
    double dBin[3];
    int nBin[3];
    for (int q = 0; q < 3; q++) {
      dBin[q] = p.bin[q] * h;
      nBin[q] = ceil((xyz[2 * q + 1] - xyz[2 * q]) / dBin[q]);
    }
    std::vector<Bin> bins((size_t)nBin[0] * nBin[1] * nBin[2]);
    // attach_elements (:294-352)
    for (int e = 0; e < nEle; e++) {
      double r[6];
      ele_range(e, r);
      bool inside = true;
      for (int q = 0; q < 3; q++)
        if (r[2 * q] < xyz[2 * q] || r[2 * q + 1] > xyz[2 * q + 1]) {
          inside = false;
          break;
        }
      if (!inside) continue;
      int imin = std::floor((r[0] - xyz[0]) / dBin[0]);
      int imax = std::ceil((r[1] - xyz[0]) / dBin[0]);
      int jmin = std::floor((r[2] - xyz[2]) / dBin[1]);
      int jmax = std::ceil((r[3] - xyz[2]) / dBin[1]);
      int kmin = std::floor((r[4] - xyz[4]) / dBin[2]);
      int kmax = std::ceil((r[5] - xyz[4]) / dBin[2]);
      for (int kk = kmin; kk < kmax; kk++)
        for (int jj = jmin; jj < jmax; jj++)
          for (int ii = imin; ii < imax; ii++) {
            Bin& b = bins[ii + jj * nBin[0] + (size_t)kk * nBin[0] * nBin[1]];
            b.has = true;
            b.elements.push_back(e);
          }
    }
    // tag_voxels (:357-482)
    double org[3] = {(0 - 0.5) * h + p.x0, (0 - 0.5) * h + p.y0, (0 - 0.5) * h + p.z0};
    for (int v = 0; v < NG; v++) {
      int idx[3];
      idx[2] = v / nxyG;
      idx[1] = (v % nxyG) / nxG;
      idx[0] = (v % nxyG) % nxG;
      double cv[3];
      int mm[6];
      bool out = false;
      for (int q = 0; q < 3; q++) {
        cv[q] = org[q] + idx[q] * h;
        if (cv[q] < xyz[2 * q] || cv[q] > xyz[2 * q + 1]) {
          out = true;
          break;
        }
        double dist = cv[q] - xyz[2 * q];
        mm[2 * q] = std::floor(dist / dBin[q]);
        mm[2 * q + 1] = std::ceil(dist / dBin[q]);
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
            const Bin& b = bins[ii + jj * nBin[0] + (size_t)kk * nBin[0] * nBin[1]];
            if (!b.has) continue;
            for (int e : b.elements) {
              double r[6];
              ele_range(e, r);
              if (cv[0] < r[0] || cv[0] > r[1] || cv[1] < r[2] || cv[1] > r[3] || cv[2] < r[4] ||
                  cv[2] > r[5])
                continue;
              double pc[3];
              if (is_in_element(e, cv, pc)) {
                tag[v] = e;
                for (int q = 0; q < 3; q++) par[3 * (size_t)v + q] = pc[q];
                got = true;
              }
              if (got) break;
            }
          }
    }
  }

  // evaluate_temperature (MapVoxelManager.cpp:746-799); every element is born
  bool eval_T(int v, double& out) const {
    double nm = p.void_temp + 0.00001;
    if (tag[v] < 0) return false;
    const int* nid = &eleNodes[8 * (size_t)tag[v]];
    for (int m = 0; m < 8; m++)
      if (theta[nid[m]] < nm) return false;
    double xi = par[3 * (size_t)v], eta = par[3 * (size_t)v + 1], zeta = par[3 * (size_t)v + 2];
    double sf[8];
    sf[0] = 0.125 * (1.0 - eta) * (1.0 - xi) * (1.0 - zeta);
    sf[1] = 0.125 * (1.0 - eta) * (1.0 + xi) * (1.0 - zeta);
    sf[2] = 0.125 * (1.0 + eta) * (1.0 + xi) * (1.0 - zeta);
    sf[3] = 0.125 * (1.0 + eta) * (1.0 - xi) * (1.0 - zeta);
    sf[4] = 0.125 * (1.0 - eta) * (1.0 - xi) * (1.0 + zeta);
    sf[5] = 0.125 * (1.0 - eta) * (1.0 + xi) * (1.0 + zeta);
    sf[6] = 0.125 * (1.0 + eta) * (1.0 + xi) * (1.0 + zeta);
    sf[7] = 0.125 * (1.0 + eta) * (1.0 - xi) * (1.0 + zeta);
    double th = 0.0;
    for (int m = 0; m < 8; m++) th += sf[m] * theta[nid[m]];
    out = th;
    return true;
  }

  // ---- remelting CA ------------------------------------------------------
  void setup() {
    nxG = p.nx + 2;
    nyG = p.ny + 2;
    nzG = p.nz + 2;
    nxyG = nxG * nyG;
    NG = nxyG * nzG;
    ori.assign(NG, -1);
    gid.assign(NG, -1);
    nucDT.assign(NG, -1.0);
    T.assign(NG, -1.0);  // nullMatTemp_ (CellularAutomataManager_3DRemelting.cpp:29, 96-105)
    cooling.assign(NG, 0);
    melted.assign(NG, p.initial_melted ? 1 : 0);
    flag.assign(NG, 0);
    cutoff = p.cutoff_factor * p.h;  // :89
  }

  // load_microstructure_information (:165-307) from in-memory arrays, then
  // retrieve_grain_information_before_simulation (:312-332)
  void load_state() {
    for (int c = 0; c < NG; c++) {
      gid[c] = p.init_gid[c];
      ori[c] = p.init_ori[c];
      melted[c] = p.init_melted[c] ? 1 : 0;
    }
    for (int o = 0; o < p.n_init_orient; o++) {
      Orient r;
      for (int q = 0; q < 3; q++) r.euler[q] = p.init_euler[3 * o + q];
      rotation_from_euler(r.euler, r.l2g);
      orients.push_back(r);
    }
    for (int g = 0; g < p.n_init_grains; g++) {
      Grain gr;
      gr.id = g;
      gr.cell = p.init_grain_cell[g];
      double x[3];
      centre(gr.cell, x);
      gr.xc = x[0];
      gr.yc = x[1];
      gr.zc = x[2];
      gr.orient = p.init_grain_orient[g];
      gr.length = 0.0;
      gr.active = false;
      gr.wave = 0;
      grains.push_back(gr);
    }
    for (int c = 0; c < NG; c++) {
      int i, j, k;
      ijk(c, i, j, k);
      if (i == 0 || i == nxG - 1 || j == 0 || j == nyG - 1 || k == 0 || k == nzG - 1) {
        ori[c] = -1;  // global ghost
        continue;
      }
      if (ori[c] > -1) continue;
      if (gid[c] < 0 || gid[c] >= (int)grains.size())
        throw std::runtime_error("a cell without orientation has no grain record");
      ori[c] = grains[gid[c]].orient;
    }
  }

  // random_microstructure_information (CellularAutomataManager_3D.cpp:1928-1984):
  // every interior cell, in cell order, gets a new orientation (three rand()
  // draws, get_random_orientation_angle :838-850) and its own grain at the
  // cell centre; ids in that order
  void random_micro() {
    for (int c = 0; c < NG; c++) {
      if (ghost(c)) {
        gid[c] = -1;
        ori[c] = -1;
        continue;
      }
      Orient o;
      double U = static_cast<double>(rand()) / RAND_MAX;
      double V = static_cast<double>(rand()) / RAND_MAX;
      o.euler[0] = U * 2 * kPi;
      o.euler[1] = acos(2 * V - 1);
      o.euler[2] = static_cast<double>(rand()) / RAND_MAX * 0.5 * kPi;
      rotation_from_euler(o.euler, o.l2g);
      const int oid = (int)orients.size();
      orients.push_back(o);
      Grain g;
      g.id = (int)grains.size();
      g.cell = c;
      double x[3];
      centre(c, x);
      g.xc = x[0];
      g.yc = x[1];
      g.zc = x[2];
      g.orient = oid;
      g.length = 0.0;
      g.active = false;
      g.wave = 0;
      grains.push_back(g);
      gid[c] = g.id;
      ori[c] = oid;
    }
  }

  // load_microstructure_information with load_seeds (CellularAutomataManager_3DRemelting.cpp:244-249):
  // every cell whose seed undercooling is > -1 joins the site list, in cell order
  void load_seeds(const double* dt) {
    for (int c = 0; c < NG; c++)
      if (dt[c] > -1) {
        nucDT[c] = dt[c];
        sites.push_back(c);
      }
  }

  void create_sites(double rho, double mean, double sd, bool surface) {
    // CellularAutomataManager::create_nucleation_sites (CellularAutomataManager.cpp:153-213)
    std::vector<int> ids;
    if (surface) {
      for (int c = 0; c < NG; ++c) {
        int K = c / nxyG;
        if (K == 0 || K == nzG - 1) continue;
        int I = (c % nxyG) % nxG;
        if (I == 0 || I == nxG - 1) continue;
        int J = (c % nxyG) / nxG;
        if (J == 0 || J == nxG - 1) continue;
        if (K == 1 || K == p.nz)
          if (I != 0 && I != p.nx + 1 && J != 0 && J != p.ny + 1) ids.push_back(c);
        if (I == 1 || I == p.nx)
          if (J != 0 && J != p.ny + 1 && K != 0 && K != p.nz + 1) ids.push_back(c);
        if (J == 1 || J == p.ny)
          if (I != 0 && I != p.nx + 1 && K != 0 && K != p.nz + 1) ids.push_back(c);
      }
    } else {
      for (int k = 1; k < nzG - 1; ++k)
        for (int j = 1; j < nyG - 1; ++j)
          for (int i = 1; i < nxG - 1; ++i) ids.push_back(nxyG * k + nxG * j + i);
    }
    double vpc = surface ? p.h * p.h : p.h * p.h * p.h;
    int numCells = (int)ids.size();
    int numSites = round(vpc * rho * numCells) + 0.5;
    std::default_random_engine gen;
    std::normal_distribution<double> dist(mean, sd);
    for (int s = 0; s < numSites; ++s) {
      int c = ids[rand() % numCells];
      double dTc = dist(gen);
      dTc = (dTc < 0) ? 0 : dTc;
      double cur = nucDT[c];
      if (cur < 0) {
        sites.push_back(c);
        nucDT[c] = dTc;
      } else if (dTc < cur) {
        nucDT[c] = dTc;
      }
    }
    sites.sort();
  }

  // checkup_current_temperature_to_update_phase_state (:338-404)
  void phase_update() {
    for (int c = 0; c < NG; c++) flag[c] = 0;
    for (int i = 0; i < nxG; i++)
      for (int j = 0; j < nyG; j++)
        for (int k = 0; k < nzG; k++) {
          int c = cell(i, j, k);
          double Tc = -1.0;
          bool withMat = eval_T(c, Tc);
          if (step > 0) {
            double rate = 410;
            if (T[c] > 0) rate = (Tc - T[c]) / dT;
            if (Tc <= 0) rate = 0;
            cooling[c] = withMat ? (rate < 0) : 0;
          }
          T[c] = Tc;
          if (i == 0 || i == nxG - 1 || j == 0 || j == nyG - 1 || k == 0 || k == nzG - 1) continue;
          if (withMat && Tc < p.Tm) continue;
          ori[c] = -1;
          melted[c] = withMat ? 1 : 0;
          if (gid[c] > -1) grains[gid[c]].active = false;
        }
  }

  // checkup_liquid_cell_neigh_to_growth (:485-542)
  void liquid_neigh(double tcafe) {
    int nb[26];
    for (int c = 0; c < NG; c++) flag[c] = 0;
    for (int i = 1; i < nxG - 1; i++)
      for (int j = 1; j < nyG - 1; j++)
        for (int k = 1; k < nzG - 1; k++) {
          int c = cell(i, j, k);
          double Tc = T[c];
          bool withMat = true;
          if (Tc < -1.0 + 0.01) withMat = false;
          if (ori[c] == -1 && withMat && tcafe >= maxDt) melted[c] = 1;
          if (!withMat) {
            melted[c] = 0;
            continue;
          }
          if (ori[c] != -1 || ghost(c)) continue;
          if (Tc >= p.Tm) continue;
          neighbours(c, nb);
          for (int s = 0; s < 26; s++)
            if (ori[nb[s]] > -1) flag[nb[s]] = 1;  // c is never a global ghost here
        }
  }

  // nucleate_at_nucleation_sites, remelting variant (:936-999)
  int nucleate() {
    int made = 0;
    for (int c : sites) {
      bool liquid = (ori[c] == -1);
      bool beenMelted = melted[c];
      bool mushy = (T[c] > p.Tmin_nuc);
      if (cooling[c] && liquid && beenMelted && mushy) {
        double dTu = p.Tm - T[c];
        if (dTu > nucDT[c]) {
          Orient o;
          double U = static_cast<double>(rand()) / RAND_MAX;
          double V = static_cast<double>(rand()) / RAND_MAX;
          o.euler[0] = U * 2 * kPi;
          o.euler[1] = acos(2 * V - 1);
          o.euler[2] = static_cast<double>(rand()) / RAND_MAX * 0.5 * kPi;
          rotation_from_euler(o.euler, o.l2g);
          int oid = (int)orients.size();
          orients.push_back(o);
          double x[3];
          centre(c, x);
          if (gid[c] == -1) {
            Grain g;
            g.id = (int)grains.size();
            g.cell = c;
            g.xc = x[0];
            g.yc = x[1];
            g.zc = x[2];
            g.orient = oid;
            g.length = 0;
            g.active = true;
            g.wave = 0;
            grains.push_back(g);
            gid[c] = g.id;
          } else {
            Grain& g = grains[gid[c]];
            g.active = true;
            g.length = 0.0;
            g.xc = x[0];
            g.yc = x[1];
            g.zc = x[2];
            g.orient = oid;
          }
          ori[c] = oid;
          made++;
        }
      }
    }
    return made;
  }

  // calculate_time_step, remelting variant (:656-727)
  double time_step(int* nuc) {
    *nuc = p.use_sites ? nucleate() : 0;
    double minT = p.Tm;
    for (int i = 1; i < nxG - 1; i++)
      for (int j = 1; j < nyG - 1; j++)
        for (int k = 1; k < nzG - 1; k++) {
          int c = cell(i, j, k);
          if (!flag[c] || ghost(c)) continue;
          if (gid[c] < 0) continue;
          double Tc = T[c];
          Grain& g = grains[gid[c]];
          if (!(g.length < cutoff)) g.length = p.h;
          if (Tc <= p.Ts) continue;
          if (minT > Tc) minT = Tc;
        }
    double maxV = (velocity(minT) > 0 ? velocity(minT) : 0);
    return std::min(p.tsf * p.h / maxV, maxDt);
  }

  // single_integration_step, remelting variant (:544-652) with capture_cell (:807-847)
  int sweep(int* fresh) {
    int nb[26];
    int caps = 0;
    *fresh = 0;
    for (int i = 1; i < nxG - 1; i++)
      for (int j = 1; j < nyG - 1; j++)
        for (int k = 1; k < nzG - 1; k++) {
          int c = cell(i, j, k);
          if (!flag[c] || ghost(c)) continue;
          if (gid[c] < 0) continue;
          Grain& g = grains[gid[c]];
          double Tc = T[c];
          double vel = (Tc < p.Ts) ? velocity(p.Ts + 0.01) : velocity(Tc);
          g.length += dT * vel;
          neighbours(c, nb);
          const double* l2g = orients[g.orient].l2g;
          for (int s = 0; s < 26; s++) {
            int n = nb[s];
            double Tn = T[n];
            if (ori[n] > -1) continue;
            if (Tn > p.Tm) continue;
            if (!(Tn > -1.0)) continue;
            if (ghost(n)) continue;
            double cc[3];
            centre(n, cc);
            Grain par = grains[gid[c]];
            if (!capture_test(par, l2g, cc)) continue;
            double gc[3] = {cc[0] - par.xc, cc[1] - par.yc, cc[2] - par.zc};
            double info[4];
            child_envelope(gc, par.length, p.h, p.trunc, l2g, info);
            double nc[3] = {info[0] + par.xc, info[1] + par.yc, info[2] + par.zc};
            if (gid[n] < 0) {
              Grain ng;
              ng.id = (int)grains.size();
              ng.cell = n;
              ng.xc = nc[0];
              ng.yc = nc[1];
              ng.zc = nc[2];
              ng.orient = par.orient;
              ng.length = info[3];
              ng.active = true;
              ng.wave = 0;
              grains.push_back(ng);
              gid[n] = ng.id;
              (*fresh)++;
            } else {
              Grain& og = grains[gid[n]];
              og.active = true;
              og.length = info[3];
              og.xc = nc[0];
              og.yc = nc[1];
              og.zc = nc[2];
              og.orient = par.orient;
            }
            ori[n] = par.orient;
            caps++;
          }
        }
    time += dT;
    step++;
    return caps;
  }
};

}  // namespace

extern "C" int oracle_remelt_run(const oracle_remelt_params* prm, oracle_remelt_out* out) {
  RemeltCA ca;
  ca.p = *prm;
  ca.setup();
  srand(prm->seed);
  try {
    ca.fe_setup();
    ca.map_voxels();
    if (prm->random_micro)
      ca.random_micro();
    else if (prm->init_ori)
      ca.load_state();
    if (prm->init_site_dt) ca.load_seeds(prm->init_site_dt);
    if (prm->use_sites) {
      ca.create_sites(prm->ns, prm->dTs_mean, prm->dTs_sd, true);
      ca.create_sites(prm->nv, prm->dTv_mean, prm->dTv_sd, false);
    }
    // CafeSolver::time_integration_weak_coupling (CafeSolver.cpp:178-219)
    double dTFE = ca.dTFE;
    ca.maxDt = dTFE;
    double tcafe = 0.0;
    int nrec = 0;
    while (tcafe < prm->final_time) {
      if (prm->max_steps >= 0 && ca.step >= prm->max_steps) break;
      ca.phase_update();
      ca.liquid_neigh(tcafe);
      int nuc = 0;
      double dTCA = ca.time_step(&nuc);
      double dT = ((dTFE) < (dTCA) ? (dTFE) : (dTCA));
      ca.dT = dT;
      double t0 = ca.time;
      int fresh = 0;
      int caps = ca.sweep(&fresh);
      tcafe += dT;
      if (out->steps && nrec < out->steps_cap) {
        oracle_remelt_step* r = &out->steps[nrec++];
        r->time_before = t0;
        r->dt = dT;
        r->n_nucleated = nuc;
        r->n_captured = caps;
        r->n_fresh = fresh;
        int fl = 0, liq = 0;
        for (int c = 0; c < ca.NG; c++) {
          fl += ca.flag[c] ? 1 : 0;
          if (!ca.ghost(c) && ca.ori[c] < 0) liq++;
        }
        r->n_flagged = fl;
        r->n_liquid = liq;
      }
      if (!ca.fe_advance(dT)) {
        out->frames_exhausted = 1;
        break;
      }
      dTFE = ca.inT[0] - ca.timeFE;  // get_max_time_step
      ca.maxDt = dTFE;
    }
    out->n_records = nrec;
  } catch (const std::exception&) {
    return 2;
  }
  out->n_steps = ca.step;
  out->time = ca.time;
  out->time_fe = ca.timeFE;
  out->n_grains = (int)ca.grains.size();
  out->n_orient = (int)ca.orients.size();
  const size_t NG = (size_t)ca.NG;
  if (out->ori) memcpy(out->ori, ca.ori.data(), sizeof(int32_t) * NG);
  if (out->gid) memcpy(out->gid, ca.gid.data(), sizeof(int32_t) * NG);
  if (out->T) memcpy(out->T, ca.T.data(), sizeof(double) * NG);
  if (out->melted)
    for (size_t c = 0; c < NG; c++) out->melted[c] = ca.melted[c];
  if (out->cooling)
    for (size_t c = 0; c < NG; c++) out->cooling[c] = ca.cooling[c];
  if (out->tag) memcpy(out->tag, ca.tag.data(), sizeof(int32_t) * NG);
  if (out->par) memcpy(out->par, ca.par.data(), sizeof(double) * 3 * NG);
  int ng = std::min<int>(out->grain_cap, (int)ca.grains.size());
  for (int g = 0; g < ng; g++) {
    const Grain& gr = ca.grains[g];
    if (out->grain_env) {
      out->grain_env[4 * g + 0] = gr.xc;
      out->grain_env[4 * g + 1] = gr.yc;
      out->grain_env[4 * g + 2] = gr.zc;
      out->grain_env[4 * g + 3] = gr.length;
    }
    if (out->grain_cell) out->grain_cell[g] = gr.cell;
    if (out->grain_orient) out->grain_orient[g] = gr.orient;
  }
  int no = std::min<int>(out->orient_cap, (int)ca.orients.size());
  for (int o = 0; o < no; o++)
    if (out->euler) memcpy(out->euler + 3 * o, ca.orients[o].euler, sizeof(double) * 3);
  if (out->theta) memcpy(out->theta, ca.theta.data(), sizeof(double) * ca.nNodes);
  out->n_sites = (int)ca.sites.size();
  int ns = 0;
  for (int c : ca.sites) {
    if (ns >= out->site_cap) break;
    if (out->site_cell) out->site_cell[ns] = c;
    if (out->site_dt) out->site_dt[ns] = ca.nucDT[c];
    ns++;
  }
  return 0;
}
