// ORACLE -- test infrastructure only.
//
// A sequential CPU restatement of AMCA3D's plain 3-D cellular-automaton path at
// one MPI rank (realm type `cellular_automata`).  It exists to check the HIP
// implementation: only the test-suite, the smoke check and the CPU-baseline leg
// of the benchmark load it.  Nothing in the product links or calls it.
//
// Each function cites the reference file:line it restates.  The arithmetic is
// written in the same operation order as the reference, and the file is built
// with -O2 -ffp-contract=off (no FMA contraction, SSE2 doubles), which is what a
// default g++ build of the reference does on x86-64.
//
// Pinning (what checks each part independently of this file):
//   * RNG draws:     glibc rand() and libstdc++ minstd_rand0/normal_distribution
//                    known-answer values (tests/test_oracle_kat.py).
//   * rotation:      orthonormality and det = +1 (closed form).
//   * capture test:  brute-force L1-ball membership against an independent
//                    numpy evaluation (tests/test_oracle_geometry.py).
//   * child envelope: containment in the parent octahedron, length bound
//                    sqrt(2)*trunc*h, shift along the chosen corner.
//   * time stepping: closed-form first nucleation step and dt sequence.
//   * capture sweep: symmetry of a single nucleus with identity orientation
//                    under the 48 cube symmetries; brute-force list semantics on
//                    tiny grids (tests/test_oracle_plain.py).
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <list>
#include <random>
#include <vector>
#include <algorithm>

#include "oracle.h"
#include "ca_common.h"

namespace {

using namespace oracle_ca;

struct PlainCA {
  oracle_plain_params p;
  int nxG, nyG, nzG, nxyG, NG;
  std::vector<int> ori, gid;
  std::vector<double> nucDT;
  std::vector<Grain> grains;
  std::vector<Orient> orients;
  std::list<int> active;
  std::list<int> sites;
  double time = 0.0, dT = 0.0;
  int step = 0;
  bool done = false;
  long long siteDraws = 0;
  int nSurface = 0, nBulk = 0;
  // explicit orientation rows (angles) consumed in nucleation order instead of
  // rand() draws; used for hand cases and device-only parity checks
  const double* givenEuler = nullptr;
  int givenCount = 0;

  void centre(int c, double* x) const {
    // CellularAutomataManager_3D::get_cell_center (:876-888), xStart = 0 at np=1
    int k = c / nxyG;
    int i = (c % nxyG) % nxG;
    int j = (c % nxyG) / nxG;
    x[0] = p.x0 + p.h * (i - 0.5);
    x[1] = p.y0 + p.h * (j - 0.5);
    x[2] = p.z0 + p.h * (k - 0.5);
  }

  void neighbours(int c, int* nb) const {
    // CellularAutomataManager_3D::get_cell_neighbors (:893-929), 26-stencil k,j,i
    int n = 0;
    for (int k = -1; k < 2; k++)
      for (int j = -1; j < 2; j++)
        for (int i = -1; i < 2; i++) {
          if (k == 0 && j == 0 && i == 0) continue;
          nb[n++] = c + k * nxyG + j * nxG + i;
        }
  }

  double velocity(double t) const {
    // ProblemPhysics_RappazGandin1993::compute_growth_velocity (ProblemPhysics.cpp:117-135)
    double T = p.T0 + p.Tdot * t;
    double dT = p.Tm - T;
    double v = 0.0;
    if (dT > 0) {
      v = p.a1 * dT + p.a2 * dT * dT + p.a3 * dT * dT * dT;
      if (v < 0.0) v = 0.0;
    }
    return v;
  }

  void setup() {
    // setup_cells / allocate_arrays / initialize_cells / initialize_boundaries
    // (CellularAutomataManager_3D.cpp:404-435, 735-833; CellularAutomataManager.cpp:303-326)
    nxG = p.nx + 2; nyG = p.ny + 2; nzG = p.nz + 2;
    nxyG = nxG * nyG;
    NG = nxyG * nzG;
    ori.assign(NG, -1);
    gid.assign(NG, -1);
    nucDT.assign(NG, -1.0);
    for (int k = 0; k < nzG; k++)
      for (int j = 0; j < nyG; j++)
        for (int i = 0; i < nxG; i++) {
          if (i == 0 || j == 0 || k == 0 || i == nxG - 1 || j == nyG - 1 || k == nzG - 1) {
            int c = k * nxyG + j * nxG + i;
            ori[c] = 0;
            gid[c] = 0;
          }
        }
    grains.reserve((size_t)p.nx * p.ny * p.nz);
  }

  void surface_cells(std::vector<int>& ids) const {
    // CellularAutomataManager_3D::get_surface_cells (:958-1000), including the
    // duplicate edge/corner entries and the y-bound test against nxLocGhost.
    ids.clear();
    if (p.surface_mode == 1) {
      // extension: the z = 0 face only, one entry per cell
      for (int j = 1; j <= p.ny; j++)
        for (int i = 1; i <= p.nx; i++) ids.push_back(1 * nxyG + j * nxG + i);
      return;
    }
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
  }

  void bulk_cells(std::vector<int>& ids) const {
    // CellularAutomataManager_3D::get_bulk_cells (:934-953)
    ids.clear();
    for (int k = 1; k < nzG - 1; ++k)
      for (int j = 1; j < nyG - 1; ++j)
        for (int i = 1; i < nxG - 1; ++i) ids.push_back(nxyG * k + nxG * j + i);
  }

  void create_sites(double rho, double mean, double sd, bool surface) {
    // CellularAutomataManager::create_nucleation_sites (CellularAutomataManager.cpp:153-213)
    std::vector<int> ids;
    if (surface) surface_cells(ids); else bulk_cells(ids);
    double vpc = surface ? p.h * p.h : p.h * p.h * p.h;
    int numCells = (int)ids.size();
    int numSites = round(vpc * rho * numCells) + 0.5;
    std::default_random_engine gen;
    std::normal_distribution<double> dist(mean, sd);
    int made = 0;
    for (int s = 0; s < numSites; ++s) {
      int c = ids[rand() % numCells];
      siteDraws++;
      double dT = dist(gen);
      dT = (dT < 0) ? 0 : dT;
      double cur = nucDT[c];
      if (cur < 0) {
        sites.push_back(c);
        nucDT[c] = dT;
        made++;
      } else if (dT < cur) {
        nucDT[c] = dT;
      }
    }
    sites.sort();
    if (surface) nSurface = made; else nBulk = made;
  }

  int new_grain(int cell, int orient, const double* x, double L) {
    // CellularAutomataManager_3D::create_new_grain (:855-871) + Grain ctor
    // appending to the active list (Grain.cpp:13-18)
    Grain g;
    g.id = (int)grains.size();
    g.cell = cell;
    g.xc = x[0]; g.yc = x[1]; g.zc = x[2];
    g.orient = orient;
    g.length = L;
    g.active = true;
    g.wave = 0;
    grains.push_back(g);
    active.push_back(g.id);
    return g.id;
  }

  int nucleate(double t) {
    // CellularAutomataManager::nucleate_at_nucleation_sites (CellularAutomataManager.cpp:219-297)
    // with the orientation draw of CellularAutomataManager_3D.cpp:838-850; at
    // np=1 the global orientation id is the running count (:1129-1215).
    int made = 0;
    auto it = sites.begin();
    while (it != sites.end()) {
      int c = *it;
      bool liquid = (gid[c] == -1);
      if (liquid) {
        double dTc = nucDT[c];
        double T = p.T0 + p.Tdot * t;  // ProblemPhysics.cpp:163-169
        double dT = p.Tm - T;
        if (dT > dTc) {
          Orient o;
          if (givenEuler) {
            int k = (int)orients.size();
            if (k >= givenCount) return -1;
            for (int q = 0; q < 3; q++) o.euler[q] = givenEuler[3 * k + q];
          } else {
            double U = static_cast<double>(rand()) / RAND_MAX;
            double V = static_cast<double>(rand()) / RAND_MAX;
            o.euler[0] = U * 2 * kPi;
            o.euler[1] = acos(2 * V - 1);
            o.euler[2] = static_cast<double>(rand()) / RAND_MAX * 0.5 * kPi;
          }
          rotation_from_euler(o.euler, o.l2g);
          int oid = (int)orients.size();
          orients.push_back(o);
          double x[3];
          centre(c, x);
          int g = new_grain(c, oid, x, 0.0);
          gid[c] = g;
          ori[c] = oid;
          liquid = false;
          made++;
        }
      }
      if (liquid) {
        ++it;
      } else {
        nucDT[c] = 0;
        it = sites.erase(it);
      }
    }
    return made;
  }

  void capture(const Grain& parent, int c) {
    // CellularAutomataManager_3D::capture_cell (:676-707)
    const double* l2g = orients[parent.orient].l2g;
    double cc[3], gc[3];
    centre(c, cc);
    double pcen[3] = {parent.xc, parent.yc, parent.zc};
    for (int i = 0; i < 3; i++) gc[i] = cc[i] - pcen[i];
    double info[4];
    child_envelope(gc, parent.length, p.h, p.trunc, l2g, info);
    double x[3];
    for (int i = 0; i < 3; i++) x[i] = info[i] + pcen[i];
    int g = new_grain(c, parent.orient, x, info[3]);
    grains[g].wave = parent.wave + 1;
    gid[c] = g;
    ori[c] = parent.orient;
  }

  // One pass of calculate_time_step + single_integration_step
  // (CellularAutomataManager_3D.cpp:440-669).  Returns captures made.
  int do_step(oracle_step_record* rec) {
    double t0 = time;
    int nuc = nucleate(time);
    if (nuc < 0) return -1;
    double maxV = 1e-10;
    for (int g : active) {
      double v = velocity(time);
      if (v > maxV) maxV = v;
    }
    double dtLoc = std::min(p.tsf * p.h / maxV, p.max_dt);
    dT = dtLoc;
    for (int g : active) {
      double v = velocity(time);
      grains[g].length += dT * v;  // Grain::grow (Grain.h:27)
    }
    int nActive0 = (int)active.size();
    for (int g : active) grains[g].wave = 0;
    int captures = 0, maxWave = 0, wave1 = 0;
    int nb[26];
    auto it = active.begin();
    while (it != active.end()) {
      int g = *it;
      Grain& gr = grains[g];
      neighbours(gr.cell, nb);
      bool open = false;
      const double* l2g = orients[gr.orient].l2g;
      for (int s = 0; s < 26; ++s) {
        int n = nb[s];
        if (gid[n] == -1) {
          double cc[3];
          centre(n, cc);
          if (capture_test(grains[g], l2g, cc)) {
            Grain par = grains[g];  // grains may reallocate on push_back
            capture(par, n);
            captures++;
            if (par.wave == 0) wave1++;
            maxWave = std::max(maxWave, par.wave + 1);
          } else {
            open = true;
          }
        }
      }
      if (!open) {
        grains[g].active = false;
        it = active.erase(it);
      } else {
        ++it;
      }
    }
    time += dT;
    step++;
    int liquid = 0;
    for (int k = 1; k <= p.nz; k++)
      for (int j = 1; j <= p.ny; j++)
        for (int i = 1; i <= p.nx; i++)
          if (ori[k * nxyG + j * nxG + i] < 0) liquid++;
    if (liquid == 0) done = true;
    if (rec) {
      rec->time_before = t0;
      rec->dt = dT;
      rec->n_nucleated = nuc;
      rec->n_active_before = nActive0;
      rec->n_captured = captures;
      rec->n_active_after = (int)active.size();
      rec->n_liquid = liquid;
      rec->max_wave = maxWave;
      rec->n_wave1 = wave1;
    }
    return 0;
  }
};

}  // namespace

extern "C" {

namespace {

int run_loop(PlainCA& ca, const oracle_plain_params* prm, oracle_plain_out* out) {
  out->n_sites = (int)ca.sites.size();
  out->n_site_draws = ca.siteDraws;
  if (out->site_cell && out->site_dt) {
    int i = 0;
    for (int c : ca.sites) {
      if (i >= out->site_cap) break;
      out->site_cell[i] = c;
      out->site_dt[i] = ca.nucDT[c];
      i++;
    }
  }
  // CellularAutomataManager::time_integration (CellularAutomataManager.cpp:134-147)
  int nrec = 0;
  while (ca.time < prm->final_time && !ca.done) {
    if (prm->max_steps >= 0 && ca.step >= prm->max_steps) break;
    oracle_step_record* rec = (out->steps && nrec < out->steps_cap) ? &out->steps[nrec] : nullptr;
    if (ca.do_step(rec) != 0) return 2;
    if (rec) nrec++;
  }
  out->n_steps = ca.step;
  out->n_records = nrec;
  out->time = ca.time;
  out->done = ca.done ? 1 : 0;
  out->n_grains = (int)ca.grains.size();
  out->n_orient = (int)ca.orients.size();
  out->n_active = (int)ca.active.size();
  out->ng = ca.NG;
  if (out->ori) memcpy(out->ori, ca.ori.data(), sizeof(int32_t) * ca.NG);
  if (out->gid) memcpy(out->gid, ca.gid.data(), sizeof(int32_t) * ca.NG);
  if (out->nuc_dt) memcpy(out->nuc_dt, ca.nucDT.data(), sizeof(double) * ca.NG);
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
    if (out->grain_active) out->grain_active[g] = gr.active ? 1 : 0;
  }
  int no = std::min<int>(out->orient_cap, (int)ca.orients.size());
  for (int o = 0; o < no; o++) {
    if (out->euler) memcpy(out->euler + 3 * o, ca.orients[o].euler, sizeof(double) * 3);
  }
  return 0;
}

}  // namespace

int oracle_plain_run(const oracle_plain_params* prm, oracle_plain_out* out) {
  PlainCA ca;
  ca.p = *prm;
  ca.setup();
  srand(prm->seed);  // CellularAutomataManager_3D.cpp:40 (srand(rank*10000))
  // surface sites first, then bulk (CellularAutomataManager_3D.cpp:366-379)
  ca.create_sites(prm->ns, prm->dTs_mean, prm->dTs_sd, true);
  ca.create_sites(prm->nv, prm->dTv_mean, prm->dTv_sd, false);
  return run_loop(ca, prm, out);
}

int oracle_plain_run_with(const oracle_plain_params* prm, int n_sites, const int32_t* site_cell,
                          const double* site_dt, int n_orient, const double* euler3,
                          oracle_plain_out* out) {
  PlainCA ca;
  ca.p = *prm;
  ca.setup();
  for (int i = 0; i < n_sites; i++) {
    ca.sites.push_back(site_cell[i]);
    ca.nucDT[site_cell[i]] = site_dt[i];
  }
  ca.sites.sort();
  ca.givenEuler = euler3;
  ca.givenCount = n_orient;
  return run_loop(ca, prm, out);
}

void oracle_rotation(const double* euler3, double* l2g9) { rotation_from_euler(euler3, l2g9); }

int oracle_capture_test(const double* env4, const double* l2g9, const double* cell3) {
  Grain g;
  g.xc = env4[0]; g.yc = env4[1]; g.zc = env4[2]; g.length = env4[3];
  return capture_test(g, l2g9, cell3) ? 1 : 0;
}

void oracle_child_envelope(const double* gc3, double L, double h, double trunc,
                           const double* l2g9, double* out4) {
  child_envelope(gc3, L, h, trunc, l2g9, out4);
}

void oracle_glibc_rand(unsigned seed, int n, int32_t* out) {
  srand(seed);
  for (int i = 0; i < n; i++) out[i] = rand();
}

void oracle_std_normal(double mean, double sd, int n, double* out) {
  std::default_random_engine gen;
  std::normal_distribution<double> dist(mean, sd);
  for (int i = 0; i < n; i++) out[i] = dist(gen);
}

void oracle_minstd(int n, int64_t* out) {
  std::default_random_engine gen;
  for (int i = 0; i < n; i++) out[i] = (int64_t)gen();
}

float oracle_sqrtf(float x) { return sqrtf(x); }

void oracle_envelope_batch(int n, const double* env4, const double* l2g9, const double* cell3,
                           double h, double trunc, int32_t* inside, double* child4) {
  for (int i = 0; i < n; i++) {
    const double* e = env4 + 4 * (size_t)i;
    const double* r = l2g9 + 9 * (size_t)i;
    const double* x = cell3 + 3 * (size_t)i;
    Grain g;
    g.xc = e[0]; g.yc = e[1]; g.zc = e[2]; g.length = e[3];
    bool in = capture_test(g, r, x);
    inside[i] = in ? 1 : 0;
    double out[4] = {0.0, 0.0, 0.0, 0.0};
    double gc[3] = {x[0] - e[0], x[1] - e[1], x[2] - e[2]};
    if (in) child_envelope(gc, e[3], h, trunc, r, out);
    child4[4 * (size_t)i + 0] = in ? out[0] + e[0] : 0.0;
    child4[4 * (size_t)i + 1] = in ? out[1] + e[1] : 0.0;
    child4[4 * (size_t)i + 2] = in ? out[2] + e[2] : 0.0;
    child4[4 * (size_t)i + 3] = out[3];
  }
}

}  // extern "C"
