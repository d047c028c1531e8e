// C ABI of a coupled CA-FE (remelting) domain (include/amca3d.h, amca_remelt_*).
//
// The reference's CafeSolver owns a FiniteElementManager that streams text
// frames and a MapVoxelManager that interpolates them per cell every step
// (CafeSolver.cpp:159-219).  Here the FE frames, the voxel map and the whole CA
// state are resident in HBM; the host only maps voxels once, refills the
// orientation rows it draws from the rand() stream, and reads counters back
// between batches of steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/amca3d.h"
#include "host/fe_text.h"
#include "host/generators.h"
#include "host/sites.h"
#include "host/voxel_map.h"
#include "env.h"
#include "libm_pow.h"
#include "plain_step.h"
#include "remelt_step.h"
#include "surrogate.h"

namespace {

thread_local std::string g_rm_error;

amca_status rm_fail(amca_status st, const std::string& msg) {
  g_rm_error = msg;
  return st;
}

#define RM_HIP(expr)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return rm_fail(AMCA_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
  } while (0)

#define RM_TRY(expr)              \
  do {                            \
    amca_status s_ = (expr);      \
    if (s_ != AMCA_OK) return s_; \
  } while (0)

}  // namespace

struct amca_remelt {
  hipStream_t stream = nullptr;
  amca_remelt_desc desc{};
  amca::RmParams P{};
  amca::RmBuffers B{};
  std::vector<void*> allocs;
  bool fe_ready = false, state_ready = false;
  int site_blocks = 1;
  int rows_pushed = 0;
  amca_orientation_source source = nullptr;
  void* source_user = nullptr;
  std::unique_ptr<amca_host::GlibcRand> rng;  // the rand() stream after amca_remelt_draw_sites
  std::vector<int32_t> site_cells;            // nucleation sites as last set
  std::vector<double> site_dts;
  std::vector<double> euler;                  // Euler angles of every row, 3 per row
  size_t grain_cap = 0;  // records genv/gco can hold
  std::vector<int32_t> host_tag;
  std::vector<double> host_par;
  std::vector<double> node_xyz;
  // per-kernel event timing (off by default)
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, size_t>> ev_pending;
  size_t ev_used = 0;
  double kernel_ms[amca::kRmKernels] = {};
  int64_t kernel_launches[amca::kRmKernels] = {};

  template <class T>
  amca_status alloc(T** p, size_t count) {
    void* q = nullptr;
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess)
      return rm_fail(AMCA_ERR_HIP, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
    allocs.push_back(q);
    *p = static_cast<T*>(q);
    return AMCA_OK;
  }
  void release(void* p) {
    if (!p) return;
    for (auto& a : allocs)
      if (a == p) {
        (void)hipFree(a);
        a = nullptr;
      }
    allocs.erase(std::remove(allocs.begin(), allocs.end(), nullptr), allocs.end());
  }
  // kGraphSteps steps captured as one HIP graph, valid while the kernels'
  // arguments (parameters, buffers, site blocks) are the ones captured
  int graph_steps = 8;  // 0: every kernel launched on the stream
  int row_reserve = 4096;  // orientation rows drawn ahead of the rows in use
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<unsigned char> graph_sig;

  void drop_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
    graph_sig.clear();
  }
  std::vector<unsigned char> launch_sig() const {
    std::vector<unsigned char> v(sizeof(P) + sizeof(B) + sizeof(int));
    std::memcpy(v.data(), &P, sizeof(P));
    std::memcpy(v.data() + sizeof(P), &B, sizeof(B));
    std::memcpy(v.data() + sizeof(P) + sizeof(B), &site_blocks, sizeof(int));
    return v;
  }
  ~amca_remelt() {
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graph();
    for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
    for (void* p : allocs) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

amca_status rm_counters(amca_remelt* d, amca::RmCounters* c) {
  RM_HIP(hipMemcpyAsync(c, d->B.ctr, sizeof(*c), hipMemcpyDeviceToHost, d->stream));
  RM_HIP(hipStreamSynchronize(d->stream));
  return AMCA_OK;
}

amca_status rm_put_counters(amca_remelt* d, const amca::RmCounters& c) {
  RM_HIP(hipMemcpyAsync(d->B.ctr, &c, sizeof(c), hipMemcpyHostToDevice, d->stream));
  RM_HIP(hipStreamSynchronize(d->stream));
  return AMCA_OK;
}

// the next step runs over the whole grid and restarts the box bookkeeping
void rm_restart_box(amca::RmCounters& c) {
  for (int q = 0; q < 2; q++) {
    for (int r = 0; r < 3; r++)
      for (int a = 0; a < 3; a++) {
        c.acc[q][r].lo[a] = INT_MAX;
        c.acc[q][r].hi[a] = INT_MIN;
      }
  }
  c.force_full = 1;
}

void rm_fill(const amca_remelt* d, const amca::RmCounters& c, amca_remelt_counters* o) {
  if (!o) return;
  o->step = c.step;
  o->time = c.t;
  o->time_fe = c.time_fe;
  o->dt = c.dt;
  o->max_dt = c.max_dt;
  o->n_grains = c.n_grains;
  o->n_orient = c.n_orient;
  o->frames_exhausted = c.exhausted;
  o->done = (c.exhausted || !(c.tcafe < d->desc.final_time)) ? 1 : 0;
}

amca_status push_rows(amca_remelt* d, int n, const double* l2g9) {
  if (d->rows_pushed + n > d->P.orient_cap)
    return rm_fail(AMCA_ERR_CAPACITY, "orientation table capacity exceeded");
  if (n > 0)
    RM_HIP(hipMemcpyAsync(d->B.l2g + 9 * (size_t)d->rows_pushed, l2g9, sizeof(double) * 9 * n,
                          hipMemcpyHostToDevice, d->stream));
  d->rows_pushed += n;
  // the device runs a step only when its armed sites fit in the pushed rows
  amca::RmCounters c;
  RM_TRY(rm_counters(d, &c));
  c.rows_avail = d->rows_pushed;
  RM_TRY(rm_put_counters(d, c));
  RM_HIP(hipStreamSynchronize(d->stream));
  return AMCA_OK;
}

// make sure `need` rows exist, pulling from the orientation source
amca_status ensure_rows(amca_remelt* d, long long need) {
  if (need <= d->rows_pushed) return AMCA_OK;
  if (!d->source && !d->rng)
    return rm_fail(AMCA_ERR_CAPACITY, "orientation rows exhausted and no orientation source set");
  need = std::min<long long>(need, d->P.orient_cap);
  const int n = (int)(need - d->rows_pushed);
  if (n <= 0) return AMCA_OK;
  std::vector<double> eu(3 * (size_t)n), r(9 * (size_t)n);
  if (d->rng) {
    for (int q = 0; q < n; q++) {
      amca_host::random_orientation(*d->rng, &eu[3 * (size_t)q]);
      amca_host::rotation_from_euler(&eu[3 * (size_t)q], &r[9 * (size_t)q]);
    }
  } else if (d->source(d->source_user, n, eu.data(), r.data()) != 0) {
    return rm_fail(AMCA_ERR_INVALID, "orientation source failed");
  }
  d->euler.insert(d->euler.end(), eu.begin(), eu.end());
  return push_rows(d, n, r.data());
}

amca_status rm_resolve_events(amca_remelt* d) {
  RM_HIP(hipStreamSynchronize(d->stream));
  for (auto& p : d->ev_pending) {
    float ms = 0.f;
    RM_HIP(hipEventElapsedTime(&ms, d->ev_pool[p.second], d->ev_pool[p.second + 1]));
    d->kernel_ms[p.first] += ms;
    d->kernel_launches[p.first] += 1;
  }
  d->ev_pending.clear();
  d->ev_used = 0;
  return AMCA_OK;
}

// k steps without per-kernel timing: whole graphs of graph_steps steps, then
// single steps for the rest
amca_status rm_enqueue_steps(amca_remelt* d, long long k) {
  const int G = d->graph_steps > 0 ? d->graph_steps : 1;
  if (d->graph_steps > 0 && k >= G) {
    std::vector<unsigned char> sig = d->launch_sig();
    if (!d->graph_exec || sig != d->graph_sig) {
      d->drop_graph();
      RM_HIP(hipStreamBeginCapture(d->stream, hipStreamCaptureModeThreadLocal));
      hipError_t e = hipSuccess;
      for (int q = 0; q < G && e == hipSuccess; q++)
        e = amca::rm_launch_step(d->P, d->B, d->site_blocks, d->stream);
      hipGraph_t g = nullptr;
      const hipError_t e2 = hipStreamEndCapture(d->stream, &g);
      if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return rm_fail(AMCA_ERR_HIP, std::string("graph capture: ") +
                                         hipGetErrorString(e != hipSuccess ? e : e2));
      }
      d->graph = g;
      RM_HIP(hipGraphInstantiate(&d->graph_exec, g, nullptr, nullptr, 0));
      d->graph_sig = std::move(sig);
    }
    for (long long q = 0; q < k / G; q++) RM_HIP(hipGraphLaunch(d->graph_exec, d->stream));
  }
  const long long rest = d->graph_steps > 0 ? k % G : k;
  for (long long q = 0; q < rest; q++)
    RM_HIP(amca::rm_launch_step(d->P, d->B, d->site_blocks, d->stream));
  return AMCA_OK;
}

// one step; with timing on, each kernel sits between two events on the stream
amca_status rm_enqueue_step(amca_remelt* d) {
  if (!d->timing) {
    RM_HIP(amca::rm_launch_step(d->P, d->B, d->site_blocks, d->stream));
    return AMCA_OK;
  }
  for (int q = 0; q < amca::kRmKernels; q++) {
    if (d->ev_used + 2 > d->ev_pool.size()) {
      if (d->ev_pool.size() >= (1u << 16)) RM_TRY(rm_resolve_events(d));
      while (d->ev_used + 2 > d->ev_pool.size()) {
        hipEvent_t e;
        RM_HIP(hipEventCreate(&e));
        d->ev_pool.push_back(e);
      }
    }
    const size_t i = d->ev_used;
    d->ev_used += 2;
    RM_HIP(hipEventRecord(d->ev_pool[i], d->stream));
    RM_HIP(amca::rm_launch_kernel(q, d->P, d->B, d->site_blocks, d->stream));
    RM_HIP(hipEventRecord(d->ev_pool[i + 1], d->stream));
    d->ev_pending.emplace_back(q, i);
  }
  return AMCA_OK;
}

}  // namespace

extern "C" {

const char* amca_remelt_last_error(void) { return g_rm_error.c_str(); }

amca_status amca_remelt_enable_timing(amca_remelt* d, int32_t on) {
  if (!d) return rm_fail(AMCA_ERR_INVALID, "null domain");
  RM_TRY(rm_resolve_events(d));
  d->timing = on != 0;
  for (int q = 0; q < amca::kRmKernels; q++) {
    d->kernel_ms[q] = 0.0;
    d->kernel_launches[q] = 0;
  }
  return AMCA_OK;
}

int32_t amca_remelt_kernel_count(void) { return amca::kRmKernels; }

const char* amca_remelt_kernel_name(int32_t i) {
  return (i >= 0 && i < amca::kRmKernels) ? amca::kRmKernelNames[i] : "";
}

amca_status amca_remelt_kernel_times(amca_remelt* d, double* ms, int64_t* launches) {
  if (!d || !ms || !launches) return rm_fail(AMCA_ERR_INVALID, "null argument");
  RM_TRY(rm_resolve_events(d));
  for (int q = 0; q < amca::kRmKernels; q++) {
    ms[q] = d->kernel_ms[q];
    launches[q] = d->kernel_launches[q];
  }
  return AMCA_OK;
}

amca_status amca_remelt_create(const amca_remelt_desc* desc, int32_t orient_capacity,
                               amca_remelt** out) {
  if (!desc || !out) return rm_fail(AMCA_ERR_INVALID, "null argument");
  *out = nullptr;
  if (desc->nx < 1 || desc->ny < 1 || desc->nz < 1 || !(desc->h > 0))
    return rm_fail(AMCA_ERR_INVALID, "bad grid");
  for (double b : {desc->b1, desc->b2, desc->b3})
    if (!amca_libm::pow_exponent_supported(b))
      return rm_fail(AMCA_ERR_INVALID,
                     "kinetics exponents b1, b2, b3 must be 0 or of magnitude in [2^-65, 2^63)");
  const long long nxG = desc->nx + 2LL, nyG = desc->ny + 2LL, nzG = desc->nz + 2LL;
  const long long NG = nxG * nyG * nzG;
  if (NG > (long long)INT_MAX) return rm_fail(AMCA_ERR_INVALID, "grid exceeds int32 cell indices");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return rm_fail(AMCA_ERR_NO_DEVICE, "no HIP device visible");
  auto* d = new amca_remelt();
  d->desc = *desc;
  if (amca::env_int("AMCA_REMELT_GRAPH_STEPS", 0, 1024, &d->graph_steps) < 0) {
    delete d;
    return rm_fail(AMCA_ERR_INVALID, "AMCA_REMELT_GRAPH_STEPS must be an integer in [0, 1024]");
  }
  if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess) {
    delete d;
    return rm_fail(AMCA_ERR_HIP, "stream creation failed");
  }
  amca::RmParams& P = d->P;
  P.nx = desc->nx; P.ny = desc->ny; P.nz = desc->nz;
  P.nxG = (int)nxG; P.nyG = (int)nyG; P.nzG = (int)nzG;
  P.nxyG = (int)(nxG * nyG);
  P.NG = (int)NG;
  P.N = (long long)desc->nx * desc->ny * desc->nz;
  P.h = desc->h; P.x0 = desc->x0; P.y0 = desc->y0; P.z0 = desc->z0;
  P.Tm = desc->Tm; P.Ts = desc->Ts; P.Tmin_nuc = desc->Tmin_nuc;
  P.a1 = desc->a1; P.a2 = desc->a2; P.a3 = desc->a3;
  P.b1 = desc->b1; P.b2 = desc->b2; P.b3 = desc->b3;
  P.trunc = desc->trunc;
  P.tsf = desc->time_step_factor;
  P.sqrt2f = (double)sqrtf(2.0f);
  P.cutoff = desc->cutoff_factor * desc->h;  // cellGrainLegnthCutoff_ *= h_
  P.final_time = desc->final_time;
  P.void_temp = desc->void_temperature;
  P.orient_cap = orient_capacity;
  P.window = 1;
  P.hot_margin = 1.0;
  P.sur_hot[0] = 1;
  P.sur_hot[1] = 0;
  {
    int s = 0;
    for (int dk = -1; dk <= 1; dk++)
      for (int dj = -1; dj <= 1; dj++)
        for (int di = -1; di <= 1; di++) {
          if (!dk && !dj && !di) continue;
          P.off[s++] = dk * P.nxyG + dj * P.nxG + di;
        }
  }
  amca::RmBuffers& B = d->B;
  const size_t ng = (size_t)NG, nc = (size_t)P.N + 1;
  amca_status st = AMCA_OK;
  auto A = [&](auto** p, size_t n) {
    if (st == AMCA_OK) st = d->alloc(p, n);
  };
  A(&B.gid, ng);
  A(&B.oid, ng);
  A(&B.T, ng);
  A(&B.cooling, ng);
  A(&B.melted, ng);
  A(&B.flag, ng);
  A(&B.tag, ng);
  A(&B.par, 3 * ng);
  A(&B.claim, ng);
  A(&B.msk, ng);
  A(&B.genv, nc);
  A(&B.gco, nc);
  d->grain_cap = nc;
  A(&B.l2g, 9 * (size_t)std::max(orient_capacity, 1));
  A(&B.rowcnt, (size_t)P.nx * P.ny);
  A(&B.rowoff, (size_t)P.nx * P.ny);
  A(&B.lb_flag, (size_t)amca::kMaxBlocks);
  A(&B.lb_agg, (size_t)amca::kMaxBlocks);
  A(&B.lb_pref, (size_t)amca::kMaxBlocks);
  A(&B.ctr, 1);
  A(&B.blk_box, 6 * 16384);
  A(&B.site_cell, 1);
  A(&B.site_dt, 1);
  if (st != AMCA_OK) {
    delete d;
    return st;
  }
  B.log = nullptr;
  B.frames = nullptr;
  B.frame_time = nullptr;
  B.theta = nullptr;
  // every cell liquid without a grain until a state is loaded
  (void)hipMemsetAsync(B.gid, 0xff, sizeof(int) * ng, d->stream);
  (void)hipMemsetAsync(B.oid, 0xff, sizeof(int) * ng, d->stream);
  (void)hipMemsetAsync(B.melted, desc->initial_melted ? 1 : 0, ng, d->stream);
  (void)hipMemsetAsync(B.lb_flag, 0, sizeof(unsigned) * amca::kMaxBlocks, d->stream);
  amca::RmCounters c{};
  for (int q = 0; q < 64; q++) c.min_t_keys[q] = ~0ull;
  c.last_step = -1;
  for (int q = 0; q < 2; q++) {
    c.box[q].lo[0] = c.box[q].lo[1] = c.box[q].lo[2] = 0;
    c.box[q].hi[0] = P.nxG - 1;
    c.box[q].hi[1] = P.nyG - 1;
    c.box[q].hi[2] = P.nzG - 1;
  }
  rm_restart_box(c);
  // the velocity at Ts + 0.01 K, through the step's own kinetics code
  P.vts = 0.0;
  double* vts_dev = nullptr;
  const bool ok = hipMalloc(&vts_dev, sizeof(double)) == hipSuccess &&
                  hipMemcpyAsync(B.ctr, &c, sizeof(c), hipMemcpyHostToDevice, d->stream) == hipSuccess &&
                  amca::rm_launch_init(P, B, d->stream) == hipSuccess &&
                  amca::rm_launch_vts(P, vts_dev, d->stream) == hipSuccess &&
                  hipMemcpyAsync(&P.vts, vts_dev, sizeof(double), hipMemcpyDeviceToHost,
                                 d->stream) == hipSuccess &&
                  hipStreamSynchronize(d->stream) == hipSuccess;
  if (vts_dev) (void)hipFree(vts_dev);
  if (!ok) {
    delete d;
    return rm_fail(AMCA_ERR_HIP, "initialisation failed");
  }
  *out = d;
  return AMCA_OK;
}

amca_status amca_remelt_destroy(amca_remelt* d) {
  delete d;
  return AMCA_OK;
}

namespace {

// tag_voxels on the device: one thread per CA cell (ghosts included), the
// arithmetic of the host mapper (voxel_math.h), so tags and parametric
// coordinates equal the host's bit for bit
__global__ void k_map_voxels(amca_vox::MeshView M, amca_vox::VoxelGrid G,
                             const int* __restrict__ bin_start, const int* __restrict__ bin_elems,
                             long long NG, double tol, int max_iter, int32_t* __restrict__ tag,
                             double* __restrict__ par) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < NG; v += stride) {
    double pc[3];
    const int e = amca_vox::tag_voxel(M, G, bin_start, bin_elems, v, tol, max_iter, pc);
    tag[v] = e;
    for (int q = 0; q < 3; q++) par[3 * v + q] = e >= 0 ? pc[q] : -2.0;
  }
}

// map_voxels on the device: the plan (search box, bins) on the host, the
// per-cell search in k_map_voxels; results copied into host vectors.
// Returns false when no cell found an element.
amca_status map_voxels_device(const amca_host::FeMesh& mesh, const amca_host::VoxelMapParams& vp,
                              std::vector<int32_t>& tag, std::vector<double>& par, bool& any) {
  amca_host::VoxelPlan plan;
  amca_host::plan_voxels(mesh, vp, plan);
  const amca_vox::VoxelGrid& G = plan.grid;
  const long long NG = (long long)G.nxG * G.nyG * G.nzG;
  if (plan.bin_elems.empty()) plan.bin_elems.push_back(0);  // a valid pointer; no bin reads it
  double* dxyz = nullptr;
  int *dbs = nullptr, *dbe = nullptr;
  int32_t* dtag = nullptr;
  double* dpar = nullptr;
  amca_status st = AMCA_OK;
  auto hip = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && st == AMCA_OK) st = rm_fail(AMCA_ERR_HIP, std::string(what) + ": " +
                                                              hipGetErrorString(e));
    return e == hipSuccess;
  };
  if (hip(hipMalloc(&dxyz, sizeof(double) * mesh.xyz.size()), "voxel map alloc") &&
      hip(hipMalloc(&dbs, sizeof(int) * plan.bin_start.size()), "voxel map alloc") &&
      hip(hipMalloc(&dbe, sizeof(int) * plan.bin_elems.size()), "voxel map alloc") &&
      hip(hipMalloc(&dtag, sizeof(int32_t) * (size_t)NG), "voxel map alloc") &&
      hip(hipMalloc(&dpar, sizeof(double) * 3 * (size_t)NG), "voxel map alloc") &&
      hip(hipMemcpy(dxyz, mesh.xyz.data(), sizeof(double) * mesh.xyz.size(),
                    hipMemcpyHostToDevice), "voxel map upload") &&
      hip(hipMemcpy(dbs, plan.bin_start.data(), sizeof(int) * plan.bin_start.size(),
                    hipMemcpyHostToDevice), "voxel map upload") &&
      hip(hipMemcpy(dbe, plan.bin_elems.data(), sizeof(int) * plan.bin_elems.size(),
                    hipMemcpyHostToDevice), "voxel map upload")) {
    const amca_vox::MeshView mv{dxyz, mesh.fnx, mesh.fny, mesh.fnz};
    long long blocks = (NG + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_map_voxels, dim3((unsigned)blocks), dim3(256), 0, 0, mv, G, dbs, dbe, NG,
                       vp.tol, vp.max_iter, dtag, dpar);
    tag.resize((size_t)NG);
    par.resize(3 * (size_t)NG);
    if (hip(hipGetLastError(), "k_map_voxels") &&
        hip(hipMemcpy(tag.data(), dtag, sizeof(int32_t) * (size_t)NG, hipMemcpyDeviceToHost),
            "voxel map readback"))
      hip(hipMemcpy(par.data(), dpar, sizeof(double) * 3 * (size_t)NG, hipMemcpyDeviceToHost),
          "voxel map readback");
  }
  for (void* ptr : {(void*)dxyz, (void*)dbs, (void*)dbe, (void*)dtag, (void*)dpar})
    if (ptr) (void)hipFree(ptr);
  any = false;
  if (st == AMCA_OK)
    for (long long v = 0; v < NG && !any; v++) any = tag[(size_t)v] >= 0;
  return st;
}

}  // namespace

amca_status amca_remelt_set_fe(amca_remelt* d, int32_t fnx, int32_t fny, int32_t fnz,
                               const double* node_xyz, int32_t n_frames, const double* frame_time,
                               const double* frame_theta) {
  if (!d || fnx < 2 || fny < 2 || fnz < 2 || !node_xyz || n_frames < 2 || !frame_time ||
      !frame_theta)
    return rm_fail(AMCA_ERR_INVALID, "FE mesh needs >= 2 nodes per axis and >= 2 frames");
  amca::RmParams& P = d->P;
  amca_host::FeMesh mesh;
  mesh.fnx = fnx;
  mesh.fny = fny;
  mesh.fnz = fnz;
  const size_t nodes = (size_t)fnx * fny * fnz;
  mesh.xyz.assign(node_xyz, node_xyz + 3 * nodes);
  d->node_xyz = mesh.xyz;
  amca_host::VoxelMapParams vp;
  vp.nx = P.nx; vp.ny = P.ny; vp.nz = P.nz;
  vp.h = P.h; vp.x0 = P.x0; vp.y0 = P.y0; vp.z0 = P.z0;
  for (int q = 0; q < 3; q++) vp.bin[q] = d->desc.bin[q] > 0 ? d->desc.bin[q] : 10;
  vp.tol = d->desc.tolerance;
  vp.max_iter = d->desc.max_iterations;
  bool any = false;
  RM_TRY(map_voxels_device(mesh, vp, d->host_tag, d->host_par, any));
  if (!any) return rm_fail(AMCA_ERR_INVALID, "no CA cell lies in the FE mesh");
  P.fnx = fnx;
  P.fny = fny;
  P.fnz = fnz;
  P.nodes = (int)nodes;
  P.ncx = fnx - 1;
  P.ncy = fny - 1;
  P.n_frames = n_frames;
  d->release((void*)d->B.frames);
  d->release((void*)d->B.frame_time);
  d->release(d->B.theta);
  d->release(const_cast<int*>(d->B.frame_hot));
  double *fr = nullptr, *ft = nullptr, *th = nullptr;
  int* fh = nullptr;
  const size_t ne = (size_t)(fnx - 1) * (fny - 1) * (fnz - 1);
  RM_TRY(d->alloc(&fr, nodes * (size_t)n_frames));
  RM_TRY(d->alloc(&ft, (size_t)n_frames));
  RM_TRY(d->alloc(&th, nodes));
  RM_TRY(d->alloc(&fh, 6 * (size_t)n_frames));
  // For the active box: the cell box of the elements around each node, then
  // per frame the union over its hot nodes (at or above Tm - hot_margin), and
  // the union over nodes whose void state can change between frames.
  {
    std::vector<int32_t> box(6 * ne);
    for (size_t e = 0; e < ne; e++) {
      box[6 * e + 0] = box[6 * e + 1] = box[6 * e + 2] = INT_MAX;
      box[6 * e + 3] = box[6 * e + 4] = box[6 * e + 5] = INT_MIN;
    }
    for (int c = 0; c < P.NG; c++) {
      const int e = d->host_tag[c];
      if (e < 0) continue;
      const int k = c / P.nxyG, j = (c % P.nxyG) / P.nxG, i = (c % P.nxyG) % P.nxG;
      int32_t* b = &box[6 * (size_t)e];
      b[0] = std::min(b[0], i);
      b[1] = std::min(b[1], j);
      b[2] = std::min(b[2], k);
      b[3] = std::max(b[3], i);
      b[4] = std::max(b[4], j);
      b[5] = std::max(b[5], k);
    }
    std::vector<int32_t> nbox(6 * nodes);
    for (size_t n = 0; n < nodes; n++) {
      nbox[6 * n + 0] = nbox[6 * n + 1] = nbox[6 * n + 2] = INT_MAX;
      nbox[6 * n + 3] = nbox[6 * n + 4] = nbox[6 * n + 5] = INT_MIN;
    }
    const int ncx = fnx - 1, ncy = fny - 1;
    for (size_t e = 0; e < ne; e++) {
      const int32_t* b = &box[6 * e];
      if (b[0] > b[3]) continue;
      const int ei = (int)(e % ncx), ej = (int)((e / ncx) % ncy), ek = (int)(e / ((size_t)ncx * ncy));
      for (int dk = 0; dk < 2; dk++)
        for (int dj = 0; dj < 2; dj++)
          for (int di = 0; di < 2; di++) {
            int32_t* nb = &nbox[6 * (((size_t)(ek + dk) * fny + (ej + dj)) * fnx + (ei + di))];
            for (int a = 0; a < 3; a++) {
              nb[a] = std::min(nb[a], b[a]);
              nb[3 + a] = std::max(nb[3 + a], b[3 + a]);
            }
          }
    }
    auto add = [](int32_t* dst, const int32_t* b) {
      if (b[0] > b[3]) return;
      for (int a = 0; a < 3; a++) {
        dst[a] = std::min(dst[a], b[a]);
        dst[3 + a] = std::max(dst[3 + a], b[3 + a]);
      }
    };
    const double hot = P.Tm - P.hot_margin;
    const double nm = P.void_temp + 0.00001;
    std::vector<int32_t> fhot(6 * (size_t)n_frames);
    std::vector<unsigned char> maybe_void(nodes, 0), always_void(nodes, 1);
    for (int f = 0; f < n_frames; f++) {
      int32_t* fb = &fhot[6 * (size_t)f];
      fb[0] = fb[1] = fb[2] = INT_MAX;
      fb[3] = fb[4] = fb[5] = INT_MIN;
      const double* v = frame_theta + (size_t)f * nodes;
      for (size_t n = 0; n < nodes; n++) {
        if (v[n] >= hot) add(fb, &nbox[6 * n]);
        // the field is void_temp where a frame is at or below it, and is
        // compared with void_temp + 1e-5: a node near that value in some
        // frame may change between void and material
        if (!(v[n] > nm + 1e-6)) maybe_void[n] = 1;
        if (!(v[n] <= P.void_temp)) always_void[n] = 0;
      }
    }
    int32_t vb[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN};
    for (size_t n = 0; n < nodes; n++)
      if (maybe_void[n] && !always_void[n]) add(vb, &nbox[6 * n]);
    for (int a = 0; a < 6; a++) P.void_box[a] = vb[a];
    RM_HIP(hipMemcpyAsync(fh, fhot.data(), sizeof(int32_t) * fhot.size(), hipMemcpyHostToDevice,
                          d->stream));
    RM_HIP(hipStreamSynchronize(d->stream));
  }
  std::vector<double> rel(n_frames);
  for (int f = 0; f < n_frames; f++) rel[f] = frame_time[f] - frame_time[0];  // inTimeOffset_
  RM_HIP(hipMemcpyAsync(fr, frame_theta, sizeof(double) * nodes * n_frames, hipMemcpyHostToDevice, d->stream));
  RM_HIP(hipMemcpyAsync(ft, rel.data(), sizeof(double) * n_frames, hipMemcpyHostToDevice, d->stream));
  // thetaArray_ starts as theta_[1], the first frame
  RM_HIP(hipMemcpyAsync(th, frame_theta, sizeof(double) * nodes, hipMemcpyHostToDevice, d->stream));
  RM_HIP(hipMemcpyAsync(d->B.tag, d->host_tag.data(), sizeof(int32_t) * P.NG, hipMemcpyHostToDevice, d->stream));
  RM_HIP(hipMemcpyAsync(d->B.par, d->host_par.data(), sizeof(double) * 3 * (size_t)P.NG,
                        hipMemcpyHostToDevice, d->stream));
  d->B.frames = fr;
  d->B.frame_time = ft;
  d->B.theta = th;
  d->B.frame_hot = fh;
  amca::RmCounters c;
  RM_TRY(rm_counters(d, &c));
  // the coming step's field is the first frame as loaded
  c.fs[c.step % 3] = amca::RmField{0, 1, 0.0};
  rm_restart_box(c);
  c.frame = 1;
  c.in_t1 = rel[0];
  c.in_t0 = rel[1];
  c.time_fe = 0.0;
  c.max_dt = c.in_t0 - c.in_t1;  // get_output_time_step_from_txt_file
  RM_TRY(rm_put_counters(d, c));
  d->fe_ready = true;
  return AMCA_OK;
}

amca_status amca_remelt_voxel_map(amca_remelt* d, int32_t* tag, double* par) {
  if (!d || !d->fe_ready) return rm_fail(AMCA_ERR_INVALID, "FE mesh not set");
  if (tag) std::memcpy(tag, d->host_tag.data(), sizeof(int32_t) * d->P.NG);
  if (par) std::memcpy(par, d->host_par.data(), sizeof(double) * 3 * (size_t)d->P.NG);
  return AMCA_OK;
}

amca_status amca_fe_map_voxels(int32_t nx, int32_t ny, int32_t nz, double h, double x0, double y0,
                               double z0, const int32_t* bin, double tolerance,
                               int32_t max_iterations, int32_t fnx, int32_t fny, int32_t fnz,
                               const double* node_xyz, int32_t* tag, double* par) {
  if (nx < 1 || ny < 1 || nz < 1 || fnx < 2 || fny < 2 || fnz < 2 || !node_xyz || !tag || !par)
    return rm_fail(AMCA_ERR_INVALID, "bad arguments");
  amca_host::FeMesh mesh;
  mesh.fnx = fnx;
  mesh.fny = fny;
  mesh.fnz = fnz;
  mesh.xyz.assign(node_xyz, node_xyz + 3 * (size_t)fnx * fny * fnz);
  amca_host::VoxelMapParams vp;
  vp.nx = nx; vp.ny = ny; vp.nz = nz;
  vp.h = h; vp.x0 = x0; vp.y0 = y0; vp.z0 = z0;
  for (int q = 0; q < 3; q++) vp.bin[q] = (bin && bin[q] > 0) ? bin[q] : 10;
  vp.tol = tolerance;
  vp.max_iter = max_iterations;
  std::vector<int32_t> t;
  std::vector<double> p;
  const bool any = amca_host::map_voxels(mesh, vp, t, p);
  std::memcpy(tag, t.data(), sizeof(int32_t) * t.size());
  std::memcpy(par, p.data(), sizeof(double) * p.size());
  return any ? AMCA_OK : rm_fail(AMCA_ERR_INVALID, "no CA cell lies in the FE mesh");
}

amca_status amca_fe_map_voxels_device(int32_t nx, int32_t ny, int32_t nz, double h, double x0,
                                      double y0, double z0, const int32_t* bin, double tolerance,
                                      int32_t max_iterations, int32_t fnx, int32_t fny,
                                      int32_t fnz, const double* node_xyz, int32_t* tag,
                                      double* par) {
  if (nx < 1 || ny < 1 || nz < 1 || fnx < 2 || fny < 2 || fnz < 2 || !node_xyz || !tag || !par)
    return rm_fail(AMCA_ERR_INVALID, "bad arguments");
  amca_host::FeMesh mesh;
  mesh.fnx = fnx;
  mesh.fny = fny;
  mesh.fnz = fnz;
  mesh.xyz.assign(node_xyz, node_xyz + 3 * (size_t)fnx * fny * fnz);
  amca_host::VoxelMapParams vp;
  vp.nx = nx; vp.ny = ny; vp.nz = nz;
  vp.h = h; vp.x0 = x0; vp.y0 = y0; vp.z0 = z0;
  for (int q = 0; q < 3; q++) vp.bin[q] = (bin && bin[q] > 0) ? bin[q] : 10;
  vp.tol = tolerance;
  vp.max_iter = max_iterations;
  std::vector<int32_t> t;
  std::vector<double> p;
  bool any = false;
  RM_TRY(map_voxels_device(mesh, vp, t, p, any));
  std::memcpy(tag, t.data(), sizeof(int32_t) * t.size());
  std::memcpy(par, p.data(), sizeof(double) * p.size());
  return any ? AMCA_OK : rm_fail(AMCA_ERR_INVALID, "no CA cell lies in the FE mesh");
}

amca_status amca_remelt_set_state(amca_remelt* d, const int32_t* orient, const int32_t* grain,
                                  const uint8_t* melted, int32_t n_orient, int32_t n_grains,
                                  const int32_t* grain_cell, const int32_t* grain_orient) {
  if (!d || !orient || !grain || !melted || n_grains < 0 || (n_grains > 0 && (!grain_cell || !grain_orient)))
    return rm_fail(AMCA_ERR_INVALID, "bad state arrays");
  amca::RmParams& P = d->P;
  if (n_orient < 0 || n_orient > d->rows_pushed)
    return rm_fail(AMCA_ERR_INVALID, "push the loaded orientation rows before the state");
  for (int g = 0; g < n_grains; g++)
    if (grain_orient[g] < 0 || grain_orient[g] >= n_orient)
      return rm_fail(AMCA_ERR_INVALID, "grain orientation out of range");
  // Every interior cell owns at most one grain record and no record is shared:
  // the device grows a record once per flagged cell, as the reference does when
  // each cell carries its own grain.  Records for grainless cells are added by
  // captures and nucleation, so the table needs n_grains + (grainless cells).
  std::vector<uint8_t> owned((size_t)std::max(n_grains, 1), 0);
  long long grainless = 0;
  for (int c = 0; c < P.NG; c++) {
    const int k = c / P.nxyG, j = (c % P.nxyG) / P.nxG, i = (c % P.nxyG) % P.nxG;
    if (i == 0 || j == 0 || k == 0 || i == P.nxG - 1 || j == P.nyG - 1 || k == P.nzG - 1) continue;
    const int g = grain[c];
    if (orient[c] < -1 || orient[c] >= n_orient)
      return rm_fail(AMCA_ERR_INVALID, "cell orientation out of range");
    if (g < 0) {
      grainless++;
      continue;
    }
    if (g >= n_grains) return rm_fail(AMCA_ERR_INVALID, "cell grain id out of range");
    if (owned[g]) return rm_fail(AMCA_ERR_INVALID, "a grain record is shared by two cells");
    owned[g] = 1;
  }
  const size_t need = (size_t)n_grains + (size_t)grainless;
  if (need > d->grain_cap) {
    RM_HIP(hipStreamSynchronize(d->stream));
    d->release(d->B.genv);
    d->release(d->B.gco);
    d->B.genv = nullptr;
    d->B.gco = nullptr;
    RM_TRY(d->alloc(&d->B.genv, need));
    RM_TRY(d->alloc(&d->B.gco, need));
    d->grain_cap = need;
  }
  // retrieve_grain_information_before_simulation (:312-332): ghosts liquid,
  // other cells without an orientation take their grain's
  std::vector<int32_t> o(orient, orient + P.NG);
  const int nxG = P.nxG, nyG = P.nyG, nzG = P.nzG;
  for (int c = 0; c < P.NG; c++) {
    const int k = c / P.nxyG, j = (c % P.nxyG) / nxG, i = (c % P.nxyG) % nxG;
    if (i == 0 || j == 0 || k == 0 || i == nxG - 1 || j == nyG - 1 || k == nzG - 1) {
      o[c] = -1;
      continue;
    }
    if (o[c] > -1) continue;
    const int g = grain[c];
    if (g < 0 || g >= n_grains)
      return rm_fail(AMCA_ERR_INVALID, "a cell without orientation has no grain record");
    o[c] = grain_orient[g];
  }
  // grain records: centre reset to the cell centre, zero length (:282-300)
  std::vector<double> env(4 * (size_t)std::max(n_grains, 1));
  std::vector<int32_t> co(2 * (size_t)std::max(n_grains, 1));
  for (int g = 0; g < n_grains; g++) {
    const int c = grain_cell[g];
    if (c < 0 || c >= P.NG) return rm_fail(AMCA_ERR_INVALID, "grain cell out of range");
    const int k = c / P.nxyG, j = (c % P.nxyG) / nxG, i = (c % P.nxyG) % nxG;
    env[4 * (size_t)g + 0] = P.x0 + P.h * (i - 0.5);
    env[4 * (size_t)g + 1] = P.y0 + P.h * (j - 0.5);
    env[4 * (size_t)g + 2] = P.z0 + P.h * (k - 0.5);
    env[4 * (size_t)g + 3] = 0.0;
    co[2 * (size_t)g] = c;
    co[2 * (size_t)g + 1] = grain_orient[g];
  }
  RM_HIP(hipMemcpyAsync(d->B.oid, o.data(), sizeof(int32_t) * P.NG, hipMemcpyHostToDevice, d->stream));
  RM_HIP(hipMemcpyAsync(d->B.gid, grain, sizeof(int32_t) * P.NG, hipMemcpyHostToDevice, d->stream));
  RM_HIP(hipMemcpyAsync(d->B.melted, melted, P.NG, hipMemcpyHostToDevice, d->stream));
  if (n_grains > 0) {
    RM_HIP(hipMemcpyAsync(d->B.genv, env.data(), sizeof(double) * 4 * n_grains, hipMemcpyHostToDevice, d->stream));
    RM_HIP(hipMemcpyAsync(d->B.gco, co.data(), sizeof(int32_t) * 2 * n_grains, hipMemcpyHostToDevice, d->stream));
  }
  amca::RmCounters c;
  RM_TRY(rm_counters(d, &c));
  c.n_grains = n_grains;
  c.n_orient = n_orient;
  rm_restart_box(c);
  RM_TRY(rm_put_counters(d, c));
  d->state_ready = true;
  return AMCA_OK;
}

amca_status amca_remelt_push_orientations(amca_remelt* d, int32_t n, const double* euler3,
                                          const double* l2g9) {
  if (!d || n < 0 || (n > 0 && !l2g9 && !euler3)) return rm_fail(AMCA_ERR_INVALID, "bad rows");
  if (!l2g9 && n > 0) {
    // Octahedron::set_coordinate_transformation_matrix on the host (glibc cos/sin)
    std::vector<double> r(9 * (size_t)n);
    for (int q = 0; q < n; q++) amca_host::rotation_from_euler(euler3 + 3 * q, &r[9 * (size_t)q]);
    RM_TRY(push_rows(d, n, r.data()));
  } else {
    RM_TRY(push_rows(d, n, l2g9));
  }
  for (int q = 0; q < 3 * n; q++) d->euler.push_back(euler3 ? euler3[q] : 0.0);
  return AMCA_OK;
}

amca_status amca_remelt_set_orientation_source(amca_remelt* d, amca_orientation_source src,
                                               void* user) {
  if (!d) return rm_fail(AMCA_ERR_INVALID, "null domain");
  d->source = src;
  d->source_user = user;
  d->rng.reset();
  return AMCA_OK;
}

amca_status amca_remelt_draw_sites_after(amca_remelt* d, uint32_t seed, int64_t skip_draws,
                                         double ns, double dTs_mean, double dTs_sd, double nv,
                                         double dTv_mean, double dTv_sd, int32_t* n_surface_new,
                                         int32_t* n_bulk_new) {
  if (!d || skip_draws < 0) return rm_fail(AMCA_ERR_INVALID, "bad arguments");
  amca_host::Grid g{d->P.nx, d->P.ny, d->P.nz};
  auto rng = std::make_unique<amca_host::GlibcRand>(seed);
  for (int64_t q = 0; q < skip_draws; q++) (void)rng->next();
  // create_nucleation_sites merges into the list already held (the seeds of a
  // loaded microstructure file): a draw on a listed cell only lowers its value
  std::unordered_map<int, double> cell_dt;
  std::vector<int> sites;
  for (size_t q = 0; q < d->site_cells.size(); q++) {
    if (!cell_dt.emplace(d->site_cells[q], d->site_dts[q]).second || d->site_dts[q] < 0)
      return rm_fail(AMCA_ERR_INVALID, "sites to merge with must be distinct cells with undercooling >= 0");
    sites.push_back(d->site_cells[q]);
  }
  const double h = d->P.h;
  amca_host::SiteDraw ds, db;
  {
    std::vector<int> ids;
    amca_host::surface_cells(g, amca_host::SurfaceFaces::kReferenceSix, ids);
    ds = amca_host::create_nucleation_sites((int)ids.size(), [&](int i) { return ids[i]; }, h * h,
                                            ns, dTs_mean, dTs_sd, *rng, cell_dt, sites);
  }
  db = amca_host::create_nucleation_sites((int)(g.nx * (long long)g.ny * g.nz),
                                          amca_host::BulkCellAt{g}, h * h * h, nv, dTv_mean,
                                          dTv_sd, *rng, cell_dt, sites);
  std::vector<double> dts(sites.size());
  for (size_t q = 0; q < sites.size(); q++) dts[q] = cell_dt[sites[q]];
  RM_TRY(amca_remelt_set_sites(d, (int32_t)sites.size(), sites.data(), dts.data()));
  d->source = nullptr;
  d->source_user = nullptr;
  d->rng = std::move(rng);
  if (n_surface_new) *n_surface_new = ds.n_new;
  if (n_bulk_new) *n_bulk_new = db.n_new;
  return AMCA_OK;
}

amca_status amca_remelt_draw_sites(amca_remelt* d, uint32_t seed, double ns, double dTs_mean,
                                   double dTs_sd, double nv, double dTv_mean, double dTv_sd,
                                   int32_t* n_sites) {
  int32_t a = 0, b = 0;
  RM_TRY(amca_remelt_draw_sites_after(d, seed, 0, ns, dTs_mean, dTs_sd, nv, dTv_mean, dTv_sd, &a,
                                      &b));
  if (n_sites) *n_sites = (int32_t)d->site_cells.size();
  return AMCA_OK;
}

amca_status amca_remelt_sites(amca_remelt* d, int32_t cap, int32_t* cell, double* dT_crit,
                              int32_t* n) {
  if (!d || !n) return rm_fail(AMCA_ERR_INVALID, "null argument");
  *n = (int32_t)d->site_cells.size();
  const int m = std::max(0, std::min(cap, *n));
  if (m > 0 && cell) std::memcpy(cell, d->site_cells.data(), sizeof(int32_t) * m);
  if (m > 0 && dT_crit) std::memcpy(dT_crit, d->site_dts.data(), sizeof(double) * m);
  return AMCA_OK;
}

amca_status amca_remelt_euler(amca_remelt* d, int32_t cap, double* euler3, int32_t* n) {
  if (!d || !n) return rm_fail(AMCA_ERR_INVALID, "null argument");
  amca::RmCounters c;
  RM_TRY(rm_counters(d, &c));
  *n = c.n_orient;
  const int m = std::max(0, std::min(std::min(cap, c.n_orient), (int)(d->euler.size() / 3)));
  if (euler3 && m > 0) std::memcpy(euler3, d->euler.data(), sizeof(double) * 3 * m);
  return AMCA_OK;
}

amca_status amca_rosenthal_frames(int64_t n_nodes, const double* node_xyz, int32_t n_frames,
                                  double t0, double dt, double P, double eta, double v, double k,
                                  double alpha, double T0, double x_start, double y_c,
                                  double z_top, double r_min, double* out) {
  if (n_nodes < 1 || n_frames < 1 || !node_xyz || !out || !(k > 0) || !(alpha > 0))
    return rm_fail(AMCA_ERR_INVALID, "bad arguments");
  amca_host::RosenthalParams p{P, eta, v, k, alpha, T0, x_start, y_c, z_top, r_min};
  amca_host::rosenthal_frames(n_nodes, node_xyz, n_frames, t0, dt, p, out);
  return AMCA_OK;
}

amca_status amca_write_fe_frames(const char* path, int32_t title_lines, int32_t fnx, int32_t fny,
                                 int32_t fnz, const double* node_xyz, int32_t n_frames,
                                 const double* times, const double* theta) {
  if (!path || title_lines < 0 || fnx < 2 || fny < 2 || fnz < 2 || n_frames < 1 || !node_xyz ||
      !times || !theta)
    return rm_fail(AMCA_ERR_INVALID, "bad arguments");
  if ((long long)fnx * fny * fnz > INT_MAX) return rm_fail(AMCA_ERR_INVALID, "mesh too large");
  try {
    amca_host::write_fe_text(path, title_lines, fnx, fny, fnz, node_xyz, n_frames, times, theta);
  } catch (const std::exception& e) {
    return rm_fail(AMCA_ERR_IO, std::string(path) + ": " + e.what());
  }
  return AMCA_OK;
}

amca_status amca_voronoi_substrate(int32_t nx, int32_t ny, int32_t nz, double h, double x0,
                                   double y0, double z0, int32_t n_seeds, uint64_t seed,
                                   int32_t* label, double* euler3) {
  if (nx < 1 || ny < 1 || nz < 1 || n_seeds < 1 || !label || !euler3 || !(h > 0))
    return rm_fail(AMCA_ERR_INVALID, "bad arguments");
  amca_host::voronoi_labels(nx, ny, nz, h, x0, y0, z0, n_seeds, seed, label, euler3);
  return AMCA_OK;
}

amca_status amca_remelt_set_sites(amca_remelt* d, int32_t n, const int32_t* cell,
                                  const double* dT_crit) {
  if (!d || n < 0 || (n > 0 && (!cell || !dT_crit))) return rm_fail(AMCA_ERR_INVALID, "bad sites");
  for (int i = 0; i < n; i++) {
    if (cell[i] < 0 || cell[i] >= d->P.NG) return rm_fail(AMCA_ERR_INVALID, "site out of range");
    if (i > 0 && cell[i] <= cell[i - 1]) return rm_fail(AMCA_ERR_INVALID, "sites must ascend");
  }
  d->site_cells.assign(cell, cell + n);
  d->site_dts.assign(dT_crit, dT_crit + n);
  d->release(d->B.site_cell);
  d->release(d->B.site_dt);
  RM_TRY(d->alloc(&d->B.site_cell, (size_t)std::max(n, 1)));
  RM_TRY(d->alloc(&d->B.site_dt, (size_t)std::max(n, 1)));
  if (n > 0) {
    RM_HIP(hipMemcpyAsync(d->B.site_cell, cell, sizeof(int32_t) * n, hipMemcpyHostToDevice, d->stream));
    RM_HIP(hipMemcpyAsync(d->B.site_dt, dT_crit, sizeof(double) * n, hipMemcpyHostToDevice, d->stream));
  }
  amca::RmCounters c;
  RM_TRY(rm_counters(d, &c));
  c.n_sites = n;
  RM_TRY(rm_put_counters(d, c));
  d->site_blocks = std::max(1, std::min((n + 255) / 256, (int)amca::kMaxBlocks));
  return AMCA_OK;
}

amca_status amca_remelt_enable_log(amca_remelt* d, int32_t capacity) {
  if (!d || capacity < 0) return rm_fail(AMCA_ERR_INVALID, "bad arguments");
  RM_HIP(hipStreamSynchronize(d->stream));
  d->release(d->B.log);
  d->B.log = nullptr;
  if (capacity > 0) RM_TRY(d->alloc(&d->B.log, (size_t)capacity));
  amca::RmCounters c;
  RM_TRY(rm_counters(d, &c));
  c.log_cap = capacity;
  return rm_put_counters(d, c);
}

amca_status amca_remelt_read_log(amca_remelt* d, int32_t cap, amca_remelt_step_record* rows,
                                 int32_t* n_rows) {
  if (!d || !n_rows) return rm_fail(AMCA_ERR_INVALID, "null argument");
  amca::RmCounters c;
  RM_TRY(rm_counters(d, &c));
  const long long avail = std::min<long long>(c.step, c.log_cap);
  *n_rows = (int32_t)avail;
  const long long n = std::min<long long>(avail, std::max(cap, 0));
  if (n > 0 && rows && d->B.log) {
    std::vector<amca::RmLogRow> tmp((size_t)n);
    RM_HIP(hipMemcpy(tmp.data(), d->B.log, sizeof(amca::RmLogRow) * n, hipMemcpyDeviceToHost));
    for (long long i = 0; i < n; i++) {
      rows[i].time_before = tmp[i].time_before;
      rows[i].dt = tmp[i].dt;
      rows[i].n_nucleated = tmp[i].n_nucleated;
      rows[i].n_captured = tmp[i].n_captured;
      rows[i].n_fresh = tmp[i].n_fresh;
      rows[i].n_flagged = tmp[i].n_flagged;
      rows[i].n_fallback = tmp[i].n_fallback;
      rows[i].n_box = tmp[i].n_box;
    }
  }
  return AMCA_OK;
}

amca_status amca_pow_selftest(int64_t n, const double* x, double y, double* out) {
  if (n < 0 || (n > 0 && (!x || !out))) return rm_fail(AMCA_ERR_INVALID, "bad arguments");
  if (!amca_libm::pow_exponent_supported(y))
    return rm_fail(AMCA_ERR_INVALID, "exponent outside [2^-65, 2^63) in magnitude");
  for (int64_t i = 0; i < n; i++)
    if (!(x[i] > 0) || !(x[i] < HUGE_VAL))
      return rm_fail(AMCA_ERR_INVALID, "bases must be finite and positive");
  if (n == 0) return AMCA_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return rm_fail(AMCA_ERR_NO_DEVICE, "no HIP device visible");
  double *dx = nullptr, *dout = nullptr;
  hipError_t e = hipMalloc(&dx, sizeof(double) * n);
  if (e == hipSuccess) e = hipMalloc(&dout, sizeof(double) * n);
  if (e == hipSuccess) e = hipMemcpy(dx, x, sizeof(double) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = amca::launch_pow_selftest(n, dx, y, dout, 0);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost);
  if (dx) (void)hipFree(dx);
  if (dout) (void)hipFree(dout);
  if (e != hipSuccess) return rm_fail(AMCA_ERR_HIP, std::string("pow self-test: ") + hipGetErrorString(e));
  return AMCA_OK;
}

}  // extern "C"

namespace {

// One domain's run, in rounds: plan (the next batch of steps, after any rows
// a stopped step asked for), enqueue, collect the counters.
struct RunState {
  long long start = 0, batch = 4, k = 0;
  bool done = false;
  amca::RmCounters c{};
};

amca_status run_begin(amca_remelt* d, RunState& st) {
  if (!d || !d->fe_ready) return rm_fail(AMCA_ERR_INVALID, "FE mesh not set");
  RM_TRY(rm_counters(d, &st.c));
  st.start = st.c.step;
  return AMCA_OK;
}

amca_status run_plan(amca_remelt* d, RunState& st, int64_t max_steps) {
  amca::RmCounters& c = st.c;
  st.k = 0;
  if (c.err == 1) return rm_fail(AMCA_ERR_CAPACITY, "orientation table exhausted");
  if (c.err == 2) return rm_fail(AMCA_ERR_INVALID, "FE time ratio below zero");
  const long long taken = c.step - st.start;
  if (c.exhausted || !(c.tcafe < d->desc.final_time) || (max_steps >= 0 && taken >= max_steps)) {
    st.done = true;
    return AMCA_OK;
  }
  if (c.need_rows) {
    // a step stopped: its armed sites could need more rows than were pushed
    const long long want = (long long)c.n_orient + c.armed[c.step & 1];
    if (want > d->P.orient_cap) return rm_fail(AMCA_ERR_CAPACITY, "orientation table exhausted");
    RM_TRY(ensure_rows(d, want + d->row_reserve));
    if (d->rows_pushed < want)
      return rm_fail(AMCA_ERR_CAPACITY, "orientation rows exhausted and no orientation source set");
    RM_TRY(rm_counters(d, &c));
    c.need_rows = 0;
    c.armed[c.step & 1] = 0;
    RM_TRY(rm_put_counters(d, c));
  }
  st.k = st.batch;
  if (max_steps >= 0) st.k = std::min<long long>(st.k, max_steps - taken);
  if (c.n_sites > 0 && (d->source || d->rng))
    RM_TRY(ensure_rows(d, (long long)c.n_orient + d->row_reserve));
  return AMCA_OK;
}

amca_status run_enqueue(amca_remelt* d, RunState& st) {
  if (st.k <= 0) return AMCA_OK;
  if (d->timing) {
    for (long long q = 0; q < st.k; q++) RM_TRY(rm_enqueue_step(d));
  } else {
    RM_TRY(rm_enqueue_steps(d, st.k));
  }
  return AMCA_OK;
}

amca_status run_collect(amca_remelt* d, RunState& st) {
  if (st.k <= 0) return AMCA_OK;
  RM_TRY(rm_counters(d, &st.c));
  st.batch = std::min<long long>(st.batch * 2, 64);
  return AMCA_OK;
}

amca_status run_end(amca_remelt* d, RunState& st, amca_remelt_counters* out) {
  // cells outside the last step's box take that step's T and cooling
  if (st.c.step > st.start) {
    RM_HIP(amca::rm_launch_complete(d->P, d->B, d->stream));
    RM_HIP(hipStreamSynchronize(d->stream));
  }
  rm_fill(d, st.c, out);
  return AMCA_OK;
}

}  // namespace

extern "C" {

amca_status amca_remelt_run(amca_remelt* d, int64_t max_steps, amca_remelt_counters* out) {
  RunState st;
  RM_TRY(run_begin(d, st));
  while (true) {
    RM_TRY(run_plan(d, st, max_steps));
    if (st.done) break;
    RM_TRY(run_enqueue(d, st));
    RM_TRY(run_collect(d, st));
  }
  return run_end(d, st, out);
}

amca_status amca_remelt_run_many(amca_remelt* const* ds, int32_t n, int64_t max_steps,
                                 amca_remelt_counters* outs) {
  if (!ds || n < 1) return rm_fail(AMCA_ERR_INVALID, "no domains");
  for (int i = 0; i < n; i++)
    for (int j = 0; j < i; j++)
      if (ds[i] == ds[j]) return rm_fail(AMCA_ERR_INVALID, "a domain is listed twice");
  std::vector<RunState> st((size_t)n);
  for (int i = 0; i < n; i++) RM_TRY(run_begin(ds[i], st[i]));
  while (true) {
    int live = 0;
    // every domain's next batch is queued on its own stream before any is waited for
    for (int i = 0; i < n; i++) {
      if (st[i].done) continue;
      RM_TRY(run_plan(ds[i], st[i], max_steps));
      if (st[i].done) continue;
      RM_TRY(run_enqueue(ds[i], st[i]));
      live++;
    }
    if (!live) break;
    for (int i = 0; i < n; i++)
      if (!st[i].done) RM_TRY(run_collect(ds[i], st[i]));
  }
  for (int i = 0; i < n; i++) RM_TRY(run_end(ds[i], st[i], outs ? &outs[i] : nullptr));
  return AMCA_OK;
}

amca_status amca_remelt_set_window(amca_remelt* d, int32_t on) {
  if (!d) return rm_fail(AMCA_ERR_INVALID, "null domain");
  d->P.window = on ? 1 : 0;
  amca::RmCounters c;
  RM_TRY(rm_counters(d, &c));
  rm_restart_box(c);
  RM_TRY(rm_put_counters(d, c));
  return AMCA_OK;
}

amca_status amca_remelt_set_row_reserve(amca_remelt* d, int32_t rows) {
  if (!d || rows < 0) return rm_fail(AMCA_ERR_INVALID, "bad row reserve");
  d->row_reserve = rows;
  return AMCA_OK;
}

amca_status amca_remelt_set_graph_steps(amca_remelt* d, int32_t steps) {
  if (!d || steps < 0 || steps > 1024) return rm_fail(AMCA_ERR_INVALID, "bad graph step count");
  RM_HIP(hipStreamSynchronize(d->stream));
  d->drop_graph();
  d->graph_steps = steps;
  return AMCA_OK;
}

amca_status amca_remelt_last_box(amca_remelt* d, int32_t* box6) {
  if (!d || !box6) return rm_fail(AMCA_ERR_INVALID, "null argument");
  amca::RmCounters c;
  RM_TRY(rm_counters(d, &c));
  if (c.last_step < 0) return rm_fail(AMCA_ERR_INVALID, "no step has run");
  const amca::RmBox& b = c.box[c.last_step & 1];
  for (int a = 0; a < 3; a++) {
    box6[a] = b.lo[a];
    box6[3 + a] = b.hi[a];
  }
  return AMCA_OK;
}

amca_status amca_remelt_get_counters(amca_remelt* d, amca_remelt_counters* out) {
  if (!d || !out) return rm_fail(AMCA_ERR_INVALID, "null argument");
  amca::RmCounters c;
  RM_TRY(rm_counters(d, &c));
  rm_fill(d, c, out);
  return AMCA_OK;
}

amca_status amca_remelt_download(amca_remelt* d, int32_t* orient, int32_t* grain, double* T,
                                 uint8_t* melted, uint8_t* cooling, int32_t grain_cap,
                                 double* genv, int32_t* gcell, int32_t* gorient, int32_t* n_grains) {
  if (!d) return rm_fail(AMCA_ERR_INVALID, "null domain");
  const size_t ng = (size_t)d->P.NG;
  amca::RmCounters c;
  RM_TRY(rm_counters(d, &c));
  if (orient) RM_HIP(hipMemcpy(orient, d->B.oid, sizeof(int32_t) * ng, hipMemcpyDeviceToHost));
  if (grain) RM_HIP(hipMemcpy(grain, d->B.gid, sizeof(int32_t) * ng, hipMemcpyDeviceToHost));
  if (T) RM_HIP(hipMemcpy(T, d->B.T, sizeof(double) * ng, hipMemcpyDeviceToHost));
  if (melted) RM_HIP(hipMemcpy(melted, d->B.melted, ng, hipMemcpyDeviceToHost));
  if (cooling) RM_HIP(hipMemcpy(cooling, d->B.cooling, ng, hipMemcpyDeviceToHost));
  if (n_grains) *n_grains = c.n_grains;
  const size_t m = (size_t)std::max(0, std::min(grain_cap, c.n_grains));
  if (m > 0) {
    if (genv) RM_HIP(hipMemcpy(genv, d->B.genv, sizeof(double) * 4 * m, hipMemcpyDeviceToHost));
    if (gcell || gorient) {
      std::vector<int32_t> co(2 * m);
      RM_HIP(hipMemcpy(co.data(), d->B.gco, sizeof(int32_t) * 2 * m, hipMemcpyDeviceToHost));
      for (size_t g = 0; g < m; g++) {
        if (gcell) gcell[g] = co[2 * g];
        if (gorient) gorient[g] = co[2 * g + 1];
      }
    }
  }
  return AMCA_OK;
}

amca_status amca_remelt_use_surrogate(amca_remelt* d, amca_mlp* m, const amca_surrogate_desc* sd,
                                      const amca_scan* scan, double band, double resid) {
  if (!d || !d->fe_ready) return rm_fail(AMCA_ERR_INVALID, "FE mesh not set");
  if (!m || !sd || !scan || !(band >= 0) || !(resid >= 0))
    return rm_fail(AMCA_ERR_INVALID, "bad surrogate arguments");
  amca::RmParams& P = d->P;
  if (sd->fnx != P.fnx || sd->fny != P.fny || sd->fnz != P.fnz)
    return rm_fail(AMCA_ERR_INVALID, "surrogate mesh differs from the domain's FE mesh");
  // the FE mesh must be the uniform lattice the description names
  for (int n = 0; n < P.nodes; n++) {
    const int i = n % P.fnx, j = (n / P.fnx) % P.fny, k = n / (P.fnx * P.fny);
    const double e[3] = {sd->fx0 + i * sd->fdx, sd->fy0 + j * sd->fdx, sd->fz0 + k * sd->fdx};
    for (int q = 0; q < 3; q++)
      if (std::fabs(d->node_xyz[3 * (size_t)n + q] - e[q]) > 1e-9 * sd->fdx)
        return rm_fail(AMCA_ERR_INVALID, "the FE mesh is not the uniform lattice of the description");
  }
  RM_HIP(hipStreamSynchronize(d->stream));
  // time span of the frames
  std::vector<double> ft(P.n_frames);
  RM_HIP(hipMemcpy(ft.data(), d->B.frame_time, sizeof(double) * P.n_frames, hipMemcpyDeviceToHost));
  const double tmax = ft.back();
  const double h = P.h;
  const double xlo = std::min(P.x0 - 1.0 * h, sd->fx0), xhi = std::max(P.x0 + (P.nx + 1) * h, sd->fx0 + (P.fnx - 1) * sd->fdx);
  const double a = xlo - scan->x_start - std::max(0.0, scan->v) * tmax - 2 * h;
  const double b = xhi - scan->x_start - std::min(0.0, scan->v) * tmax + 2 * h;
  const int nxi = (int)std::ceil((b - a) / h) + 2;
  d->release(const_cast<float*>(d->B.sur_cell));
  d->release(const_cast<float*>(d->B.sur_node));
  d->release(d->B.node_bad);
  d->release(d->B.elem_fb);
  d->release(d->B.sur_stats);
  d->release(d->B.fb_slots);
  float *tc = nullptr, *tn = nullptr;
  unsigned char *nb = nullptr, *eb = nullptr;
  double* stats = nullptr;
  RM_TRY(d->alloc(&tc, (size_t)nxi * P.nyG * P.nzG));
  RM_TRY(d->alloc(&tn, (size_t)nxi * P.fny * P.fnz));
  RM_TRY(d->alloc(&nb, (size_t)P.nodes));
  RM_TRY(d->alloc(&eb, 2 * (size_t)(P.fnx - 1) * (P.fny - 1) * (P.fnz - 1)));
  RM_TRY(d->alloc(&stats, 8));
  int* slots = nullptr;
  RM_TRY(d->alloc(&slots, 256));
  RM_HIP(hipMemset(slots, 0, sizeof(int) * 256));
  // cell lattice: y, z at the cell centres (ghosts included); node lattice at the nodes
  SurLattice gc{a, h, nxi, P.y0 - 0.5 * h, h, P.nyG, P.z0 - 0.5 * h, h, P.nzG};
  SurLattice gn{a, h, nxi, sd->fy0, sd->fdx, P.fny, sd->fz0, sd->fdx, P.fnz};
  amca_status st = surrogate_eval_lattice(m, *sd, *scan, gc, tc);
  if (st == AMCA_OK) st = surrogate_eval_lattice(m, *sd, *scan, gn, tn);
  if (st != AMCA_OK) return rm_fail(st, std::string("surrogate evaluation: ") + surrogate_error());
  d->B.sur_cell = tc;
  d->B.sur_node = tn;
  d->B.node_bad = nb;
  d->B.elem_fb = eb;
  d->B.sur_stats = stats;
  d->B.fb_slots = slots;
  P.sur = 1;
  P.sur_xi0 = a;
  P.sur_dxi = h;
  P.sur_nxi = nxi;
  P.sur_nxi_nodes = nxi;
  P.sur_x_start = scan->x_start;
  P.sur_v = scan->v;
  P.sur_band = band;
  P.sur_resid = resid;
  P.fx0 = sd->fx0;
  P.fy0 = sd->fy0;
  P.fz0 = sd->fz0;
  P.fdx = sd->fdx;
  // hot region of the cell table (xi index, j, k), for the active box
  {
    int* hb = nullptr;
    RM_TRY(d->alloc(&hb, 6));
    const int init[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN};
    RM_HIP(hipMemcpy(hb, init, sizeof(init), hipMemcpyHostToDevice));
    RM_HIP(amca::rm_launch_sur_hot(P, d->B, hb, d->stream));
    int got[6];
    RM_HIP(hipMemcpyAsync(got, hb, sizeof(got), hipMemcpyDeviceToHost, d->stream));
    RM_HIP(hipStreamSynchronize(d->stream));
    d->release(hb);
    P.sur_hot[0] = got[0];
    P.sur_hot[1] = got[3];
    P.sur_hot[2] = got[1];
    P.sur_hot[3] = got[4];
    P.sur_hot[4] = got[2];
    P.sur_hot[5] = got[5];
  }
  amca::RmCounters c;
  RM_TRY(rm_counters(d, &c));
  rm_restart_box(c);
  RM_TRY(rm_put_counters(d, c));
  return AMCA_OK;
}

amca_status amca_remelt_surrogate_error(amca_remelt* d, double* out6) {
  if (!d || !out6 || !d->P.sur) return rm_fail(AMCA_ERR_INVALID, "no surrogate in use");
  RM_HIP(hipMemsetAsync(d->B.sur_stats, 0, sizeof(double) * 8, d->stream));
  RM_HIP(amca::rm_launch_sur_error(d->P, d->B, d->stream));
  double s[8];
  RM_HIP(hipMemcpyAsync(s, d->B.sur_stats, sizeof(s), hipMemcpyDeviceToHost, d->stream));
  RM_HIP(hipStreamSynchronize(d->stream));
  out6[0] = s[1] > 0 ? std::sqrt(s[0] / s[1]) : 0.0;  // RMS over all cells with material
  out6[1] = s[1];
  out6[2] = s[3] > 0 ? std::sqrt(s[2] / s[3]) : 0.0;  // RMS over the cells the surrogate serves
  out6[3] = s[3];
  out6[4] = s[4];                                     // max |error| over those cells
  out6[5] = s[5];                                     // cells on the FE path
  return AMCA_OK;
}

amca_status amca_remelt_theta(amca_remelt* d, double* theta) {
  if (!d || !d->fe_ready || !theta) return rm_fail(AMCA_ERR_INVALID, "FE mesh not set");
  RM_HIP(hipStreamSynchronize(d->stream));
  RM_HIP(hipMemcpy(theta, d->B.theta, sizeof(double) * d->P.nodes, hipMemcpyDeviceToHost));
  return AMCA_OK;
}

}  // extern "C"
