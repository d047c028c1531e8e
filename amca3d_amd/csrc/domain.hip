// C ABI of one device-resident CA domain (include/amca3d.h).
//
// The reference keeps the grid as raw host arrays owned by
// CellularAutomataManager (CellularAutomataManager.cpp:303-326) and one heap
// Grain per solid cell.  Here the grid and the per-cell envelopes live in HBM
// for the lifetime of the handle, a step is a fixed sequence of six kernel
// launches that reads every scalar it needs from device memory, and the host
// only reads the counters back between batches of steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/amca3d.h"
#include "env.h"
#include "plain_step.h"

namespace {

thread_local std::string g_last_error;

amca_status fail(amca_status st, const std::string& msg) {
  g_last_error = msg;
  return st;
}

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(AMCA_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));     \
  } while (0)

enum KernelId {
  K_NUCLEATE = 0, K_CLAIM, K_RESOLVE, K_EMIT, K_TAIL, K_IDSCAN, K_COMPACT, K_CNT, K_COUNT
};
// named as the kernels are, so event times line up with a kernel trace
const char* kKernelNames[K_COUNT] = {"k_nucleate", "k_claim",   "k_resolve", "k_emit",
                                     "k_tail",     "k_idscan",  "k_compact", "k_count"};

}  // namespace

struct amca_domain {
  int device = 0;
  hipStream_t stream = nullptr;
  amca_domain_desc desc{};
  amca::DevParams P{};
  amca::DevBuffers B{};
  int site_cap = 0;
  int orient_pushed = 0;
  int nuc_blocks = 1;
  int list_blocks = 1;
  int emit_blocks = 1;
  int ids_blocks = 1;
  int cmp_blocks = 1;
  bool timing = false;
  // event pairs recorded around timed launches, resolved when times are read
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, size_t>> ev_pending;  // (kernel id, index of first event)
  size_t ev_used = 0;
  double kernel_ms[K_COUNT] = {};
  int64_t kernel_launches[K_COUNT] = {};
  std::vector<void*> allocs;
  // graph_steps steps captured once and replayed by run / run_fixed when
  // kernel timing is off; dropped whenever a launch argument changes.
  // AMCA_DOMAIN_GRAPH_STEPS sets the count (0: no graphs, every step launched
  // on its own, for tools that do not follow graph launches).
  int graph_steps = 16;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;

  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
  }

  template <class T>
  amca_status alloc(T** p, size_t count) {
    void* q = nullptr;
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess)
      return fail(AMCA_ERR_HIP, std::string("hipMalloc(") + std::to_string(bytes) +
                                    "): " + hipGetErrorString(e));
    allocs.push_back(q);
    *p = static_cast<T*>(q);
    return AMCA_OK;
  }

  ~amca_domain() {
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graph();
    for (void* p : allocs) (void)hipFree(p);
    for (auto& e : ev_pool)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

#define AMCA_TRY(expr)                  \
  do {                                  \
    amca_status s_ = (expr);            \
    if (s_ != AMCA_OK) return s_;       \
  } while (0)

amca_status read_counters(amca_domain* d, amca::DevCounters* c) {
  HIP_TRY(hipMemcpyAsync(c, d->B.ctr, sizeof(*c), hipMemcpyDeviceToHost, d->stream));
  HIP_TRY(hipStreamSynchronize(d->stream));
  return AMCA_OK;
}

void fill_counters(const amca_domain* d, const amca::DevCounters& c, amca_counters* out) {
  if (!out) return;
  out->step = c.step;
  out->time = c.t;
  out->dt = c.dt;
  out->n_liquid = d->P.N - c.n_solid;
  out->n_grains = c.n_grains;
  out->n_orient = c.n_orient;
  out->n_active = c.nA_in;
  out->n_sites = c.n_sites;
  out->done = (c.done || !(c.t < d->desc.final_time)) ? 1 : 0;
  out->max_wave = c.max_wave;
}

amca_status resolve_events(amca_domain* d) {
  if (d->ev_pending.empty()) return AMCA_OK;
  HIP_TRY(hipStreamSynchronize(d->stream));
  for (auto& p : d->ev_pending) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, d->ev_pool[p.second], d->ev_pool[p.second + 1]));
    d->kernel_ms[p.first] += ms;
    d->kernel_launches[p.first] += 1;
  }
  d->ev_pending.clear();
  d->ev_used = 0;
  return AMCA_OK;
}

template <class F>
amca_status timed(amca_domain* d, int k, F&& launch) {
  if (!d->timing) {
    HIP_TRY(launch());
    return AMCA_OK;
  }
  if (d->ev_used + 2 > d->ev_pool.size()) {
    if (d->ev_pool.size() >= 1u << 16) AMCA_TRY(resolve_events(d));
    while (d->ev_used + 2 > d->ev_pool.size()) {
      hipEvent_t e;
      HIP_TRY(hipEventCreate(&e));
      d->ev_pool.push_back(e);
    }
  }
  const size_t i = d->ev_used;
  d->ev_used += 2;
  HIP_TRY(hipEventRecord(d->ev_pool[i], d->stream));
  HIP_TRY(launch());
  HIP_TRY(hipEventRecord(d->ev_pool[i + 1], d->stream));
  d->ev_pending.emplace_back(k, i);
  return AMCA_OK;
}

// The fixed launch sequence of one step.
amca_status enqueue_step(amca_domain* d) {
  const amca::DevParams& P = d->P;
  const amca::DevBuffers& B = d->B;
  hipStream_t s = d->stream;
  AMCA_TRY(timed(d, K_NUCLEATE, [&] { return amca::launch_nucleate(P, B, d->nuc_blocks, s); }));
  AMCA_TRY(timed(d, K_CNT, [&] { return amca::launch_count(P, B, d->cmp_blocks, s); }));
  AMCA_TRY(timed(d, K_COMPACT, [&] { return amca::launch_compact(P, B, d->cmp_blocks, s); }));
  for (int wave = 0; wave < 2; wave++) {
    AMCA_TRY(timed(d, K_CLAIM, [&] { return amca::launch_claim(P, B, wave, d->list_blocks, s); }));
    AMCA_TRY(timed(d, K_RESOLVE,
                   [&] { return amca::launch_resolve(P, B, wave, d->emit_blocks, s); }));
    AMCA_TRY(timed(d, K_IDSCAN,
                   [&] { return amca::launch_idscan(P, B, wave, d->emit_blocks, s); }));
    AMCA_TRY(timed(d, K_EMIT, [&] { return amca::launch_emit(P, B, wave, d->emit_blocks, s); }));
  }
  AMCA_TRY(timed(d, K_TAIL, [&] { return amca::launch_tail(P, B, s); }));
  return AMCA_OK;
}

amca_status capture_steps(amca_domain* d) {
  d->drop_graph();
  HIP_TRY(hipStreamBeginCapture(d->stream, hipStreamCaptureModeThreadLocal));
  amca_status st = AMCA_OK;
  for (int q = 0; q < d->graph_steps && st == AMCA_OK; q++) st = enqueue_step(d);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(d->stream, &g);
  if (st != AMCA_OK || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return st != AMCA_OK ? st : fail(AMCA_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
  }
  d->graph = g;
  HIP_TRY(hipGraphInstantiate(&d->gexec, g, nullptr, nullptr, 0));
  return AMCA_OK;
}

// n steps: whole graphs of graph_steps steps, then single launches (every
// launch is single, with its events, when kernel timing is on).  Steps past
// termination do nothing, so a graph may run beyond the last live step.
amca_status enqueue_steps(amca_domain* d, long long n) {
  if (!d->timing && d->graph_steps > 0) {
    for (; n >= d->graph_steps; n -= d->graph_steps) {
      if (!d->gexec) AMCA_TRY(capture_steps(d));
      HIP_TRY(hipGraphLaunch(d->gexec, d->stream));
    }
  }
  for (; n > 0; n--) AMCA_TRY(enqueue_step(d));
  return AMCA_OK;
}

// Every live site can nucleate once and takes the next orientation row, so a
// run may start only when the pushed rows cover the orientations in use plus
// one per live site.  The kernels never read a row that was not pushed.
amca_status check_rows(amca_domain* d, const amca::DevCounters& c) {
  if ((long long)c.n_orient + c.n_sites > d->orient_pushed)
    return fail(AMCA_ERR_CAPACITY,
                "orientation rows pushed (" + std::to_string(d->orient_pushed) +
                    ") do not cover the orientations in use plus one per live site (" +
                    std::to_string((long long)c.n_orient + c.n_sites) +
                    "): push more rows with amca_domain_push_orientations");
  return AMCA_OK;
}

amca_status check_err(amca_domain* d, const amca::DevCounters& c) {
  if (c.err)
    return fail(AMCA_ERR_CAPACITY,
                "orientation table exhausted: push more rows with amca_domain_push_orientations");
  return AMCA_OK;
}

}  // namespace

namespace {
template <class F>
struct ScopeExit {
  F f;
  ~ScopeExit() { f(); }
};
template <class F>
ScopeExit<F> on_scope_exit(F f) {
  return {f};
}
}  // namespace

extern "C" {

const char* amca_last_error(void) { return g_last_error.c_str(); }
const char* amca_version(void) { return "amca3d-mi355x 0.1 (gfx950)"; }

amca_status amca_domain_create(const amca_domain_desc* desc, int32_t orient_capacity,
                               amca_domain** out) {
  if (!desc || !out) return fail(AMCA_ERR_INVALID, "null argument");
  *out = nullptr;
  if (desc->nx < 1 || desc->ny < 1 || desc->nz < 1)
    return fail(AMCA_ERR_INVALID, "domain needs at least one cell per axis");
  const long long nxG = desc->nx + 2LL, nyG = desc->ny + 2LL, nzG = desc->nz + 2LL;
  const long long NG = nxG * nyG * nzG;
  if (NG > (long long)INT_MAX)
    return fail(AMCA_ERR_INVALID, "ghosted grid exceeds 2^31-1 cells (int32 cell index)");
  if (!(desc->h > 0)) return fail(AMCA_ERR_INVALID, "cell size must be positive");
  if (orient_capacity < 0) return fail(AMCA_ERR_INVALID, "negative orientation capacity");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(AMCA_ERR_NO_DEVICE, "no HIP device visible");

  auto* d = new amca_domain();
  HIP_TRY(hipGetDevice(&d->device));
  HIP_TRY(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
  d->desc = *desc;
  amca::DevParams& P = d->P;
  P.nx = desc->nx; P.ny = desc->ny; P.nz = desc->nz;
  P.nxG = (int)nxG; P.nyG = (int)nyG; P.nzG = (int)nzG;
  P.nxyG = (int)(nxG * nyG);
  P.NG = (int)NG;
  P.nbX = (int)((nxG + 3) / 4);
  P.nbY = (int)((nyG + 3) / 4);
  P.nbZ = (int)((nzG + 3) / 4);
  P.fdX = amca::make_fastdiv((unsigned)P.nbX);
  P.fdY = amca::make_fastdiv((unsigned)P.nbY);
  P.N = (long long)desc->nx * desc->ny * desc->nz;
  const long long nB = (long long)P.nbX * P.nbY * P.nbZ;
  if (64 * nB > (long long)INT_MAX) {
    delete d;
    return fail(AMCA_ERR_INVALID, "grid exceeds 2^31-1 brick slots (int32 slot index)");
  }
  P.h = desc->h; P.x0 = desc->x0; P.y0 = desc->y0; P.z0 = desc->z0;
  P.T0 = desc->T0; P.Tm = desc->Tm; P.Tdot = desc->Tdot;
  P.a1 = desc->a1; P.a2 = desc->a2; P.a3 = desc->a3;
  P.trunc = desc->trunc;
  P.tsf = desc->time_step_factor;
  P.max_dt = desc->max_dt;
  P.final_time = desc->final_time;
  P.sqrt2f = (double)sqrtf(2.0f);
  P.orient_cap = orient_capacity;
  {
    int s = 0;
    for (int dk = -1; dk <= 1; dk++)
      for (int dj = -1; dj <= 1; dj++)
        for (int di = -1; di <= 1; di++) {
          if (!dk && !dj && !di) continue;
          P.off[s++] = dk * P.nxyG + dj * P.nxG + di;
        }
  }
  amca::DevBuffers& B = d->B;
  const size_t N = (size_t)P.N;
  amca_status st;
#define ALLOC(ptr, n)                 \
  st = d->alloc(&(ptr), (n));         \
  if (st != AMCA_OK) {                \
    delete d;                         \
    return st;                        \
  }
  ALLOC(B.gid, (size_t)NG);
  ALLOC(B.oid, (size_t)NG);
  ALLOC(B.claim, 64 * (size_t)nB);  // brick order (claim_index)
  ALLOC(B.bmap, 2 * (size_t)nB);
  ALLOC(B.amap, (size_t)nB);
  ALLOC(B.rec, 64 * (size_t)nB);
  ALLOC(B.sL, 64 * (size_t)nB);
  ALLOC(B.skey, 64 * (size_t)nB);
  ALLOC(B.genv, N);
  ALLOC(B.gco, N);
  ALLOC(B.alist, N);
  ALLOC(B.plist, N);
  ALLOC(B.cnt, N + 64);  // the id scan reads whole 16-byte groups past the last id
  ALLOC(B.coff, N);
  ALLOC(B.idtot, amca::kMaxIdsBlocks);
  ALLOC(B.cmptot, amca::kMaxBlocks);
  ALLOC(B.listW[0], N);
  ALLOC(B.listW[1], N);
  ALLOC(B.masks, N);
  ALLOC(B.l2g, 9 * (size_t)std::max(orient_capacity, 1));
  ALLOC(B.lb_flag, (size_t)amca::kLbSites * amca::kMaxBlocks);
  ALLOC(B.lb_agg, (size_t)amca::kLbSites * amca::kMaxBlocks);
  ALLOC(B.lb_pref, (size_t)amca::kLbSites * amca::kMaxBlocks);
  ALLOC(B.ctr, 1);
  ALLOC(B.site_cell[0], 1);
  ALLOC(B.site_cell[1], 1);
  ALLOC(B.site_dt[0], 1);
  ALLOC(B.site_dt[1], 1);
#undef ALLOC
  B.log = nullptr;
  d->site_cap = 1;
  HIP_TRY(hipMemsetAsync(B.lb_flag, 0, sizeof(unsigned) * amca::kLbSites * amca::kMaxBlocks, d->stream));
  HIP_TRY(hipMemsetAsync(B.cnt, 0, N + 64, d->stream));
  amca::DevCounters c{};
  c.t = 0.0;
  HIP_TRY(hipMemcpyAsync(B.ctr, &c, sizeof(c), hipMemcpyHostToDevice, d->stream));
  HIP_TRY(amca::launch_init_grid(P, B, d->stream));
  // Grid sizes: enough blocks to fill 256 CUs several times over.
  int cus = 256;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, d->device) == hipSuccess && prop.multiProcessorCount > 0)
    cus = prop.multiProcessorCount;
  long long want = (P.N + amca::kBlock - 1) / amca::kBlock;
  d->list_blocks = (int)std::max(8LL, std::min<long long>((want + 7) / 8 * 8, 16LL * cus));
  d->emit_blocks = (int)std::max(8LL, std::min<long long>((want + 7) / 8 * 8, std::min(16LL * cus, (long long)amca::kMaxBlocks)));
  d->nuc_blocks = 1;
  // id scan and compaction: eight blocks per CU at most (each block's tiles
  // run one after another), at least 32 K ids / 512 bricks per block
  d->ids_blocks = (int)std::max(1LL, std::min<long long>((P.N + 32767) / 32768,
                                                         std::min(8LL * cus, (long long)amca::kMaxIdsBlocks)));
  P.ids_blocks = d->ids_blocks;
  d->cmp_blocks = (int)std::max(1LL, std::min<long long>((nB + 511) / 512,
                                                         std::min(8LL * cus, (long long)amca::kMaxBlocks)));
  if (amca::env_int("AMCA_DOMAIN_GRAPH_STEPS", 0, 1024, &d->graph_steps) < 0) {
    delete d;
    return fail(AMCA_ERR_INVALID, "AMCA_DOMAIN_GRAPH_STEPS must be an integer in [0, 1024]");
  }
  HIP_TRY(hipStreamSynchronize(d->stream));
  *out = d;
  return AMCA_OK;
}

amca_status amca_domain_destroy(amca_domain* d) {
  if (!d) return AMCA_OK;
  (void)hipStreamSynchronize(d->stream);  // the handle goes either way
  delete d;
  return AMCA_OK;
}

amca_status amca_domain_set_sites(amca_domain* d, int32_t n, const int32_t* cell,
                                  const double* dT_crit) {
  if (!d || n < 0 || (n > 0 && (!cell || !dT_crit))) return fail(AMCA_ERR_INVALID, "bad sites");
  for (int i = 0; i < n; i++) {
    if (cell[i] < 0 || cell[i] >= d->P.NG) return fail(AMCA_ERR_INVALID, "site cell out of range");
    if (i > 0 && cell[i] <= cell[i - 1])
      return fail(AMCA_ERR_INVALID, "sites must be strictly ascending by cell");
  }
  HIP_TRY(hipStreamSynchronize(d->stream));
  d->drop_graph();  // site buffers and the nucleation grid may change
  int cap = std::max(n, 1);
  if (cap > d->site_cap) {
    for (int b = 0; b < 2; b++) {
      for (auto& p : d->allocs) {
        if (p == d->B.site_cell[b] || p == d->B.site_dt[b]) {
          (void)hipFree(p);
          p = nullptr;
        }
      }
    }
    d->allocs.erase(std::remove(d->allocs.begin(), d->allocs.end(), nullptr), d->allocs.end());
    for (int b = 0; b < 2; b++) {
      AMCA_TRY(d->alloc(&d->B.site_cell[b], (size_t)cap));
      AMCA_TRY(d->alloc(&d->B.site_dt[b], (size_t)cap));
    }
    d->site_cap = cap;
  }
  amca::DevCounters c;
  AMCA_TRY(read_counters(d, &c));
  if (n > 0) {
    HIP_TRY(hipMemcpyAsync(d->B.site_cell[c.site_sel], cell, sizeof(int32_t) * n,
                           hipMemcpyHostToDevice, d->stream));
    HIP_TRY(hipMemcpyAsync(d->B.site_dt[c.site_sel], dT_crit, sizeof(double) * n,
                           hipMemcpyHostToDevice, d->stream));
  }
  c.n_sites = n;
  HIP_TRY(hipMemcpyAsync(d->B.ctr, &c, sizeof(c), hipMemcpyHostToDevice, d->stream));
  long long want = ((long long)n + amca::kBlock - 1) / amca::kBlock;
  d->nuc_blocks = (int)std::max(1LL, std::min<long long>(want, amca::kMaxBlocks));
  HIP_TRY(hipStreamSynchronize(d->stream));
  return AMCA_OK;
}

amca_status amca_domain_push_orientations(amca_domain* d, int32_t n, const double* euler3,
                                          const double* l2g9) {
  (void)euler3;
  if (!d || n < 0 || (n > 0 && !l2g9)) return fail(AMCA_ERR_INVALID, "bad orientation rows");
  if (d->orient_pushed + n > d->P.orient_cap)
    return fail(AMCA_ERR_CAPACITY, "orientation table capacity exceeded");
  if (n > 0)
    HIP_TRY(hipMemcpyAsync(d->B.l2g + 9 * (size_t)d->orient_pushed, l2g9, sizeof(double) * 9 * n,
                           hipMemcpyHostToDevice, d->stream));
  d->orient_pushed += n;
  HIP_TRY(hipStreamSynchronize(d->stream));
  return AMCA_OK;
}

amca_status amca_domain_step(amca_domain* d, amca_counters* out) {
  if (!d) return fail(AMCA_ERR_INVALID, "null domain");
  amca::DevCounters c;
  AMCA_TRY(read_counters(d, &c));
  AMCA_TRY(check_rows(d, c));
  AMCA_TRY(enqueue_step(d));
  AMCA_TRY(read_counters(d, &c));
  fill_counters(d, c, out);
  return check_err(d, c);
}

amca_status amca_domain_run(amca_domain* d, int64_t max_steps, amca_counters* out) {
  if (!d) return fail(AMCA_ERR_INVALID, "null domain");
  amca::DevCounters c;
  AMCA_TRY(read_counters(d, &c));
  AMCA_TRY(check_rows(d, c));
  const long long start = c.step;
  long long batch = 8;
  while (true) {
    bool finished = c.done || !(c.t < d->desc.final_time);
    if (finished || c.err) break;
    long long taken = c.step - start;
    if (max_steps >= 0 && taken >= max_steps) break;
    long long k = batch;
    if (max_steps >= 0) k = std::min<long long>(k, max_steps - taken);
    AMCA_TRY(enqueue_steps(d, k));
    AMCA_TRY(read_counters(d, &c));
    batch = std::min<long long>(batch * 2, 64);
  }
  fill_counters(d, c, out);
  return check_err(d, c);
}

amca_status amca_domain_run_fixed(amca_domain* d, int64_t n_steps) {
  if (!d || n_steps < 0) return fail(AMCA_ERR_INVALID, "bad arguments");
  amca::DevCounters c;
  AMCA_TRY(read_counters(d, &c));
  AMCA_TRY(check_rows(d, c));
  AMCA_TRY(enqueue_steps(d, n_steps));
  HIP_TRY(hipStreamSynchronize(d->stream));
  return AMCA_OK;
}

amca_status amca_domain_counters(amca_domain* d, amca_counters* out) {
  if (!d || !out) return fail(AMCA_ERR_INVALID, "null argument");
  amca::DevCounters c;
  AMCA_TRY(read_counters(d, &c));
  fill_counters(d, c, out);
  return check_err(d, c);
}

amca_status amca_domain_enable_log(amca_domain* d, int32_t capacity) {
  if (!d || capacity < 0) return fail(AMCA_ERR_INVALID, "bad arguments");
  HIP_TRY(hipStreamSynchronize(d->stream));
  d->drop_graph();  // the log pointer is a launch argument
  if (d->B.log) {
    for (auto& p : d->allocs)
      if (p == d->B.log) {
        (void)hipFree(p);
        p = nullptr;
      }
    d->allocs.erase(std::remove(d->allocs.begin(), d->allocs.end(), nullptr), d->allocs.end());
    d->B.log = nullptr;
  }
  amca::DevCounters c;
  AMCA_TRY(read_counters(d, &c));
  c.log_cap = capacity;
  if (capacity > 0) AMCA_TRY(d->alloc(&d->B.log, (size_t)capacity));
  HIP_TRY(hipMemcpyAsync(d->B.ctr, &c, sizeof(c), hipMemcpyHostToDevice, d->stream));
  HIP_TRY(hipStreamSynchronize(d->stream));
  return AMCA_OK;
}

amca_status amca_domain_read_log(amca_domain* d, int32_t cap, amca_step_record* rows,
                                 int32_t* n_rows) {
  if (!d || !n_rows) return fail(AMCA_ERR_INVALID, "null argument");
  amca::DevCounters c;
  AMCA_TRY(read_counters(d, &c));
  const long long avail = std::min<long long>(c.step, c.log_cap);
  *n_rows = (int32_t)avail;  // rows available; at most `cap` are copied
  const long long n = std::min<long long>(avail, std::max(cap, 0));
  if (n > 0 && rows && d->B.log) {
    std::vector<amca::StepLogRow> tmp((size_t)n);
    HIP_TRY(hipMemcpy(tmp.data(), d->B.log, sizeof(amca::StepLogRow) * n, hipMemcpyDeviceToHost));
    for (long long i = 0; i < n; i++) {
      rows[i].time_before = tmp[i].time_before;
      rows[i].dt = tmp[i].dt;
      rows[i].n_nucleated = tmp[i].n_nucleated;
      rows[i].n_active_before = tmp[i].n_active_before;
      rows[i].n_captured = tmp[i].n_captured;
      rows[i].n_active_after = tmp[i].n_active_after;
      rows[i].n_liquid = tmp[i].n_liquid;
      rows[i].n_wave1 = tmp[i].n_wave1;
      rows[i].max_wave = tmp[i].max_wave;
      rows[i].n_sites = tmp[i].n_sites;
    }
  }
  return AMCA_OK;
}

amca_status amca_domain_download(amca_domain* d, int32_t* orient, int32_t* grain, double* env,
                                 int32_t site_cap, int32_t* site_cell, double* site_dt,
                                 int32_t* n_sites) {
  if (!d) return fail(AMCA_ERR_INVALID, "null domain");
  const size_t NG = (size_t)d->P.NG;
  amca::DevCounters c;
  AMCA_TRY(read_counters(d, &c));
  // the grain-ID fields and the records by id, materialized from the slots
  if (orient || grain || env) {
    HIP_TRY(amca::launch_materialize(d->P, d->B, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
  }
  if (orient) HIP_TRY(hipMemcpy(orient, d->B.oid, sizeof(int32_t) * NG, hipMemcpyDeviceToHost));
  if (grain) HIP_TRY(hipMemcpy(grain, d->B.gid, sizeof(int32_t) * NG, hipMemcpyDeviceToHost));
  if (env) {
    // grain records by id, scattered to their cells
    std::memset(env, 0, sizeof(double) * 4 * NG);
    const size_t ng = (size_t)std::max(c.n_grains, 0);
    if (ng > 0) {
      std::vector<double> ge(4 * ng);
      std::vector<int32_t> gc(2 * ng);
      HIP_TRY(hipMemcpy(ge.data(), d->B.genv, sizeof(double) * 4 * ng, hipMemcpyDeviceToHost));
      HIP_TRY(hipMemcpy(gc.data(), d->B.gco, sizeof(int32_t) * 2 * ng, hipMemcpyDeviceToHost));
      for (size_t g = 0; g < ng; g++)
        std::memcpy(env + 4 * (size_t)gc[2 * g], &ge[4 * g], sizeof(double) * 4);
    }
  }
  if (n_sites) *n_sites = c.n_sites;
  int m = std::min(c.n_sites, site_cap);
  if (m > 0 && site_cell)
    HIP_TRY(hipMemcpy(site_cell, d->B.site_cell[c.site_sel], sizeof(int32_t) * m, hipMemcpyDeviceToHost));
  if (m > 0 && site_dt)
    HIP_TRY(hipMemcpy(site_dt, d->B.site_dt[c.site_sel], sizeof(double) * m, hipMemcpyDeviceToHost));
  return AMCA_OK;
}

amca_status amca_domain_download_grains(amca_domain* d, int32_t cap, double* env4, int32_t* cell,
                                        int32_t* orient, int32_t* n_grains) {
  if (!d || !n_grains) return fail(AMCA_ERR_INVALID, "null argument");
  amca::DevCounters c;
  AMCA_TRY(read_counters(d, &c));
  *n_grains = c.n_grains;
  const size_t m = (size_t)std::max(0, std::min(cap, c.n_grains));
  if (m == 0) return AMCA_OK;
  HIP_TRY(amca::launch_materialize(d->P, d->B, d->stream));
  HIP_TRY(hipStreamSynchronize(d->stream));
  if (env4) {
    HIP_TRY(hipMemcpy(env4, d->B.genv, sizeof(double) * 4 * m, hipMemcpyDeviceToHost));
  }
  if (cell || orient) {
    std::vector<int32_t> co(2 * m);
    HIP_TRY(hipMemcpy(co.data(), d->B.gco, sizeof(int32_t) * 2 * m, hipMemcpyDeviceToHost));
    for (size_t g = 0; g < m; g++) {
      if (cell) cell[g] = co[2 * g];
      if (orient) orient[g] = co[2 * g + 1];
    }
  }
  return AMCA_OK;
}

amca_status amca_domain_device_fields(amca_domain* d, void** orient, void** grain, void** env,
                                      void** stream) {
  if (!d) return fail(AMCA_ERR_INVALID, "null domain");
  if (orient || grain || env) {
    HIP_TRY(amca::launch_materialize(d->P, d->B, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
  }
  if (orient) *orient = d->B.oid;
  if (grain) *grain = d->B.gid;
  if (env) *env = d->B.genv;
  if (stream) *stream = d->stream;
  return AMCA_OK;
}

amca_status amca_domain_enable_timing(amca_domain* d, int32_t on) {
  if (!d) return fail(AMCA_ERR_INVALID, "null domain");
  AMCA_TRY(resolve_events(d));
  d->timing = on != 0;
  for (int k = 0; k < K_COUNT; k++) {
    d->kernel_ms[k] = 0.0;
    d->kernel_launches[k] = 0;
  }
  return AMCA_OK;
}

int32_t amca_kernel_count(void) { return K_COUNT; }

const char* amca_kernel_name(int32_t i) {
  if (i < 0 || i >= K_COUNT) return "";
  return kKernelNames[i];
}

amca_status amca_domain_kernel_times(amca_domain* d, double* ms, int64_t* launches) {
  if (!d) return fail(AMCA_ERR_INVALID, "null domain");
  AMCA_TRY(resolve_events(d));
  for (int k = 0; k < K_COUNT; k++) {
    if (ms) ms[k] = d->kernel_ms[k];
    if (launches) launches[k] = d->kernel_launches[k];
  }
  return AMCA_OK;
}

amca_status amca_envelope_selftest(int32_t n, const double* env4, const double* l2g9,
                                   const double* cell3, double h, double trunc,
                                   int32_t* inside, double* child4) {
  if (n < 0 || (n > 0 && (!env4 || !l2g9 || !cell3 || !inside || !child4)))
    return fail(AMCA_ERR_INVALID, "bad arguments");
  if (n == 0) return AMCA_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(AMCA_ERR_NO_DEVICE, "no HIP device visible");
  double *de = nullptr, *dr = nullptr, *dc = nullptr, *dout = nullptr;
  int* din = nullptr;
  // device buffers freed on every path, the early error returns included
  const auto release = on_scope_exit([&] {
    for (void* q : {(void*)de, (void*)dr, (void*)dc, (void*)dout, (void*)din})
      if (q) (void)hipFree(q);
  });
  HIP_TRY(hipMalloc(&de, sizeof(double) * 4 * n));
  HIP_TRY(hipMalloc(&dr, sizeof(double) * 9 * n));
  HIP_TRY(hipMalloc(&dc, sizeof(double) * 3 * n));
  HIP_TRY(hipMalloc(&dout, sizeof(double) * 4 * n));
  HIP_TRY(hipMalloc(&din, sizeof(int) * n));
  HIP_TRY(hipMemcpy(de, env4, sizeof(double) * 4 * n, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dr, l2g9, sizeof(double) * 9 * n, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dc, cell3, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
  HIP_TRY(amca::launch_envelope_selftest(n, de, dr, dc, h, trunc, (double)sqrtf(2.0f), din, dout, 0));
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(inside, din, sizeof(int) * n, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(child4, dout, sizeof(double) * 4 * n, hipMemcpyDeviceToHost));
  return AMCA_OK;
}

}  // extern "C"

// ===========================================================================
// Batches: many independent plain domains stepped together.  Every kernel of
// a step is launched once for the whole batch; its blocks are split among the
// domains by a per-kernel (domain, block, count) map, so small domains that
// could not fill the device alone share it.  The steps are captured in a HIP
// graph of `steps_per_graph` steps that is replayed until every domain has
// finished; a finished domain's kernels return at once.  Results per domain
// do not depend on the block split (every kernel's result is independent of
// its block count), so each domain stays identical to its own run.

struct amca_batch {
  std::vector<amca_domain*> doms;
  int device = 0;
  hipStream_t stream = nullptr;
  amca::DevParams* dP = nullptr;
  amca::DevBuffers* dB = nullptr;
  int4* maps = nullptr;  // nuc | list | emit | scan | ids | compaction, back to back
  int4* status = nullptr;
  int4* hstatus = nullptr;  // pinned: two slots of n statuses, one per replay in flight
  hipEvent_t hready[2] = {nullptr, nullptr};
  amca::BatchLaunch L{};
  amca_batch_schedule sched{};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_steps = 0;
  bool use_graph = true;  // false: every replay launches its kernels one by one
  int cus = 256;
  double last_ms_per_step = 0.0;
  // domains still running when the maps were last built (finished domains get
  // no blocks) and the finished count then
  std::vector<char> active;
  int mapped_finished = 0;

  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    graph_steps = 0;
  }
  ~amca_batch() {
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graph();
    if (dP) (void)hipFree(dP);
    if (dB) (void)hipFree(dB);
    if (maps) (void)hipFree(maps);
    if (status) (void)hipFree(status);
    if (hstatus) (void)hipHostFree(hstatus);
    for (auto& e : hready)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// Per-domain block counts of the list kernels under a schedule: a budget of
// `list_blocks` blocks (and emit_blocks for resolve/emit) shared in proportion
// to the domains' cell counts, at least `min_blocks` each, capped by what a
// domain can use (one block per 256 cells of its grid, kMaxBlocks for the
// scanned kernels).
void batch_geometry(const amca_batch* b, std::vector<int>& nl, std::vector<int>& ne) {
  const size_t n = b->doms.size();
  double total = 0.0;
  for (size_t i = 0; i < n; i++)
    if (b->active[i]) total += (double)b->doms[i]->P.N;
  nl.assign(n, 0);
  ne.assign(n, 0);
  for (size_t i = 0; i < n; i++) {
    if (!b->active[i]) continue;
    const double share = (double)b->doms[i]->P.N / total;
    const long long cap = std::max(1LL, (b->doms[i]->P.N + amca::kBlock - 1) / amca::kBlock);
    long long l = (long long)std::llround(share * b->sched.list_blocks);
    long long e = (long long)std::llround(share * b->sched.emit_blocks);
    l = std::max<long long>(b->sched.min_blocks, std::min(l, cap));
    // resolve / emit blocks keep per-block stats in kMaxBlocks slots
    e = std::min(std::max<long long>(b->sched.min_blocks, std::min(e, cap)), (long long)amca::kMaxBlocks);
    nl[i] = (int)std::max(1LL, l);
    ne[i] = (int)std::max(1LL, e);
  }
}

// Block order within one launch: domain-major in the given domain order
// (`interleave` = 0), or round-robin over domains so that every domain's
// blocks are spread over the whole launch (`interleave` = 1).
void build_map(const std::vector<int>& order, const std::vector<int>& cnt, int interleave,
               std::vector<int4>& out) {
  out.clear();
  if (!interleave) {
    for (int d : order)
      for (int k = 0; k < cnt[d]; k++) out.push_back(make_int4(d, k, cnt[d], 0));
    return;
  }
  int left = 1;
  for (int k = 0; left; k++) {
    left = 0;
    for (int d : order)
      if (k < cnt[d]) {
        out.push_back(make_int4(d, k, cnt[d], 0));
        left = 1;
      }
  }
}

amca_status batch_apply_schedule(amca_batch* b) {
  const int n = (int)b->doms.size();
  std::vector<int> nl, ne, nn(n), one(n, 1), ni(n), nc(n);
  batch_geometry(b, nl, ne);
  for (int i = 0; i < n; i++) {
    // a finished domain gets no blocks (its kernels would return at once)
    nn[i] = b->active[i] ? b->doms[i]->nuc_blocks : 0;
    ni[i] = b->active[i] ? b->doms[i]->ids_blocks : 0;
    nc[i] = b->active[i] ? b->doms[i]->cmp_blocks : 0;
    one[i] = b->active[i] ? 1 : 0;
  }
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  // larger domains first: their blocks start earliest
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int c) { return b->doms[a]->P.N > b->doms[c]->P.N; });
  std::vector<int4> mn, ml, me, ms, mi, mc;
  build_map(order, nn, b->sched.interleave, mn);
  build_map(order, ni, b->sched.interleave, mi);
  for (auto& m : mi) m.w = ne[m.x];  // the id scan sums the domain's k_resolve blocks
  build_map(order, nc, b->sched.interleave, mc);
  build_map(order, nl, b->sched.interleave, ml);
  build_map(order, ne, b->sched.interleave, me);
  build_map(order, one, 0, ms);
  const size_t total = mn.size() + ml.size() + me.size() + ms.size() + mi.size() + mc.size();
  if (b->maps) HIP_TRY(hipFree(b->maps));
  b->maps = nullptr;
  HIP_TRY(hipMalloc(&b->maps, sizeof(int4) * total));
  std::vector<int4> all;
  all.insert(all.end(), mn.begin(), mn.end());
  all.insert(all.end(), ml.begin(), ml.end());
  all.insert(all.end(), me.begin(), me.end());
  all.insert(all.end(), ms.begin(), ms.end());
  all.insert(all.end(), mi.begin(), mi.end());
  all.insert(all.end(), mc.begin(), mc.end());
  HIP_TRY(hipMemcpy(b->maps, all.data(), sizeof(int4) * total, hipMemcpyHostToDevice));
  b->L.map_nuc = b->maps;
  b->L.map_list = b->maps + mn.size();
  b->L.map_emit = b->L.map_list + ml.size();
  b->L.map_one = b->L.map_emit + me.size();
  b->L.map_ids = b->L.map_one + ms.size();
  b->L.map_cmp = b->L.map_ids + mi.size();
  b->L.n_nuc = (int)mn.size();
  b->L.n_list = (int)ml.size();
  b->L.n_emit = (int)me.size();
  b->L.n_one = (int)ms.size();
  b->L.n_ids = (int)mi.size();
  b->L.n_cmp = (int)mc.size();
  b->drop_graph();
  return AMCA_OK;
}

amca_status batch_capture(amca_batch* b, int steps) {
  b->drop_graph();
  HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipSuccess;
  for (int q = 0; q < steps && e == hipSuccess; q++) e = amca::launch_batch_step(b->L, b->stream);
  if (e == hipSuccess) e = amca::launch_batch_status(b->L, (int)b->doms.size(), b->status, b->stream);
  hipGraph_t g = nullptr;
  const hipError_t e2 = hipStreamEndCapture(b->stream, &g);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return fail(AMCA_ERR_HIP, std::string("graph capture: ") +
                                  hipGetErrorString(e != hipSuccess ? e : e2));
  }
  b->graph = g;
  HIP_TRY(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  b->graph_steps = steps;
  return AMCA_OK;
}

// Replays the graph until every domain has finished or `max_graphs` replays
// ran; returns the number of finished domains and the elapsed time.  Each
// replay ends with a status kernel whose output is copied to a pinned slot;
// replay r + 1 is queued before the host reads replay r's slot, so the GPU
// does not wait for the host's check.  A replay queued after every domain
// finished runs steps that all skip.  With `remap`, once max(4, n/8) more
// domains have finished the maps are rebuilt over the domains still running,
// so that the budgets of the list kernels go to them alone (a finished
// domain's kernels return at once, so leaving it out changes no result).
amca_status batch_replay(amca_batch* b, long long max_graphs, int* finished, double* ms,
                         long long* replays, bool remap) {
  if (b->use_graph && (!b->exec || b->graph_steps != b->sched.steps_per_graph))
    AMCA_TRY(batch_capture(b, b->sched.steps_per_graph));
  const int n = (int)b->doms.size();
  struct Events {  // destroyed on every return
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Events() {
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
    }
  } ev;
  HIP_TRY(hipEventCreate(&ev.e0));
  HIP_TRY(hipEventCreate(&ev.e1));
  auto queue = [&](long long r) -> hipError_t {
    hipError_t e = hipSuccess;
    if (b->use_graph) {
      e = hipGraphLaunch(b->exec, b->stream);
    } else {
      for (int q = 0; q < b->sched.steps_per_graph && e == hipSuccess; q++)
        e = amca::launch_batch_step(b->L, b->stream);
      if (e == hipSuccess) e = amca::launch_batch_status(b->L, n, b->status, b->stream);
    }
    if (e == hipSuccess)
      e = hipMemcpyAsync(b->hstatus + (r & 1) * n, b->status, sizeof(int4) * n,
                         hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipEventRecord(b->hready[r & 1], b->stream);
    return e;
  };
  HIP_TRY(hipEventRecord(ev.e0, b->stream));
  long long r = 0, queued = 0;
  int fin = 0;
  if (max_graphs != 0) {
    HIP_TRY(queue(0));
    queued = 1;
  }
  while (r < queued) {
    if (max_graphs < 0 || queued < max_graphs) {
      HIP_TRY(queue(queued));
      queued++;
    }
    HIP_TRY(hipEventSynchronize(b->hready[r & 1]));
    const int4* st = b->hstatus + (r & 1) * n;
    r++;
    fin = 0;
    for (int i = 0; i < n; i++) {
      if (st[i].y) {
        (void)hipStreamSynchronize(b->stream);
        return fail(AMCA_ERR_CAPACITY, "orientation table exhausted in domain " + std::to_string(i));
      }
      fin += st[i].x;
    }
    if (fin == n) break;
    if (remap && fin - b->mapped_finished >= std::max(4, n / 8)) {
      // rebuild the maps over the running domains: wait for the queued replay,
      // take the finished set from its status, and capture a new graph
      HIP_TRY(hipStreamSynchronize(b->stream));
      const int4* last = b->hstatus + ((queued - 1) & 1) * n;
      fin = 0;
      for (int i = 0; i < n; i++) {
        if (last[i].y)
          return fail(AMCA_ERR_CAPACITY, "orientation table exhausted in domain " + std::to_string(i));
        fin += last[i].x;
        b->active[i] = last[i].x ? 0 : 1;
      }
      r = queued;
      if (fin == n) break;
      b->mapped_finished = fin;
      AMCA_TRY(batch_apply_schedule(b));
      if (b->use_graph) AMCA_TRY(batch_capture(b, b->sched.steps_per_graph));
      if (max_graphs < 0 || queued < max_graphs) {
        HIP_TRY(queue(queued));
        queued++;
      }
    }
  }
  HIP_TRY(hipEventRecord(ev.e1, b->stream));
  HIP_TRY(hipEventSynchronize(ev.e1));
  float t = 0.f;
  HIP_TRY(hipEventElapsedTime(&t, ev.e0, ev.e1));
  r = queued;
  if (finished) *finished = fin;
  if (ms) *ms = t;
  if (replays) *replays = r;
  return AMCA_OK;
}

}  // namespace

extern "C" {

amca_status amca_batch_create(amca_domain* const* doms, int32_t n, amca_batch** out) {
  if (!doms || n < 1 || !out) return fail(AMCA_ERR_INVALID, "need at least one domain");
  *out = nullptr;
  auto* b = new amca_batch();
  for (int i = 0; i < n; i++) {
    if (!doms[i]) {
      delete b;
      return fail(AMCA_ERR_INVALID, "null domain in batch");
    }
    if (doms[i]->timing) {
      delete b;
      return fail(AMCA_ERR_INVALID, "disable per-kernel timing before batching a domain");
    }
    amca::DevCounters c;
    amca_status rows = read_counters(doms[i], &c);
    if (rows == AMCA_OK) rows = check_rows(doms[i], c);
    if (rows != AMCA_OK) {
      delete b;
      return rows;
    }
    b->doms.push_back(doms[i]);
  }
  amca_status st = AMCA_OK;
  do {
    if (hipGetDevice(&b->device) != hipSuccess) { st = fail(AMCA_ERR_HIP, "hipGetDevice"); break; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, b->device) == hipSuccess && prop.multiProcessorCount > 0)
      b->cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
      st = fail(AMCA_ERR_HIP, "stream creation failed");
      break;
    }
    std::vector<amca::DevParams> P(n);
    std::vector<amca::DevBuffers> B(n);
    for (int i = 0; i < n; i++) {
      // every domain's own queued work must be done before the batch stream takes over
      if (hipStreamSynchronize(b->doms[i]->stream) != hipSuccess) { st = fail(AMCA_ERR_HIP, "sync"); break; }
      P[i] = b->doms[i]->P;
      B[i] = b->doms[i]->B;
    }
    if (st != AMCA_OK) break;
    if (hipMalloc(&b->dP, sizeof(amca::DevParams) * n) != hipSuccess ||
        hipMalloc(&b->dB, sizeof(amca::DevBuffers) * n) != hipSuccess ||
        hipMalloc(&b->status, sizeof(int4) * n) != hipSuccess ||
        hipHostMalloc(&b->hstatus, sizeof(int4) * 2 * n) != hipSuccess ||
        hipEventCreateWithFlags(&b->hready[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->hready[1], hipEventDisableTiming) != hipSuccess) {
      st = fail(AMCA_ERR_HIP, "batch allocation failed");
      break;
    }
    if (hipMemcpy(b->dP, P.data(), sizeof(amca::DevParams) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->dB, B.data(), sizeof(amca::DevBuffers) * n, hipMemcpyHostToDevice) != hipSuccess) {
      st = fail(AMCA_ERR_HIP, "batch upload failed");
      break;
    }
    b->L.P = b->dP;
    b->L.B = b->dB;
    b->sched.list_blocks = 16 * b->cus;
    b->sched.emit_blocks = 16 * b->cus;
    b->sched.min_blocks = 1;
    b->sched.interleave = 0;
    b->sched.steps_per_graph = 16;
    b->active.assign(n, 1);
    int g = 1;
    if (amca::env_int("AMCA_BATCH_GRAPHS", 0, 1, &g) < 0) {
      st = fail(AMCA_ERR_INVALID, "AMCA_BATCH_GRAPHS must be 0 or 1");
      break;
    }
    b->use_graph = g != 0;
    st = batch_apply_schedule(b);
  } while (0);
  if (st != AMCA_OK) {
    delete b;
    return st;
  }
  *out = b;
  return AMCA_OK;
}

amca_status amca_batch_destroy(amca_batch* b) {
  delete b;
  return AMCA_OK;
}

amca_status amca_batch_set_schedule(amca_batch* b, const amca_batch_schedule* s) {
  if (!b || !s || s->list_blocks < 1 || s->emit_blocks < 1 || s->min_blocks < 1 ||
      s->min_blocks > amca::kMaxBlocks || s->steps_per_graph < 1 || s->steps_per_graph > 4096)
    return fail(AMCA_ERR_INVALID, "bad schedule (budgets >= 1, min_blocks in [1, 4096], "
                                  "steps_per_graph in [1, 4096])");
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->sched = *s;
  return batch_apply_schedule(b);
}

amca_status amca_batch_get_schedule(amca_batch* b, amca_batch_schedule* s) {
  if (!b || !s) return fail(AMCA_ERR_INVALID, "null argument");
  *s = b->sched;
  return AMCA_OK;
}

amca_status amca_batch_run(amca_batch* b, int64_t max_graphs, int32_t* n_finished, double* ms,
                           int64_t* replays) {
  if (!b) return fail(AMCA_ERR_INVALID, "null batch");
  int fin = 0;
  double t = 0.0;
  long long r = 0;
  AMCA_TRY(batch_replay(b, max_graphs, &fin, &t, &r, true));
  if (n_finished) *n_finished = fin;
  if (ms) *ms = t;
  if (replays) *replays = r;
  return AMCA_OK;
}

// The autotuner runs the real batch: each candidate schedule takes
// `trial_graphs` replays of a short graph (kTrialSteps steps, or the
// schedule's own length if shorter) in turn, for `rounds` rounds, and the
// schedule with the least mean time per step wins.  The batch's work per step
// drifts while the trials run, so odd rounds visit the candidates in reverse
// order, which cancels a steady drift over each pair of rounds.  Stage 1 tries
// one block budget for all list kernels (4, 16, 64 blocks per CU) in both
// block orders; stage 2 sizes the claim grid and the resolve / emit grid
// apart, a factor 4 up and down from the stage-1 choice.  Schedules do not
// change results, so the trial steps are ordinary steps of the run.
amca_status amca_batch_autotune(amca_batch* b, int32_t trial_graphs, int32_t rounds,
                                amca_batch_schedule* chosen) {
  if (!b || trial_graphs < 1 || rounds < 1) return fail(AMCA_ERR_INVALID, "bad arguments");
  constexpr int kTrialSteps = 4;
  const int run_steps = b->sched.steps_per_graph;
  const int trial_steps = std::min(run_steps, kTrialSteps);
  // the schedule as it was, restored if a trial fails
  const amca_batch_schedule before = b->sched;
  struct Restore {
    amca_batch* b;
    const amca_batch_schedule* s;
    bool armed = true;
    ~Restore() {
      if (!armed) return;
      (void)hipStreamSynchronize(b->stream);
      b->sched = *s;
      (void)batch_apply_schedule(b);
    }
  } restore{b, &before};
  const int emit_cap = (int)amca::kMaxBlocks * 8;
  auto make = [&](int list_budget, int emit_budget, int inter) {
    amca_batch_schedule s = b->sched;
    s.list_blocks = list_budget * b->cus;
    s.emit_blocks = std::min(emit_budget * b->cus, emit_cap);
    s.min_blocks = 1;
    s.interleave = inter;
    return s;
  };
  bool all_done = false;
  // mean ms per step of each candidate over `rounds` passes; the best index
  auto pick = [&](const std::vector<amca_batch_schedule>& cand, size_t* best) -> amca_status {
    std::vector<double> tot(cand.size(), 0.0);
    std::vector<int> runs(cand.size(), 0);
    for (int r = 0; r < rounds && !all_done; r++)
      for (size_t u = 0; u < cand.size() && !all_done; u++) {
        const size_t c = (r & 1) ? cand.size() - 1 - u : u;
        b->sched.list_blocks = cand[c].list_blocks;
        b->sched.emit_blocks = cand[c].emit_blocks;
        b->sched.min_blocks = cand[c].min_blocks;
        b->sched.interleave = cand[c].interleave;
        b->sched.steps_per_graph = trial_steps;
        AMCA_TRY(batch_apply_schedule(b));
        int fin = 0;
        double ms = 0.0;
        long long rep = 0;
        AMCA_TRY(batch_replay(b, trial_graphs, &fin, &ms, &rep, false));
        tot[c] += ms / (double)(std::max(1LL, rep) * trial_steps);
        runs[c]++;
        if (fin == (int)b->doms.size()) all_done = true;
      }
    size_t k = 0;
    for (size_t c = 1; c < cand.size(); c++)
      if (runs[c] > 0 && (runs[k] == 0 || tot[c] / runs[c] < tot[k] / runs[k])) k = c;
    *best = k;
    return AMCA_OK;
  };
  // stage 1: one budget for every list kernel, both block orders
  std::vector<amca_batch_schedule> cand;
  for (int budget : {4, 16, 64})
    for (int inter : {0, 1}) cand.push_back(make(budget, budget, inter));
  size_t best = 0;
  AMCA_TRY(pick(cand, &best));
  amca_batch_schedule win = cand[best];
  // stage 2: the claim kernel's grid and the resolve / emit grid apart, a
  // factor 4 either way from the stage-1 choice
  if (!all_done) {
    const int lb = win.list_blocks / b->cus, eb = std::max(1, win.emit_blocks / b->cus);
    std::vector<amca_batch_schedule> c2 = {win};
    for (int f : {4, -4}) {
      const int l2 = f > 0 ? lb * f : std::max(1, lb / -f);
      const int e2 = f > 0 ? eb * f : std::max(1, eb / -f);
      if (l2 != lb) c2.push_back(make(l2, eb, win.interleave));
      if (e2 != eb) c2.push_back(make(lb, e2, win.interleave));
    }
    AMCA_TRY(pick(c2, &best));
    win = c2[best];
  }
  b->sched.list_blocks = win.list_blocks;
  b->sched.emit_blocks = win.emit_blocks;
  b->sched.min_blocks = win.min_blocks;
  b->sched.interleave = win.interleave;
  b->sched.steps_per_graph = run_steps;
  restore.armed = false;
  AMCA_TRY(batch_apply_schedule(b));
  if (chosen) *chosen = b->sched;
  return AMCA_OK;
}

}  // extern "C"
