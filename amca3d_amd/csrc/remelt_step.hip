// Kernels of the coupled CA-FE step (realm type `cellular_automata_remelting`
// driven by CafeSolver::time_integration_weak_coupling, CafeSolver.cpp:178-219).
//
// One step, in the reference's order (each grid kernel over the step's active
// box, see the active-box notes below):
//   k_rm_field          the step's nodal field, interpolated in time between
//                       the frames the last step selected
//                       (FiniteElementManager::pseudo_single_integration_step,
//                       LoadDataFromFile.cpp:207-313), and the hot nodes that
//                       seed the active box
//   k_rm_sur_nodes/elems  (surrogate only) element fallback flags
//   k_rm_phase          temperature of every cell from the FE nodal field (hex8
//                       interpolation, MapVoxelManager::evaluate_temperature),
//                       cooling flag, melting
//                       (checkup_current_temperature_to_update_phase_state,
//                       CellularAutomataManager_3DRemelting.cpp:338-404)
//   k_rm_liquid         flags the solid neighbours of undercooled liquid cells
//                       (checkup_liquid_cell_neigh_to_growth, :485-542)
//   k_rm_nucleate_mint  nucleation sites, remelting variant (:936-999), and in
//                       the same launch the length cutoff and the minimum
//                       flagged temperature (:656-727)
//   k_rm_claim          dt = min(factor*h/v(minT), dT_FE); growth of every
//                       flagged grain and claims on enclosed liquid neighbours
//                       (single_integration_step, :544-652)
//   k_rm_resolve        captures: the claim with the smallest x-major sweep key
//   k_rm_rowscan        grain ids for cells that never had a grain, in sweep
//                       order; then the time advance and FE frame selection
//   k_rm_emit           child envelopes (capture_cell, :807-847)
// The reference sweeps cells with i outermost and k innermost, and a capture in
// the sweep is final, so a liquid cell goes to the first flagged cell in that
// order whose envelope holds it: key (i*nyG + j)*nzG + k.  Per-row counts of
// fresh captures, scanned in x-major row order, give the new grain ids.
#include <hip/hip_runtime.h>
#include <climits>

#include "device_util.h"
#include "envelope.h"
#include "libm_pow.h"
#include "remelt_step.h"

namespace amca {

namespace {

constexpr int kRmBlock = 256;

__device__ __forceinline__ void rm_ijk(const RmParams& P, int c, int& i, int& j, int& k) {
  k = c / P.nxyG;
  int rem = c - k * P.nxyG;
  j = rem / P.nxG;
  i = rem - j * P.nxG;
}

__device__ __forceinline__ bool rm_ghost_ijk(const RmParams& P, int i, int j, int k) {
  return i == 0 || j == 0 || k == 0 || i == P.nxG - 1 || j == P.nyG - 1 || k == P.nzG - 1;
}

// ProblemPhysics_RappazGandin1993::compute_growth_velocity_with_temperature_input
// (ProblemPhysics.cpp:142-158): v = a1 pow(dT, b1) + a2 pow(dT, b2) + a3 pow(dT, b3)
// for dT = Tm - T > 0, summed left to right, clamped at 0.  pow is glibc's
// (libm_pow.h, with exact-product shortcuts for the exponents 1, 2, 3), so the
// velocity, and with it every growth increment and the step's dt, is the
// reference's to the bit for any exponents.
// a * pow(x, b) for x > 0.  For 0 <= b < 15 and x < 2^64, pow(x, b) is finite
// and not negative, so a zero coefficient gives a itself (a zero of a's sign)
// and the pow need not be evaluated; the reference decks set a1 or a3 to 0.
__device__ __forceinline__ double rm_term(double a, double x, double b) {
  if (a == 0.0 && x < 0x1p64 && b >= 0.0 && b < 15.0) return a;
  return a * amca_libm::glibc_pow_fast(x, b);
}

__device__ __forceinline__ double rm_velocity(const RmParams& P, double Tc) {
  const double dTu = P.Tm - Tc;
  double v = 0.0;
  if (dTu > 0) {
    v = rm_term(P.a1, dTu, P.b1) + rm_term(P.a2, dTu, P.b2) + rm_term(P.a3, dTu, P.b3);
    if (v < 0.0) v = 0.0;
  }
  return v;
}

// order-preserving map of a double to an unsigned key (for atomicMin)
__device__ __forceinline__ unsigned long long dkey(double x) {
  unsigned long long b = (unsigned long long)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double dkey_inv(unsigned long long k) {
  unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double((long long)b);
}

// thetaArray_ at node nd for field f (FiniteElementManager::
// pseudo_single_integration_step, LoadDataFromFile.cpp:207-313): the first
// frame as loaded, else void where the newer frame is void, the newer frame
// where the older is void, otherwise linear in time.
__device__ __forceinline__ double rm_theta(const RmParams& P, const RmBuffers& B, const RmField& f,
                                           int nd) {
  if (f.init) return B.frames[nd];
  const double* th0 = B.frames + (size_t)f.frame * P.nodes;
  const double a = th0[nd], b = th0[nd - P.nodes];
  if (a <= P.void_temp) return P.void_temp;
  if (b <= P.void_temp) return a;
  return f.ratio * (a - b) + b;
}

// MapVoxelManager::evaluate_temperature (MapVoxelManager.cpp:746-799) on a
// structured mesh: element e has corner node ((ek*fny + ej)*fnx + ei).
// `field` (optional): field f already evaluated at every node.
__device__ __forceinline__ bool rm_eval_T(const RmParams& P, const RmBuffers& B, const RmField& f,
                                          int c, double& out, const double* field = nullptr) {
  const double nm = P.void_temp + 0.00001;
  const int e = B.tag[c];
  if (e < 0) return false;
  const int ei = e % P.ncx;
  const int ej = (e / P.ncx) % P.ncy;
  const int ek = e / (P.ncx * P.ncy);
  const int nxy = P.fnx * P.fny;
  int nid[8];
  nid[0] = nxy * ek + P.fnx * ej + ei;
  nid[1] = nid[0] + 1;
  nid[2] = nid[1] + P.fnx;
  nid[3] = nid[2] - 1;
  nid[4] = nid[0] + nxy;
  nid[5] = nid[1] + nxy;
  nid[6] = nid[2] + nxy;
  nid[7] = nid[3] + nxy;
  double th[8];
#pragma unroll
  for (int m = 0; m < 8; m++) th[m] = field ? field[nid[m]] : rm_theta(P, B, f, nid[m]);
#pragma unroll
  for (int m = 0; m < 8; m++)
    if (th[m] < nm) return false;
  const double xi = B.par[3 * (size_t)c], eta = B.par[3 * (size_t)c + 1],
               zeta = B.par[3 * (size_t)c + 2];
  double sf[8];
  sf[0] = 0.125 * (1.0 - eta) * (1.0 - xi) * (1.0 - zeta);
  sf[1] = 0.125 * (1.0 - eta) * (1.0 + xi) * (1.0 - zeta);
  sf[2] = 0.125 * (1.0 + eta) * (1.0 + xi) * (1.0 - zeta);
  sf[3] = 0.125 * (1.0 + eta) * (1.0 - xi) * (1.0 - zeta);
  sf[4] = 0.125 * (1.0 - eta) * (1.0 - xi) * (1.0 + zeta);
  sf[5] = 0.125 * (1.0 - eta) * (1.0 + xi) * (1.0 + zeta);
  sf[6] = 0.125 * (1.0 + eta) * (1.0 + xi) * (1.0 + zeta);
  sf[7] = 0.125 * (1.0 + eta) * (1.0 - xi) * (1.0 + zeta);
  double s = 0.0;
#pragma unroll
  for (int m = 0; m < 8; m++) s += sf[m] * th[m];
  out = s;
  return true;
}

}  // namespace

// surrogate temperature of cell (i, j, k) at FE time t from the moving-frame
// table (linear in xi)
__device__ __forceinline__ double rm_sur_T(const RmParams& P, const RmBuffers& B, int i, int j,
                                           int k, double t) {
  const double xi = P.x0 + P.h * (i - 0.5) - P.sur_x_start - P.sur_v * t;
  double u = (xi - P.sur_xi0) / P.sur_dxi;
  int i0 = (int)floor(u);
  i0 = i0 < 0 ? 0 : (i0 > P.sur_nxi - 2 ? P.sur_nxi - 2 : i0);
  double f = u - i0;
  f = f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
  const float* row = B.sur_cell + ((size_t)k * P.nyG + j) * P.sur_nxi;
  return (1.0 - f) * (double)row[i0] + f * (double)row[i0 + 1];
}

// surrogate temperature of FE node (ni, nj, nk) at FE time t (node table)
__device__ __forceinline__ double rm_sur_node_T(const RmParams& P, const RmBuffers& B, int ni,
                                                int nj, int nk, double t) {
  const double xi = P.fx0 + P.fdx * ni - P.sur_x_start - P.sur_v * t;
  double u = (xi - P.sur_xi0) / P.sur_dxi;
  int i0 = (int)floor(u);
  i0 = i0 < 0 ? 0 : (i0 > P.sur_nxi_nodes - 2 ? P.sur_nxi_nodes - 2 : i0);
  double f = u - i0;
  f = f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
  const float* row = B.sur_node + ((size_t)nk * P.fny + nj) * P.sur_nxi_nodes;
  return (1.0 - f) * (double)row[i0] + f * (double)row[i0 + 1];
}

// Surrogate fallback rule of one FE node at one step: the surrogate differs
// from the FE nodal field by more than sur_resid, or the node is void.
__device__ __forceinline__ bool rm_node_bad(const RmParams& P, const RmBuffers& B, int nd,
                                            const RmField& f, double t) {
  const int ni = nd % P.fnx, nj = (nd / P.fnx) % P.fny, nk = nd / (P.fnx * P.fny);
  const double Ts = rm_sur_node_T(P, B, ni, nj, nk, t);
  const double th = rm_theta(P, B, f, nd);
  return fabs(Ts - th) > P.sur_resid || th < P.void_temp + 0.00001;
}

// An element falls back to the FE path when any of its nodes is bad.
__device__ __forceinline__ bool rm_elem_bad(const RmParams& P, const RmBuffers& B, int e,
                                            const RmField& f, double t) {
  const int ei = e % P.ncx, ej = (e / P.ncx) % P.ncy, ek = e / (P.ncx * P.ncy);
  const int nxy = P.fnx * P.fny;
  const int n0 = nxy * ek + P.fnx * ej + ei;
  const int nid[8] = {n0, n0 + 1, n0 + P.fnx, n0 + P.fnx + 1,
                      n0 + nxy, n0 + nxy + 1, n0 + nxy + P.fnx, n0 + nxy + P.fnx + 1};
  bool bad = false;
#pragma unroll
  for (int m = 0; m < 8; m++) bad |= rm_node_bad(P, B, nid[m], f, t);
  return bad;
}

// Temperature of cell c at one step: the FE path on nodal field f, or with
// the surrogate (sur) the table value at FE time tfe unless the cell's element
// falls back (efb, the step's flags; NULL: the rule evaluated here) or the
// value is within sur_band of Tm.  Returns whether the cell has material;
// n_fb counts cells of an element that took the FE path.
__device__ __forceinline__ bool rm_cell_T(const RmParams& P, const RmBuffers& B, int c,
                                          const RmField& f, const unsigned char* efb,
                                          double tfe, int sur, double& Tn, int& n_fb,
                                          const double* field = nullptr) {
  Tn = -1.0;
  if (!sur) return rm_eval_T(P, B, f, c, Tn, field);
  const int e = B.tag[c];
  bool fe = e < 0 || (efb ? efb[e] != 0 : rm_elem_bad(P, B, e, f, tfe));
  if (!fe) {
    int si, sj, sk;
    rm_ijk(P, c, si, sj, sk);
    Tn = rm_sur_T(P, B, si, sj, sk, tfe);
    fe = fabs(Tn - P.Tm) < P.sur_band;
  }
  if (!fe) return true;
  const bool mat = rm_eval_T(P, B, f, c, Tn, field);
  if (!mat) Tn = -1.0;
  if (e >= 0) n_fb++;
  return mat;
}
__device__ __forceinline__ const unsigned char* rm_fb(const RmParams& P, const RmBuffers& B, int par) {
  return B.elem_fb ? B.elem_fb + (size_t)par * (P.ncx * P.ncy * (P.fnz - 1)) : nullptr;
}

// ---------------------------------------------------------------------------
// active box
//
// A cell outside the box of step n changes nothing in step n but its
// temperature and cooling flag, which nothing reads that step: it is solid
// and has material, its elements are below Tm - hot_margin at every node (the
// hex8 value is a convex combination of them), it has no undercooled liquid
// neighbour (so it is not flagged) and no flagged neighbour (so it is not
// captured).  The box of step n is the union of
//   - the cells of elements with a node at or above Tm - hot_margin, or whose
//     nodes change between void and material, at step n (k_rm_field), and
//     with the surrogate the table's hot region at step n;
//   - the liquid cells with material after the phase of step n-1 (k_rm_liquid);
//   - the cells captured in step n-1 (a capture of a cell without material is
//     undone by the next phase) (k_rm_emit),
// widened by two cells (flags reach one cell, captures one more).  The first
// step, and any step after the state or the temperature source changed, runs
// over the whole grid.  A cell that enters the box takes its last-step
// temperature for the cooling rate from the last step's field, which is what
// it would have stored; k_rm_complete writes T and cooling of the cells
// outside the last box when a run returns.

__device__ __forceinline__ void box_empty(RmBox& b) {
  for (int a = 0; a < 3; a++) {
    b.lo[a] = INT_MAX;
    b.hi[a] = INT_MIN;
  }
}
__device__ __forceinline__ void box_add(RmBox& b, int i, int j, int k) {
  b.lo[0] = min(b.lo[0], i);
  b.lo[1] = min(b.lo[1], j);
  b.lo[2] = min(b.lo[2], k);
  b.hi[0] = max(b.hi[0], i);
  b.hi[1] = max(b.hi[1], j);
  b.hi[2] = max(b.hi[2], k);
}
__device__ __forceinline__ bool box_has(const RmBox& b, int i, int j, int k) {
  return i >= b.lo[0] && i <= b.hi[0] && j >= b.lo[1] && j <= b.hi[1] && k >= b.lo[2] &&
         k <= b.hi[2];
}
__device__ __forceinline__ void box_full(const RmParams& P, RmBox& b) {
  b.lo[0] = b.lo[1] = b.lo[2] = 0;
  b.hi[0] = P.nxG - 1;
  b.hi[1] = P.nyG - 1;
  b.hi[2] = P.nzG - 1;
}
// extents of b clipped to [lo, hi] on every axis; all 0 when the clipped box
// is empty (an empty box holds lo = INT_MAX, hi = INT_MIN: no difference of
// the raw bounds is taken, it would overflow)
__device__ __forceinline__ void box_clip(const RmBox& b, const int* lo, const int* hi, int* o0,
                                         int* ext) {
  bool empty = false;
  for (int a = 0; a < 3; a++) {
    const int l = max(b.lo[a], lo[a]);
    const int h = min(b.hi[a], hi[a]);
    o0[a] = l;
    empty |= h < l;
    ext[a] = h < l ? 0 : h - l + 1;
  }
  if (empty)
    for (int a = 0; a < 3; a++) {
      o0[a] = lo[a];
      ext[a] = 0;
    }
}

// Merge every thread's box into *dst: a min/max reduction over the block,
// then one atomic per bound.
// Box of the whole block, valid in thread 0 (a min/max reduction over the block).
template <int kThreads>
__device__ __forceinline__ RmBox block_box_reduce(RmBox b) {
  __shared__ int s_b[kThreads / 64][6];
  int v[6] = {b.lo[0], b.lo[1], b.lo[2], b.hi[0], b.hi[1], b.hi[2]};
#pragma unroll
  for (int q = 0; q < 6; q++)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const int w = __shfl_xor(v[q], o, 64);
      v[q] = q < 3 ? min(v[q], w) : max(v[q], w);
    }
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int q = 0; q < 6; q++) s_b[threadIdx.x >> 6][q] = v[q];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int w = 1; w < kThreads / 64; w++)
      for (int q = 0; q < 6; q++) v[q] = q < 3 ? min(v[q], s_b[w][q]) : max(v[q], s_b[w][q]);
  RmBox r;
  for (int q = 0; q < 3; q++) {
    r.lo[q] = v[q];
    r.hi[q] = v[q + 3];
  }
  return r;
}

// Merge every thread's box into *dst: a block reduction, then one atomic per
// bound.  For boxes that few blocks contribute to (hot nodes, captures); a box
// that most blocks contribute to goes through per-block slots instead, since
// same-address atomics from every block serialise.
__device__ __forceinline__ void block_box_merge(RmBox b, RmBox* dst) {
  const RmBox v = block_box_reduce<kRmBlock>(b);
  if (threadIdx.x == 0 && v.lo[0] <= v.hi[0]) {
    for (int q = 0; q < 3; q++) {
      atomicMin(&dst->lo[q], v.lo[q]);
      atomicMax(&dst->hi[q], v.hi[q]);
    }
  }
}

// blocks of the grid-stride cell kernels: with the active box, 1024 blocks of
// 256 (four per CU; the box is a few hundred thousand cells); over the whole
// grid, up to 16384.  k_rm_liquid writes one box slot per block, which
// k_rm_rowscan reads, so both use this count.
__host__ __device__ __forceinline__ int rm_cell_blocks(const RmParams& P) {
  const long long cb = ((long long)P.NG + kRmBlock - 1) / kRmBlock;
  const int cap = P.window ? 1024 : 16384;
  return (int)(cb < cap ? cb : cap);
}

// the box of the running step (written by k_rm_phase)
__device__ __forceinline__ RmBox rm_step_box(const RmCounters* ctr) {
  return ctr->box[ctr->step & 1];
}

__device__ __forceinline__ bool rm_skip_now(const RmParams& P, const RmCounters* ctr) {
  return !(ctr->tcafe < P.final_time && !ctr->exhausted) || ctr->need_rows;
}

__device__ __forceinline__ int block_sum_int(int v) {
  __shared__ int s_v[kRmBlock / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) s_v[threadIdx.x >> 6] = v;
  __syncthreads();
  int t = 0;
#pragma unroll
  for (int q = 0; q < kRmBlock / 64; q++) t += s_v[q];
  return t;
}

// ---------------------------------------------------------------------------

__global__ void k_rm_init(RmParams P, RmBuffers B) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < P.NG; c += stride) {
    B.T[c] = -1.0;  // nullMatTemp_ (CellularAutomataManager_3DRemelting.cpp:29)
    B.cooling[c] = 0;
    B.flag[c] = 0;
    B.claim[c] = INT_MAX;
    B.msk[c] = 0u;
  }
}

// The velocity every cell below Ts grows with (the reference evaluates it at
// Ts + 0.01 K per cell, CellularAutomataManager_3DRemelting.cpp:572, with the
// kinetics of ProblemPhysics.cpp:142-158); a constant of the domain, evaluated
// once on the device with the same kinetics code as every other velocity.
__global__ void k_rm_vts(RmParams P, double* out) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *out = rm_velocity(P, P.Ts + 0.01);
}

// Start of a step: the sites whose cell is liquid (the most nuclei the step
// can make), the hot region of the active box into acc[step & 1][0], and with
// the surrogate the step's nodal field into B.theta for the fallback test.
// A node is hot in the step's field only if it is hot in one of the two frames
// the field interpolates (the interpolation is a convex combination, and void
// nodes are cold), so the per-frame boxes of hot nodes bound the hot region.
__global__ void __launch_bounds__(kRmBlock) k_rm_field(RmParams P, RmBuffers B) {
  RmCounters* ctr = B.ctr;
  if (rm_skip_now(P, ctr)) return;
  const long long n = ctr->step;
  const int par = (int)(n & 1);
  const RmField f = ctr->fs[n % 3];
  {
    const int ns = ctr->n_sites;
    int armed = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ns; i += gridDim.x * blockDim.x)
      armed += B.oid[B.site_cell[i]] == -1 ? 1 : 0;
    if (ns > 0) {
      armed = block_sum_int(armed);
      if (threadIdx.x == 0 && armed) atomicAdd(&ctr->armed[par], armed);
    }
  }
  if (!P.window) {
    // every cell each step: the field at every node once, read by k_rm_phase
    const int stride = gridDim.x * blockDim.x;
    for (int nd = blockIdx.x * blockDim.x + threadIdx.x; nd < P.nodes; nd += stride)
      B.theta[nd] = rm_theta(P, B, f, nd);
  }
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const bool hot_test = P.window && !ctr->force_full && n > 0;
  if (!hot_test) {
    box_full(P, ctr->box[par]);
    return;
  }
  RmBox b;
  box_empty(b);
  auto add = [&](const int* q) {
    if (q[0] <= q[3]) {
      box_add(b, q[0], q[1], q[2]);
      box_add(b, q[3], q[4], q[5]);
    }
  };
  add(B.frame_hot + 6 * (size_t)(f.init ? 0 : f.frame));
  if (!f.init) add(B.frame_hot + 6 * (size_t)(f.frame - 1));
  add(P.void_box);
  if (P.sur && P.sur_hot[0] <= P.sur_hot[1]) {
    // cell i reads table entries floor(u), floor(u) + 1 (clamped), with
    // u = (x0 + h (i - 0.5) - x_start - v t - xi0) / dxi
    const double t = ctr->time_fe;
    const double base = P.sur_x_start + P.sur_v * t + P.sur_xi0 - P.x0;
    int ilo = 0, ihi = P.nxG - 1;
    if (P.sur_hot[0] > 0) {
      const double x = (P.sur_hot[0] - 1) * P.sur_dxi + base;
      ilo = max(0, (int)floor(x / P.h + 0.5) - 2);
    }
    if (P.sur_hot[1] < P.sur_nxi - 2) {
      const double x = (P.sur_hot[1] + 1) * P.sur_dxi + base;
      ihi = min(P.nxG - 1, (int)ceil(x / P.h + 0.5) + 2);
    }
    if (ilo <= ihi) {
      box_add(b, ilo, P.sur_hot[2], P.sur_hot[4]);
      box_add(b, ihi, P.sur_hot[3], P.sur_hot[5]);
    }
  }
  RmBox& a = ctr->acc[par][0];
  if (b.lo[0] <= b.hi[0])
    for (int q = 0; q < 3; q++) {
      a.lo[q] = min(a.lo[q], b.lo[q]);
      a.hi[q] = max(a.hi[q], b.hi[q]);
    }
  // the step's box: its three inputs, widened by two cells
  RmBox R;
  box_empty(R);
  for (int q = 0; q < 3; q++) {
    const RmBox& x = ctr->acc[par][q];
    if (x.lo[0] > x.hi[0]) continue;
    box_add(R, x.lo[0], x.lo[1], x.lo[2]);
    box_add(R, x.hi[0], x.hi[1], x.hi[2]);
  }
  if (R.lo[0] <= R.hi[0]) {
    const int top[3] = {P.nxG - 1, P.nyG - 1, P.nzG - 1};
    for (int q = 0; q < 3; q++) {
      R.lo[q] = max(0, R.lo[q] - 2);
      R.hi[q] = min(top[q], R.hi[q] + 2);
    }
  }
  ctr->box[par] = R;
}

__global__ void __launch_bounds__(kRmBlock) k_rm_phase(RmParams P, RmBuffers B) {
  RmCounters* ctr = B.ctr;
  const long long n = ctr->step;
  const int par = (int)(n & 1);
  const bool short_rows = ctr->n_sites > 0 && ctr->n_orient + ctr->armed[par] > ctr->rows_avail;
  const int skip = (rm_skip_now(P, ctr) || short_rows) ? 1 : 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctr->skip = skip;
    if (short_rows && !rm_skip_now(P, ctr)) ctr->need_rows = 1;
  }
  if (skip) return;
  const RmBox R = ctr->box[par];  // written by k_rm_field
  const bool later = n > 0;
  const double dt_prev = ctr->dt;
  const RmBox Rp = ctr->box[par ^ 1];
  const RmField th = ctr->fs[n % 3];
  const RmField thp = ctr->fs[(n + 2) % 3];  // step n - 1
  const unsigned char* efb = rm_fb(P, B, par);
  const double tfe = ctr->time_fe, tfe_p = ctr->time_fe_prev;
  const int sur_p = ctr->sur_used[par ^ 1];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctr->sur_used[par] = P.sur;
    ctr->last_step = n;
    ctr->last_time_fe = tfe;
    ctr->last_time_fe_prev = tfe_p;
    ctr->last_dt = dt_prev;
  }
  const int zero[3] = {0, 0, 0}, top[3] = {P.nxG - 1, P.nyG - 1, P.nzG - 1};
  int o0[3], ext[3];
  box_clip(R, zero, top, o0, ext);
  const int exy = ext[0] * ext[1];
  const int cnt = exy * ext[2];
  int n_fb = 0;
  const int stride = gridDim.x * blockDim.x;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < cnt; q += stride) {
    const int k = o0[2] + q / exy;
    const int rem = q - (k - o0[2]) * exy;
    const int j = o0[1] + rem / ext[0];
    const int i = o0[0] + rem % ext[0];
    const int c = k * P.nxyG + j * P.nxG + i;
    double Tn;
    const bool mat = rm_cell_T(P, B, c, th, efb, tfe, P.sur, Tn, n_fb, P.window ? nullptr : B.theta);
    if (later) {
      double Tp;
      if (box_has(Rp, i, j, k)) {
        Tp = B.T[c];
      } else {
        // a cell entering the box: its last-step temperature, evaluated
        // here (fallback rule included) as that step would have
        int unused = 0;
        (void)rm_cell_T(P, B, c, thp, nullptr, tfe_p, sur_p, Tp, unused);
      }
      double rate = 410;
      if (Tp > 0) rate = (Tn - Tp) / dt_prev;
      if (Tn <= 0) rate = 0;
      B.cooling[c] = mat ? (rate < 0) : 0;
    }
    B.T[c] = Tn;
    B.flag[c] = 0;
    if (rm_ghost_ijk(P, i, j, k)) continue;
    if (mat && Tn < P.Tm) continue;
    B.oid[c] = -1;
    B.melted[c] = mat ? 1 : 0;
  }
  if (P.sur) {
    // block sum, one atomic per block into one of 256 slots (summed at the end of the step)
    __shared__ int s_fb[kRmBlock / 64];
    int v = n_fb;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) s_fb[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      int t = 0;
#pragma unroll
      for (int q = 0; q < kRmBlock / 64; q++) t += s_fb[q];
      if (t) atomicAdd(&B.fb_slots[blockIdx.x & 255], t);
    }
  }
}

// T and cooling of the last step for the cells outside its box, computed as
// that step's phase would have (k_rm_phase).
__global__ void __launch_bounds__(kRmBlock) k_rm_complete(RmParams P, RmBuffers B) {
  const RmCounters* ctr = B.ctr;
  {
    // thetaArray_ of the coming step
    const RmField f = ctr->fs[ctr->step % 3];
    const int s = gridDim.x * blockDim.x;
    for (int nd = blockIdx.x * blockDim.x + threadIdx.x; nd < P.nodes; nd += s)
      B.theta[nd] = rm_theta(P, B, f, nd);
  }
  const long long n = ctr->last_step;
  if (n < 0) return;
  const int par = (int)(n & 1);
  const RmBox R = ctr->box[par];
  const RmField th = ctr->fs[n % 3];
  const RmField thp = ctr->fs[(n + 2) % 3];
  // the fallback flags are kept for the box only: the rule is evaluated here
  const double tfe = ctr->last_time_fe, tfe_p = ctr->last_time_fe_prev, dt_prev = ctr->last_dt;
  const int sur = ctr->sur_used[par], sur_p = ctr->sur_used[par ^ 1];
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long cc = (long long)blockIdx.x * blockDim.x + threadIdx.x; cc < P.NG; cc += stride) {
    const int c = (int)cc;
    int i, j, k;
    rm_ijk(P, c, i, j, k);
    if (box_has(R, i, j, k)) continue;
    int unused = 0;
    double Tn;
    const bool mat = rm_cell_T(P, B, c, th, nullptr, tfe, sur, Tn, unused);
    if (n > 0) {
      double Tp;
      (void)rm_cell_T(P, B, c, thp, nullptr, tfe_p, sur_p, Tp, unused);
      double rate = 410;
      if (Tp > 0) rate = (Tn - Tp) / dt_prev;
      if (Tn <= 0) rate = 0;
      B.cooling[c] = mat ? (rate < 0) : 0;
    }
    B.T[c] = Tn;
  }
}

// Hot region of the surrogate's cell table: entries at or above thr.
__global__ void __launch_bounds__(kRmBlock) k_rm_sur_hot(RmParams P, RmBuffers B, double thr,
                                                          RmBox* out) {
  const long long total = (long long)P.nzG * P.nyG * P.sur_nxi;
  RmBox b;
  box_empty(b);
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += stride) {
    if ((double)B.sur_cell[q] < thr) continue;
    const int x = (int)(q % P.sur_nxi);
    const int j = (int)((q / P.sur_nxi) % P.nyG);
    const int k = (int)(q / ((long long)P.sur_nxi * P.nyG));
    box_add(b, x, j, k);
  }
  block_box_merge(b, out);
}

// Surrogate fallback flags of the step, for the elements under the step's box
// (on the uniform node lattice the surrogate requires, with one element of
// margin for cells on element faces): a node is bad when the surrogate
// differs from the FE nodal field by more than sur_resid or the node is void;
// an element falls back to the FE path when any of its nodes is bad.
__device__ __forceinline__ void rm_box_nodes(const RmParams& P, const RmBox& R, int* lo, int* hi) {
  const double c0[3] = {P.x0, P.y0, P.z0}, f0[3] = {P.fx0, P.fy0, P.fz0};
  const int top[3] = {P.fnx - 1, P.fny - 1, P.fnz - 1};
  for (int a = 0; a < 3; a++) {
    if (R.lo[0] > R.hi[0]) {
      lo[a] = 1;
      hi[a] = 0;
      continue;
    }
    const double xl = c0[a] + P.h * (R.lo[a] - 0.5), xh = c0[a] + P.h * (R.hi[a] - 0.5);
    lo[a] = max(0, (int)floor((xl - f0[a]) / P.fdx) - 1);
    hi[a] = min(top[a], (int)floor((xh - f0[a]) / P.fdx) + 2);
  }
}

__global__ void __launch_bounds__(kRmBlock) k_rm_sur_nodes(RmParams P, RmBuffers B) {
  const RmCounters* ctr = B.ctr;
  if (rm_skip_now(P, ctr)) return;
  const long long n = ctr->step;
  const RmField f = ctr->fs[n % 3];
  const double t = ctr->time_fe;
  int lo[3], hi[3];
  rm_box_nodes(P, ctr->box[n & 1], lo, hi);
  const int ex = hi[0] - lo[0] + 1, ey = hi[1] - lo[1] + 1, ez = hi[2] - lo[2] + 1;
  const int total = (ex > 0 && ey > 0 && ez > 0) ? ex * ey * ez : 0;
  const int stride = gridDim.x * blockDim.x;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < total; q += stride) {
    const int ni = lo[0] + q % ex, nj = lo[1] + (q / ex) % ey, nk = lo[2] + q / (ex * ey);
    const int nd = (nk * P.fny + nj) * P.fnx + ni;
    B.node_bad[nd] = rm_node_bad(P, B, nd, f, t) ? 1 : 0;
  }
}

__global__ void __launch_bounds__(kRmBlock) k_rm_sur_elems(RmParams P, RmBuffers B) {
  const RmCounters* ctr = B.ctr;
  if (rm_skip_now(P, ctr)) return;
  const long long n = ctr->step;
  const int ne = P.ncx * P.ncy * (P.fnz - 1);
  const int nxy = P.fnx * P.fny;
  unsigned char* fb = B.elem_fb + (size_t)(n & 1) * ne;
  int lo[3], hi[3];
  rm_box_nodes(P, ctr->box[n & 1], lo, hi);
  // elements whose 8 nodes all lie in the node box
  const int ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
  const int total = (ex > 0 && ey > 0 && ez > 0) ? ex * ey * ez : 0;
  const int stride = gridDim.x * blockDim.x;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < total; q += stride) {
    const int ei = lo[0] + q % ex, ej = lo[1] + (q / ex) % ey, ek = lo[2] + q / (ex * ey);
    const int e = (ek * P.ncy + ej) * P.ncx + ei;
    const int n0 = nxy * ek + P.fnx * ej + ei;
    const unsigned char* b = B.node_bad;
    const int bad = b[n0] | b[n0 + 1] | b[n0 + P.fnx] | b[n0 + P.fnx + 1] | b[n0 + nxy] |
                    b[n0 + nxy + 1] | b[n0 + nxy + P.fnx] | b[n0 + nxy + P.fnx + 1];
    fb[e] = bad ? 1 : 0;
  }
}

// Error statistics of the surrogate against the FE path at the current FE
// time over cells with material: [0] sum sq (all), [1] count, [2] sum sq of
// the cells the surrogate would serve, [3] their count, [4] their max |err|
// (as bits of a non-negative double), [5] fallback cells.
__global__ void __launch_bounds__(kRmBlock) k_rm_sur_error(RmParams P, RmBuffers B) {
  const double t = B.ctr->time_fe;
  const RmField fcur = B.ctr->fs[B.ctr->step % 3];
  double s_all = 0.0, n_all = 0.0, s_use = 0.0, n_use = 0.0, mx = 0.0, n_fb = 0.0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long cc = (long long)blockIdx.x * blockDim.x + threadIdx.x; cc < P.NG; cc += stride) {
    const int c = (int)cc;
    double Tfe = -1.0;
    if (!rm_eval_T(P, B, fcur, c, Tfe)) continue;
    int i, j, k;
    rm_ijk(P, c, i, j, k);
    const double Ts = rm_sur_T(P, B, i, j, k, t);
    const double d = Ts - Tfe;
    s_all += d * d;
    n_all += 1.0;
    const bool fb = rm_elem_bad(P, B, B.tag[c], fcur, t) || fabs(Ts - P.Tm) < P.sur_band;
    if (fb) {
      n_fb += 1.0;
    } else {
      s_use += d * d;
      n_use += 1.0;
      mx = fmax(mx, fabs(d));
    }
  }
  atomicAdd(&B.sur_stats[0], s_all);
  atomicAdd(&B.sur_stats[1], n_all);
  atomicAdd(&B.sur_stats[2], s_use);
  atomicAdd(&B.sur_stats[3], n_use);
  atomicMax((unsigned long long*)&B.sur_stats[4], (unsigned long long)__double_as_longlong(mx));
  atomicAdd(&B.sur_stats[5], n_fb);
}

__global__ void __launch_bounds__(kRmBlock) k_rm_liquid(RmParams P, RmBuffers B) {
  RmCounters* ctr = B.ctr;
  if (ctr->skip) return;
  const double tc = ctr->tcafe, maxdt = ctr->max_dt;
  const int lo[3] = {1, 1, 1}, hi[3] = {P.nx, P.ny, P.nz};
  int o0[3], ext[3];
  box_clip(rm_step_box(ctr), lo, hi, o0, ext);
  const int exy = ext[0] * ext[1];
  const int cnt = exy * ext[2];
  RmBox liq;
  box_empty(liq);
  const int stride = gridDim.x * blockDim.x;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < cnt; q += stride) {
    const int k = o0[2] + q / exy;
    const int rem = q - (k - o0[2]) * exy;
    const int j = o0[1] + rem / ext[0];
    const int i = o0[0] + rem % ext[0];
    const int c = k * P.nxyG + j * P.nxG + i;
    const double Tc = B.T[c];
    const bool mat = !(Tc < -1.0 + 0.01);
    const int o = B.oid[c];
    if (o == -1 && mat && tc >= maxdt) B.melted[c] = 1;
    if (!mat) {
      B.melted[c] = 0;
      continue;
    }
    if (o != -1) continue;
    box_add(liq, i, j, k);  // liquid with material: part of the next step's box
    if (Tc >= P.Tm) continue;
    // all 26 loads before any store (a byte store may alias the loads)
    int no[26];
#pragma unroll
    for (int s = 0; s < 26; s++) no[s] = B.oid[c + P.off[s]];
#pragma unroll
    for (int s = 0; s < 26; s++)
      if (no[s] > -1) B.flag[c + P.off[s]] = 1;
  }
  // per-block slot, reduced by k_rm_rowscan into the next step's box inputs
  const RmBox v = block_box_reduce<kRmBlock>(liq);
  if (threadIdx.x == 0) {
    int* slot = B.blk_box + 6 * (size_t)blockIdx.x;
    for (int q = 0; q < 3; q++) {
      slot[q] = v.lo[q];
      slot[q + 3] = v.hi[q];
    }
  }
}

// nucleate_at_nucleation_sites, remelting variant: sites stay in the list;
// a nucleus reuses the cell's grain record when it has one
__device__ __forceinline__ void rm_nucleate_body(const RmParams& P, const RmBuffers& B, int nblk) {
  RmCounters* ctr = B.ctr;
  __shared__ int s_bid;
  __shared__ unsigned long long s_sm[kRmBlock / 64];
  __shared__ unsigned long long s_excl;
  if (threadIdx.x == 0) s_bid = (int)atomicAdd(&ctr->tickets[0], 1u);
  __syncthreads();
  const int bid = s_bid;
  const int n = ctr->n_sites;
  const int ng0 = ctr->n_grains, no0 = ctr->n_orient;
  auto decide = [&](int i, int& c, bool& nuc, bool& fresh) {
    c = B.site_cell[i];
    const double Tc = B.T[c];
    nuc = false;
    fresh = false;
    if (B.cooling[c] && B.oid[c] == -1 && B.melted[c] && (Tc > P.Tmin_nuc)) {
      const double dTu = P.Tm - Tc;
      if (dTu > B.site_dt[i]) {
        nuc = true;
        fresh = B.gid[c] == -1;
      }
    }
  };
  int lo, hi;
  chunk_of(n, bid, nblk, kRmBlock, lo, hi);
  unsigned long long mine = 0;
  for (int i = lo + threadIdx.x; i < hi; i += kRmBlock) {
    int c;
    bool nuc, fresh;
    decide(i, c, nuc, fresh);
    mine += ((unsigned long long)nuc << 32) | (unsigned long long)fresh;
  }
  unsigned long long agg;
  block_scan_excl(mine, &agg, s_sm);
  if (threadIdx.x < 64) {
    unsigned long long e =
        lookback(B.lb_flag, B.lb_agg, B.lb_pref, bid, (unsigned)(ctr->step * 4 + 0), agg);
    if (threadIdx.x == 0) s_excl = e;
  }
  __syncthreads();
  unsigned long long run = s_excl;
  for (int base = lo; base < hi; base += kRmBlock) {
    const int i = base + threadIdx.x;
    int c = 0;
    bool nuc = false, fresh = false;
    if (i < hi) decide(i, c, nuc, fresh);
    unsigned long long tot;
    const unsigned long long off =
        run + block_scan_excl(((unsigned long long)nuc << 32) | (unsigned long long)fresh, &tot, s_sm);
    if (nuc) {
      const int o = no0 + (int)(off >> 32);
      if (o < P.orient_cap) {
        int ci, cj, ck;
        rm_ijk(P, c, ci, cj, ck);
        const double4 e = make_double4(centre_coord(P.x0, P.h, ci), centre_coord(P.y0, P.h, cj),
                                       centre_coord(P.z0, P.h, ck), 0.0);
        int g = B.gid[c];
        if (fresh) {
          g = ng0 + (int)(off & 0xffffffffull);
          B.gid[c] = g;
        }
        B.genv[g] = e;
        if (fresh)
          B.gco[g] = make_int2(c, o);
        else
          B.gco[g].y = o;  // a reused record keeps its cell
        B.oid[c] = o;
      }
    }
    run += tot;
  }
  if (threadIdx.x == 0 && bid == nblk - 1) {
    const int nn = (int)(run >> 32);
    const int nf = (int)(run & 0xffffffffull);
    if (no0 + nn > P.orient_cap) ctr->err = 1;
    ctr->n_orient = no0 + nn;
    ctr->n_grains = ng0 + nf;
    ctr->step_nuc = nn;
    ctr->tickets[0] = 0u;
  }
}

// dt = min(dT_FE, min(factor*h/v(minT), dT_FE)) from the minimum flagged
// temperature (every block of k_rm_claim computes it the same way)
__device__ __forceinline__ double rm_step_dt(const RmParams& P, const RmCounters* ctr) {
  double minT = P.Tm;
  unsigned long long key = ~0ull;
#pragma unroll
  for (int q = 0; q < 64; q++) key = min(key, ctr->min_t_keys[q]);
  if (key != ~0ull) {
    const double m = dkey_inv(key);
    if (minT > m) minT = m;
  }
  const double v = rm_velocity(P, minT);
  const double maxV = (v > 0 ? v : 0);
  const double a = P.tsf * P.h / maxV;
  const double dtca = (ctr->max_dt < a) ? ctr->max_dt : a;  // std::min(a, maxDt)
  const double dtfe = ctr->max_dt;
  return ((dtfe) < (dtca) ? (dtfe) : (dtca));  // __min(dT_FE, dT_CA)
}

__device__ __forceinline__ void rm_mint_body(const RmParams& P, const RmBuffers& B, int blk,
                                             int nblk) {
  RmCounters* ctr = B.ctr;
  __shared__ unsigned long long s_min;
  __shared__ int s_cnt;
  if (threadIdx.x == 0) {
    s_min = ~0ull;
    s_cnt = 0;
  }
  __syncthreads();
  unsigned long long mk = ~0ull;
  int cnt = 0;
  const int zero[3] = {0, 0, 0}, top[3] = {P.nxG - 1, P.nyG - 1, P.nzG - 1};
  int o0[3], ext[3];
  box_clip(rm_step_box(ctr), zero, top, o0, ext);
  const int exy = ext[0] * ext[1];
  const int total = exy * ext[2];
  const int stride = nblk * blockDim.x;
  for (int q = blk * blockDim.x + threadIdx.x; q < total; q += stride) {
    const int k = o0[2] + q / exy;
    const int rem = q - (k - o0[2]) * exy;
    const int j = o0[1] + rem / ext[0];
    const int i = o0[0] + rem % ext[0];
    const int c = k * P.nxyG + j * P.nxG + i;
    const unsigned char fl = B.flag[c];
    const int g = B.gid[c];
    const double Tc = B.T[c];
    if (!fl) continue;
    cnt++;
    if (rm_ghost_ijk(P, i, j, k)) continue;
    if (g < 0) continue;
    const double L = B.genv[g].w;
    if (!(L < P.cutoff)) B.genv[g].w = P.h;  // length reset (:688-694)
    if (Tc <= P.Ts) continue;
    const unsigned long long k2 = dkey(Tc);
    if (k2 < mk) mk = k2;
  }
  if (mk != ~0ull) atomicMin(&s_min, mk);
  if (cnt) atomicAdd(&s_cnt, cnt);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (s_min != ~0ull) atomicMin(&ctr->min_t_keys[blk & 63], s_min);
    if (s_cnt) atomicAdd(&ctr->flagged_parts[blk & 63], s_cnt);
  }
}

// Nucleation at the sites (blocks 0 .. site_blocks-1) and the minimum flagged
// temperature (the other blocks) in one launch: nucleation touches only liquid
// site cells and their own grain records, the minimum only flagged (solid)
// cells and theirs.
__global__ void __launch_bounds__(kRmBlock) k_rm_nucleate_mint(RmParams P, RmBuffers B,
                                                                int site_blocks) {
  if (B.ctr->skip) return;
  if ((int)blockIdx.x < site_blocks)
    rm_nucleate_body(P, B, site_blocks);
  else
    rm_mint_body(P, B, (int)blockIdx.x - site_blocks, (int)gridDim.x - site_blocks);
}

// The sweep kernels run one thread per interior cell of the box (i fastest).
// Rows (i, j) keep the x-major sweep order: row r = (i - i0) * ny_box + (j - j0),
// and per-row counts of fresh captures give each row its first new grain id.
struct RmRows {
  int i0, j0, k0, ex, ey, ez;
};
__device__ __forceinline__ RmRows rm_rows_of(const RmParams& P, const RmBox& b) {
  const int lo[3] = {1, 1, 1}, hi[3] = {P.nx, P.ny, P.nz};
  int o0[3], ext[3];
  box_clip(b, lo, hi, o0, ext);
  return RmRows{o0[0], o0[1], o0[2], ext[0], ext[1], ext[2]};
}
__device__ __forceinline__ RmRows rm_rows(const RmParams& P, const RmCounters* ctr) {
  return rm_rows_of(P, rm_step_box(ctr));
}
__device__ __forceinline__ void rm_cell_of(const RmRows& w, int q, int& i, int& j, int& k) {
  const int exy = w.ex * w.ey;
  k = w.k0 + q / exy;
  const int rem = q - (k - w.k0) * exy;
  j = w.j0 + rem / w.ex;
  i = w.i0 + rem % w.ex;
}

constexpr unsigned kSlotBits = (1u << 26) - 1u;  // msk: captured slots; bits 26.. fresh count

__global__ void __launch_bounds__(kRmBlock) k_rm_claim(RmParams P, RmBuffers B) {
  const RmCounters* ctr = B.ctr;
  if (ctr->skip) return;
  const RmRows w = rm_rows(P, ctr);
  const int total = w.ex * w.ey * w.ez;
  // dt is the same for every thread: one thread per block evaluates it (three
  // pow calls) for the rest
  __shared__ double s_dt;
  if (threadIdx.x == 0) s_dt = rm_step_dt(P, ctr);
  __syncthreads();
  const double dt = s_dt, vts = P.vts;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // read by later kernels only
    B.ctr->dt = dt;
    B.ctr->step_cap = 0;
    B.ctr->step_fresh = 0;
  }
  const int stride = gridDim.x * blockDim.x;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < total; q += stride) {
    int i, j, k;
    rm_cell_of(w, q, i, j, k);
    if (k == w.k0) B.rowcnt[(i - w.i0) * w.ey + (j - w.j0)] = 0;
    const int c = k * P.nxyG + j * P.nxG + i;
    const unsigned char fl = B.flag[c];
    const int g = B.gid[c];
    if (!fl) continue;
    if (g < 0) continue;
    double4 e = B.genv[g];
    const double Tc = B.T[c];
    const int o = B.gco[g].y;
    // neighbour state, loaded before any store or atomic
    int no[26];
#pragma unroll
    for (int s = 0; s < 26; s++) no[s] = B.oid[c + P.off[s]];
    double tn[26];
#pragma unroll
    for (int s = 0; s < 26; s++) tn[s] = no[s] > -1 ? 0.0 : B.T[c + P.off[s]];
    const double vel = (Tc < P.Ts) ? vts : rm_velocity(P, Tc);
    e.w = e.w + dt * vel;  // Grain::grow
    B.genv[g].w = e.w;
    double r[9];
#pragma unroll
    for (int x = 0; x < 9; x++) r[x] = B.l2g[9 * (size_t)o + x];
    const int key = (i * P.nyG + j) * P.nzG + k;
    unsigned m = 0;
#pragma unroll
    for (int s = 0; s < 26; s++) {
      int di, dj, dk;
      slot_delta(s, di, dj, dk);
      const int nb = c + P.off[s];
      if (no[s] > -1) continue;
      const double Tn = tn[s];
      if (Tn > P.Tm) continue;
      if (!(Tn > -1.0)) continue;
      if (rm_ghost_ijk(P, i + di, j + dj, k + dk)) continue;
      const double gx = centre_coord(P.x0, P.h, i + di) - e.x;
      const double gy = centre_coord(P.y0, P.h, j + dj) - e.y;
      const double gz = centre_coord(P.z0, P.h, k + dk) - e.z;
      if (inside_envelope(r, e.w, gx, gy, gz)) {
        m |= 1u << s;
        atomicMin(&B.claim[nb], key);
      }
    }
    B.msk[c] = m;
  }
}

__global__ void __launch_bounds__(kRmBlock) k_rm_resolve(RmParams P, RmBuffers B) {
  RmCounters* ctr = B.ctr;
  if (ctr->skip) return;
  const RmRows w = rm_rows(P, ctr);
  const int total = w.ex * w.ey * w.ez;
  const int stride = gridDim.x * blockDim.x;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < total; q += stride) {
    int i, j, k;
    rm_cell_of(w, q, i, j, k);
    const int c = k * P.nxyG + j * P.nxG + i;
    const unsigned char fl = B.flag[c];
    const int g = B.gid[c];
    unsigned m = B.msk[c];
    if (!fl) continue;
    if (g < 0) continue;
    if (!m) continue;
    const int key = (i * P.nyG + j) * P.nzG + k;
    unsigned cap = 0;
    int fresh = 0;
    while (m) {
      const int s = __builtin_ctz(m);
      m &= m - 1;
      const int nb = c + P.off[s];
      if (B.claim[nb] == key) {
        cap |= 1u << s;
        if (B.gid[nb] < 0) fresh++;
      }
    }
    B.msk[c] = cap | ((unsigned)fresh << 26);
    if (fresh) {
      atomicAdd(&B.rowcnt[(i - w.i0) * w.ey + (j - w.j0)], fresh);
      atomicAdd(&ctr->step_fresh, fresh);
    }
    if (cap) atomicAdd(&ctr->step_cap, __builtin_popcount(cap));
  }
}

constexpr int kRmScan = 1024;

// time advance and FE frame selection (one thread)
__device__ void rm_finish(const RmParams& P, const RmBuffers& B) {
  RmCounters* ctr = B.ctr;
  const double t0 = ctr->t;
  const double dt = ctr->dt;
  ctr->t = t0 + dt;
  const long long st = ctr->step;
  ctr->step = st + 1;
  ctr->tcafe = ctr->tcafe + dt;
  if (B.log && st < ctr->log_cap) {
    RmLogRow row;
    row.time_before = t0;
    row.dt = dt;
    row.n_nucleated = ctr->step_nuc;
    row.n_captured = ctr->step_cap;
    row.n_fresh = ctr->step_fresh;
    row.n_flagged = 0;
    for (int q = 0; q < 64; q++) row.n_flagged += ctr->flagged_parts[q];
    row.n_fallback = 0;
    if (B.fb_slots)
      for (int q = 0; q < 256; q++) row.n_fallback += B.fb_slots[q];
    const RmBox& bx = ctr->box[st & 1];
    row.n_box = (bx.lo[0] <= bx.hi[0]) ? (bx.hi[0] - bx.lo[0] + 1) * (bx.hi[1] - bx.lo[1] + 1) *
                                             (bx.hi[2] - bx.lo[2] + 1)
                                       : 0;
    B.log[st] = row;
  }
  if (B.fb_slots)
    for (int q = 0; q < 256; q++) B.fb_slots[q] = 0;
  ctr->armed[st & 1] = 0;
  for (int q = 0; q < 64; q++) {
    ctr->flagged_parts[q] = 0;
    ctr->min_t_keys[q] = ~0ull;
  }
  ctr->n_fallback = 0;
  // this step's box inputs are spent; the slot serves step + 2
  for (int q = 0; q < 3; q++) {
    RmBox& a = ctr->acc[st & 1][q];
    for (int x = 0; x < 3; x++) {
      a.lo[x] = INT_MAX;
      a.hi[x] = INT_MIN;
    }
  }
  ctr->force_full = 0;
  ctr->time_fe_prev = ctr->time_fe;
  // FiniteElementManager::pseudo_single_integration_step
  double tf = ctr->time_fe;
  if (tf == 0) tf = ctr->in_t1;
  tf += dt;
  ctr->time_fe = tf;
  double in0 = ctr->in_t0, in1 = ctr->in_t1;
  int f = ctr->frame;
  while (tf >= in0) {
    if (f + 1 >= P.n_frames) {
      ctr->exhausted = 1;
      break;
    }
    f++;
    in1 = in0;
    in0 = B.frame_time[f];
  }
  ctr->frame = f;
  ctr->in_t0 = in0;
  ctr->in_t1 = in1;
  if (!ctr->exhausted) {
    ctr->ratio = (tf - in1) / (in0 - in1);
    if (ctr->ratio < 0.0) ctr->err = 2;
    ctr->max_dt = in0 - tf;  // get_max_time_step
  }
  // the next step's field: the frames just selected, or this step's when the
  // frames ran out (the run ends there)
  RmField nf = ctr->fs[st % 3];
  if (!ctr->exhausted) {
    nf.frame = f;
    nf.init = 0;
    nf.ratio = ctr->ratio;
  }
  ctr->fs[(st + 1) % 3] = nf;
}

// Row offsets of the fresh grain ids (when there are fresh captures), then the
// step's end: time advance and frame selection.  Nothing later in the step
// reads what rm_finish changes except k_rm_emit, which takes the step number
// as it was.
__global__ void __launch_bounds__(kRmScan) k_rm_rowscan(RmParams P, RmBuffers B) {
  RmCounters* ctr = B.ctr;
  if (ctr->skip) return;
  if (ctr->step_fresh) {
    __shared__ unsigned long long s_sm[kRmScan / 64];
    const RmRows w = rm_rows(P, ctr);
    const int nrow = w.ex * w.ey;
    const int per = (nrow + kRmScan - 1) / kRmScan;
    const int lo = min(nrow, threadIdx.x * per), hi = min(nrow, lo + per);
    unsigned long long mine = 0;
    for (int r = lo; r < hi; r++) mine += (unsigned long long)B.rowcnt[r];
    unsigned long long tot;
    unsigned long long off =
        block_scan_excl(mine, &tot, s_sm) + (unsigned long long)ctr->n_grains;
    for (int r = lo; r < hi; r++) {
      B.rowoff[r] = (int)off;
      off += (unsigned long long)B.rowcnt[r];
    }
    __syncthreads();
    if (threadIdx.x == 0) ctr->n_grains += (int)tot;
  }
  // liquid cells with material after this step's phase: the blocks' boxes
  {
    RmBox liq;
    box_empty(liq);
    const int nb = rm_cell_blocks(P);
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
      const int* slot = B.blk_box + 6 * (size_t)b;
      if (slot[0] > slot[3]) continue;
      box_add(liq, slot[0], slot[1], slot[2]);
      box_add(liq, slot[3], slot[4], slot[5]);
    }
    const RmBox v = block_box_reduce<kRmScan>(liq);
    if (threadIdx.x == 0 && v.lo[0] <= v.hi[0]) {
      RmBox& a = ctr->acc[(ctr->step + 1) & 1][1];
      for (int q = 0; q < 3; q++) {
        a.lo[q] = min(a.lo[q], v.lo[q]);
        a.hi[q] = max(a.hi[q], v.hi[q]);
      }
    }
  }
  if (threadIdx.x == 0) rm_finish(P, B);
}

__global__ void __launch_bounds__(kRmBlock) k_rm_emit(RmParams P, RmBuffers B) {
  RmCounters* ctr = B.ctr;
  if (ctr->skip || ctr->step_cap == 0) return;
  const long long st = ctr->step - 1;  // k_rm_rowscan has advanced the step
  const RmRows w = rm_rows_of(P, ctr->box[st & 1]);
  const int total = w.ex * w.ey * w.ez;
  RmBox cap_box;
  box_empty(cap_box);
  const int stride = gridDim.x * blockDim.x;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < total; q += stride) {
    int i, j, k;
    rm_cell_of(w, q, i, j, k);
    const int c = k * P.nxyG + j * P.nxG + i;
    const unsigned char fl = B.flag[c];
    const int g = B.gid[c];
    const unsigned mk = B.msk[c];
    if (!fl) continue;
    if (g < 0) continue;
    unsigned cap = mk & kSlotBits;
    if (!cap) continue;
    // first new id of this cell: the row's first id plus the fresh captures
    // of the row's flagged cells below it (cells of the row are never
    // captured in this step, so their flags, ids and masks are settled)
    int next = 0;
    if (mk >> 26) {
      next = B.rowoff[(i - w.i0) * w.ey + (j - w.j0)];
      for (int kk = w.k0; kk < k; kk++) {
        const int cc = kk * P.nxyG + j * P.nxG + i;
        if (B.flag[cc] && B.gid[cc] >= 0) next += (int)(B.msk[cc] >> 26);
      }
    }
    const double4 e = B.genv[g];
    const int o = B.gco[g].y;
    double r[9];
#pragma unroll
    for (int x = 0; x < 9; x++) r[x] = B.l2g[9 * (size_t)o + x];
    while (cap) {
      const int s = __builtin_ctz(cap);
      cap &= cap - 1;
      int di, dj, dk;
      slot_delta(s, di, dj, dk);
      const int nb = c + P.off[s];
      const double gx = centre_coord(P.x0, P.h, i + di) - e.x;
      const double gy = centre_coord(P.y0, P.h, j + dj) - e.y;
      const double gz = centre_coord(P.z0, P.h, k + dk) - e.z;
      double out[4];
      child_envelope(r, e.w, P.h, P.trunc, P.sqrt2f, gx, gy, gz, out);
      const double4 ce = make_double4(out[0] + e.x, out[1] + e.y, out[2] + e.z, out[3]);
      int id = B.gid[nb];
      if (id < 0) {
        id = next++;
        B.gid[nb] = id;
        B.gco[id] = make_int2(nb, o);
      } else {
        B.gco[id].y = o;
      }
      B.genv[id] = ce;
      B.oid[nb] = o;
      B.claim[nb] = INT_MAX;
      box_add(cap_box, i + di, j + dj, k + dk);
    }
  }
  block_box_merge(cap_box, &ctr->acc[(st + 1) & 1][2]);
}

// pow(x[i], y) on the device, for checks of the velocity's arithmetic
__global__ void k_pow_selftest(long long n, const double* x, double y, double* out) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = amca_libm::glibc_pow_fast(x[i], y);
}

// ---------------------------------------------------------------------------

hipError_t rm_launch_vts(const RmParams& P, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_rm_vts, dim3(1), dim3(64), 0, s, P, out);
  return hipGetLastError();
}

hipError_t rm_launch_init(const RmParams& P, const RmBuffers& B, hipStream_t s) {
  long long blocks = (P.NG + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_rm_init, dim3((unsigned)blocks), dim3(256), 0, s, P, B);
  return hipGetLastError();
}

const char* const kRmKernelNames[kRmKernels] = {
    "k_rm_field", "k_rm_sur_nodes", "k_rm_sur_elems", "k_rm_phase",   "k_rm_liquid",
    "k_rm_nucleate_mint", "k_rm_claim", "k_rm_resolve", "k_rm_rowscan", "k_rm_emit"};

hipError_t rm_launch_kernel(int q, const RmParams& P, const RmBuffers& B, int site_blocks,
                            hipStream_t s) {
  const unsigned cells = (unsigned)rm_cell_blocks(P);
  // the kernels that work only on flagged cells (a small part of the box):
  // with the box, two blocks per CU (their waves hold many registers)
  const unsigned sparse = P.window && cells > 512u ? 512u : cells;
  long long nbk = (P.nodes + kRmBlock - 1) / kRmBlock;
  if (nbk > 8192) nbk = 8192;
  if (nbk < 1) nbk = 1;
  switch (q) {
    case 0:
      // the site count (grid-stride), the step's box, and without the box the
      // field at every node
      hipLaunchKernelGGL(k_rm_field, dim3(P.window ? 64u : (unsigned)(nbk < 1024 ? nbk : 1024)),
                         dim3(kRmBlock), 0, s, P, B);
      break;
    case 1:
      if (!P.sur) return hipSuccess;
      hipLaunchKernelGGL(k_rm_sur_nodes, dim3((unsigned)(nbk < 1024 ? nbk : 1024)), dim3(kRmBlock),
                         0, s, P, B);
      break;
    case 2:
      if (!P.sur) return hipSuccess;
      hipLaunchKernelGGL(k_rm_sur_elems, dim3((unsigned)(nbk < 1024 ? nbk : 1024)), dim3(kRmBlock),
                         0, s, P, B);
      break;
    case 3: hipLaunchKernelGGL(k_rm_phase, dim3(cells), dim3(kRmBlock), 0, s, P, B); break;
    case 4: hipLaunchKernelGGL(k_rm_liquid, dim3(cells), dim3(kRmBlock), 0, s, P, B); break;
    case 5:
      hipLaunchKernelGGL(k_rm_nucleate_mint, dim3(sparse + (unsigned)site_blocks), dim3(kRmBlock), 0,
                         s, P, B, site_blocks);
      break;
    case 6: hipLaunchKernelGGL(k_rm_claim, dim3(sparse), dim3(kRmBlock), 0, s, P, B); break;
    case 7: hipLaunchKernelGGL(k_rm_resolve, dim3(cells), dim3(kRmBlock), 0, s, P, B); break;
    case 8: hipLaunchKernelGGL(k_rm_rowscan, dim3(1), dim3(kRmScan), 0, s, P, B); break;
    case 9: hipLaunchKernelGGL(k_rm_emit, dim3(sparse), dim3(kRmBlock), 0, s, P, B); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t rm_launch_sur_error(const RmParams& P, const RmBuffers& B, hipStream_t s) {
  long long cb = (P.NG + kRmBlock - 1) / kRmBlock;
  if (cb > 4096) cb = 4096;
  hipLaunchKernelGGL(k_rm_sur_error, dim3((unsigned)cb), dim3(kRmBlock), 0, s, P, B);
  return hipGetLastError();
}

hipError_t rm_launch_complete(const RmParams& P, const RmBuffers& B, hipStream_t s) {
  long long cb = (P.NG + kRmBlock - 1) / kRmBlock;
  if (cb > 16384) cb = 16384;
  hipLaunchKernelGGL(k_rm_complete, dim3((unsigned)cb), dim3(kRmBlock), 0, s, P, B);
  return hipGetLastError();
}

hipError_t rm_launch_sur_hot(const RmParams& P, const RmBuffers& B, int* out, hipStream_t s) {
  long long cb = ((long long)P.nzG * P.nyG * P.sur_nxi + kRmBlock - 1) / kRmBlock;
  if (cb > 8192) cb = 8192;
  const double thr = P.Tm - P.sur_band - P.hot_margin;
  hipLaunchKernelGGL(k_rm_sur_hot, dim3((unsigned)cb), dim3(kRmBlock), 0, s, P, B, thr,
                     reinterpret_cast<RmBox*>(out));
  return hipGetLastError();
}

hipError_t launch_pow_selftest(long long n, const double* x, double y, double* out,
                               hipStream_t s) {
  long long blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_pow_selftest, dim3((unsigned)blocks), dim3(256), 0, s, n, x, y, out);
  return hipGetLastError();
}

hipError_t rm_launch_step(const RmParams& P, const RmBuffers& B, int site_blocks, hipStream_t s) {
  for (int q = 0; q < kRmKernels; q++) {
    const hipError_t e = rm_launch_kernel(q, P, B, site_blocks, s);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace amca
