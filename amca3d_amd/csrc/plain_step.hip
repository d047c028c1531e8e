// Kernels of the plain CA step (realm type `cellular_automata`) for gfx950.
//
// One step is the reference's calculate_time_step + single_integration_step
// (CellularAutomataManager_3D.cpp:440-669).  The reference sweeps a std::list of
// active grains in creation order and appends every captured cell's new grain
// to the same list, so new grains are swept later in the same step.  Grains are
// always appended, so list order is grain-id order, and the sweep is a
// breadth-first cascade of "waves":
//   wave 0      = the active list after nucleation (ids ascending),
//   wave w + 1  = the grains created while sweeping wave w (ids ascending).
// Within a wave the first grain (smallest id) whose envelope holds a liquid
// neighbour takes it; new ids follow (parent id, neighbour slot) order; a grain
// stays active iff one of its liquid neighbours is outside its envelope and not
// taken by a smaller id of the same wave.  The kernels compute exactly that:
//   k_count/k_compact: wave 0's list, the active slots in brick order;
//   k_claim         : per wave entry, liquid test of the 26 neighbours on the
//                     1-bit liquid map, envelope tests, and an atomicMin of the
//                     grain id into each enclosed neighbour's claim word;
//   k_resolve       : per entry, captured / keep decisions from the claims;
//                     capturing grains go to a parent list with their counts;
//   k_idscan        : first child of each capturing grain, a scan over ids;
//   k_emit          : the child records and the next wave's list;
//   k_tail          : waves >= 2 (rare) in one work-group, then the time advance.
// Waves 0 and 1 each run k_claim, k_resolve, k_idscan and k_emit.
// Claims need no reset: every claimed cell is captured in the same wave and is
// never liquid again.
#include <hip/hip_runtime.h>
#include <climits>

#include "device_util.h"
#include "envelope.h"
#include "plain_step.h"

namespace amca {

// Growth velocity of the plain path: uniform undercooling at time t
// (ProblemPhysics.cpp:117-135 with compute_cell_temperature :163-169).
__device__ __forceinline__ double plain_velocity(const DevParams& P, double t) {
  double T = P.T0 + P.Tdot * t;
  double dT = P.Tm - T;
  double v = 0.0;
  if (dT > 0) {
    v = P.a1 * dT + P.a2 * dT * dT + P.a3 * dT * dT * dT;
    if (v < 0.0) v = 0.0;
  }
  return v;
}

// ---------------------------------------------------------------------------
// grid initialisation: CellularAutomataManager::initialize_cells and
// CellularAutomataManager_3D::initialize_boundaries (ghost faces carry 0).

__device__ __forceinline__ bool is_ghost(const DevParams& P, long long c) {
  int k = (int)(c / P.nxyG);
  int rem = (int)(c - (long long)k * P.nxyG);
  int j = rem / P.nxG;
  int i = rem - j * P.nxG;
  return i == 0 || j == 0 || k == 0 || i == P.nxG - 1 || j == P.nyG - 1 || k == P.nzG - 1;
}

__global__ void k_init_grid(DevParams P, DevBuffers B) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < P.NG; c += stride) {
    int v = is_ghost(P, c) ? 0 : -1;
    B.gid[c] = v;
    B.oid[c] = v;
  }
  // claim words, brick order, padding included: no claim
  const long long nclaim = 64ll * P.nbX * P.nbY * P.nbZ;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < nclaim; q += stride)
    B.claim[q] = INT_MAX;
  // brick map: every interior cell liquid, nothing claimed
  const long long nb = (long long)P.nbX * P.nbY * P.nbZ;
  for (long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += stride) {
    const int bx = (int)(b % P.nbX);
    const int by = (int)((b / P.nbX) % P.nbY);
    const int bz = (int)(b / ((long long)P.nbX * P.nbY));
    unsigned long long bits = 0;
    for (int q = 0; q < 64; q++) {
      const int i = 4 * bx + (q & 3), j = 4 * by + ((q >> 2) & 3), k = 4 * bz + (q >> 4);
      if (i >= 1 && i <= P.nx && j >= 1 && j <= P.ny && k >= 1 && k <= P.nz) bits |= 1ull << q;
    }
    B.bmap[2 * b] = bits;
    B.bmap[2 * b + 1] = 0ull;
    B.amap[b] = 0ull;
  }
}

// ---------------------------------------------------------------------------
// nucleation (CellularAutomataManager::nucleate_at_nucleation_sites,
// CellularAutomataManager.cpp:219-297) and the time step
// (CellularAutomataManager_3D::calculate_time_step, :440-511).
//
// Sites are visited in ascending cell order; a liquid site whose undercooling
// exceeds its critical value nucleates with the next orientation row and the
// next grain id; a non-liquid site is dropped.  Ranks come from the look-back
// scan, so ids follow cell order exactly as in the reference loop.

// Index of slot q's record.  A 128-byte line holds four 32-byte records; with
// this order (x-bit 1 and y-bit 0 of the lane exchanged) they are a 2 x 2
// patch of cells in x and y rather than a row of four in x, which holds more
// cells of a growth front per line on average.
__device__ __forceinline__ long long rec_slot(long long q) {
  return (q & ~6ll) | ((q & 2) << 1) | ((q & 4) >> 1);
}

// Whether ghosted cell c is liquid: its bit in the brick map (ghost cells
// never have one).
__device__ __forceinline__ bool cell_liquid(const DevParams& P, const DevBuffers& B, int c) {
  int i, j, k;
  cell_ijk(P, c, i, j, k);
  return (B.bmap[brick_word(P, i, j, k)] & brick_bit(i, j, k)) != 0ull;
}

// Kernel bodies take the domain by reference and their block number / block
// count explicitly, so that one launch can also serve many domains (the
// batched kernels `kb_*` map each block to a (domain, block, count) entry).
__device__ __forceinline__ void nucleate_body(const DevParams& P, const DevBuffers& B, int blk,
                                              int nblk) {
  DevCounters* ctr = B.ctr;
  const double t = ctr->t;
  // an exhausted orientation table (err) stops the run like termination does
  const int skip = (t < P.final_time && !ctr->done && !ctr->err) ? 0 : 1;
  if (blk == 0 && threadIdx.x == 0) ctr->skip = skip;
  if (skip) return;

  __shared__ int s_bid;
  __shared__ unsigned long long s_sm[kBlock / 64];
  __shared__ unsigned long long s_excl;
  if (threadIdx.x == 0) s_bid = (int)atomicAdd(&ctr->tickets[0], 1u);
  __syncthreads();
  const int bid = s_bid;

  const int sel = ctr->site_sel;
  const int n = ctr->n_sites;
  const int ng0 = ctr->n_grains, no0 = ctr->n_orient, nA0 = ctr->nA_in;
  const int* __restrict__ sc = B.site_cell[sel];
  const double* __restrict__ sd = B.site_dt[sel];
  int* __restrict__ nc = B.site_cell[sel ^ 1];
  double* __restrict__ nd = B.site_dt[sel ^ 1];
  const double T = P.T0 + P.Tdot * t;
  const double under = P.Tm - T;

  int lo, hi;
  chunk_of(n, bid, nblk, kBlock, lo, hi);
  unsigned long long mine = 0;
  for (int i = lo + threadIdx.x; i < hi; i += kBlock) {
    const bool liq = cell_liquid(P, B, sc[i]);
    const bool nuc = liq && (under > sd[i]);
    mine += ((unsigned long long)nuc << 32) | (unsigned long long)(liq && !nuc);
  }
  unsigned long long agg;
  block_scan_excl(mine, &agg, s_sm);
  if (threadIdx.x < 64) {
    unsigned long long e = lookback(B.lb_flag, B.lb_agg, B.lb_pref, bid,
                                    (unsigned)(ctr->step * kLbSites + 0), agg);
    if (threadIdx.x == 0) s_excl = e;
  }
  __syncthreads();
  unsigned long long run = s_excl;
  for (int base = lo; base < hi; base += kBlock) {
    int i = base + threadIdx.x;
    int c = 0;
    bool nuc = false, keep = false;
    if (i < hi) {
      c = sc[i];
      const bool liq = cell_liquid(P, B, c);
      nuc = liq && (under > sd[i]);
      keep = liq && !nuc;
    }
    unsigned long long tot;
    unsigned long long off =
        run + block_scan_excl(((unsigned long long)nuc << 32) | (unsigned long long)keep, &tot, s_sm);
    if (nuc) {
      int r = (int)(off >> 32);
      int o = no0 + r;
      if (o < P.orient_cap) {
        int ci, cj, ck;
        cell_ijk(P, c, ci, cj, ck);
        const int id = ng0 + r;
        const long long q = claim_index(P, ci, cj, ck);
        SlotRec r;
        r.x = centre_coord(P.x0, P.h, ci);
        r.y = centre_coord(P.y0, P.h, cj);
        r.z = centre_coord(P.z0, P.h, ck);
        r.id = id;
        r.o = o;
        B.rec[rec_slot(q)] = r;
        B.sL[q] = 0.0;
        B.skey[q] = id;
        brick_clear(P, B.bmap, 0, ci, cj, ck);
        atomicOr(&B.amap[q >> 6], 1ull << (q & 63));
      }
    }
    if (keep) {
      int k = (int)(off & 0xffffffffull);
      nc[k] = c;
      nd[k] = sd[i];
    }
    run += tot;
  }
  if (bid == nblk - 1)
    for (int q = threadIdx.x; q < 64; q += kBlock) {
      ctr->claim_lo[q] = INT_MAX;
      ctr->claim_hi[q] = -1;
    }
  if (threadIdx.x == 0 && bid == nblk - 1) {
    const int nn = (int)(run >> 32);
    const int nk = (int)(run & 0xffffffffull);
    if (no0 + nn > P.orient_cap) {
      // nuclei beyond the table were not written: stop here, so that no
      // later kernel of this step or a later one reads their list entries
      ctr->err = 1;
      ctr->skip = 1;
      ctr->tickets[0] = 0u;
      return;
    }
    const int nA = nA0 + nn;
    ctr->n_grains = ng0 + nn;
    ctr->n_orient = no0 + nn;
    ctr->nA_in = nA;
    ctr->n_sites = nk;
    ctr->site_sel = sel ^ 1;
    ctr->step_nuc = nn;
    ctr->step_sites = n;
    ctr->step_nA0 = nA;
    ctr->step_cap = 0;
    ctr->step_wave = 0;

    ctr->nA_out = 0;
    ctr->nW[0] = 0;
    ctr->nW[1] = 0;
    // maximum velocity over active grains, floor 1e-10 (:449-506), then
    // dt = min(factor*h/maxV, maxDt) (:507)
    const double v = plain_velocity(P, t);
    double maxV = 1e-10;
    if (nA > 0 && v > maxV) maxV = v;
    const double a = P.tsf * P.h / maxV;
    ctr->dt = (P.max_dt < a) ? P.max_dt : a;
    ctr->v = v;
    ctr->tickets[0] = 0u;
  }
}

// ---------------------------------------------------------------------------
// capture waves

// Waves 0 and 1 run as "list waves" over many work-groups: wave 0 over the
// active slots in brick order (k_count / k_compact), wave 1 over the slots of
// wave 0's children in the order wave 0's k_emit wrote them (parent blocks in
// brick order).  The processing order of a wave does not change its outcome
// (claims are an atomicMin, each entry's decision depends only on the claim
// words); its children are numbered by an exclusive scan over the capturing
// grains' ids (k_idscan).
__device__ __forceinline__ void list_wave(const DevBuffers& B, int wave, const int*& list,
                                          int& n) {
  const DevCounters* ctr = B.ctr;
  if (wave == 0) {
    list = B.alist;
    n = ctr->nA_in;
  } else {
    list = B.listW[0];
    n = ctr->nW[0];
  }
}

// Envelope tests of grain g (orientation o, envelope e) at cell (ci, cj, ck)
// against its liquid neighbours lm: each neighbour inside the envelope gets an
// atomicMin of g into its claim word and its claimed bit.  Returns the inside
// bits (none when L == 0: such an envelope captures nothing).
//
// inside_envelope's arithmetic, reorganised without changing a value: along
// each axis the neighbour centres sit at three offsets from the envelope
// centre, and to_local multiplies each offset by the same three rotation
// entries, so the 27 products are formed once; a neighbour's local
// coordinates are then sums of three of them, added in to_local's order
// (its leading 0.0 + p is p up to the sign of a zero, which fabs drops).
__device__ __forceinline__ unsigned claim_tests(const DevParams& P, const DevBuffers& B, int g,
                                                int o, int ci, int cj, int ck, unsigned lm,
                                                const double4& e, WordStage* st) {
  if (lm == 0 || e.w == 0) return 0u;
  double r[9];
  const double* rp = B.l2g + 9 * (size_t)o;
#pragma unroll
  for (int q = 0; q < 9; q++) r[q] = rp[q];
  double px[3][3], py[3][3], pz[3][3];  // [offset + 1][local axis]
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double gx = centre_coord(P.x0, P.h, ci + d - 1) - e.x;
    const double gy = centre_coord(P.y0, P.h, cj + d - 1) - e.y;
    const double gz = centre_coord(P.z0, P.h, ck + d - 1) - e.z;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      px[d][c] = gx * r[c];
      py[d][c] = gy * r[3 + c];
      pz[d][c] = gz * r[6 + c];
    }
  }
  // x + y parts once per (di, dj) column, then the three z offsets
  unsigned im = 0;
#pragma unroll
  for (int dj = -1; dj <= 1; dj++) {
#pragma unroll
    for (int di = -1; di <= 1; di++) {
      unsigned col = 0;
#pragma unroll
      for (int dk = -1; dk <= 1; dk++) {
        const int idx = (dk + 1) * 9 + (dj + 1) * 3 + (di + 1);
        if (idx != 13) col |= 1u << (idx < 13 ? idx : idx - 1);
      }
      if (!(lm & col)) continue;
      const double sx = px[di + 1][0] + py[dj + 1][0];
      const double sy = px[di + 1][1] + py[dj + 1][1];
      const double sz = px[di + 1][2] + py[dj + 1][2];
#pragma unroll
      for (int dk = -1; dk <= 1; dk++) {
        const int idx = (dk + 1) * 9 + (dj + 1) * 3 + (di + 1);
        if (idx == 13) continue;
        const int s = idx < 13 ? idx : idx - 1;
        if (!((lm >> s) & 1u)) continue;
        const double lx = sx + pz[dk + 1][0];
        const double ly = sy + pz[dk + 1][1];
        const double lz = sz + pz[dk + 1][2];
        if ((fabs(lx) + fabs(ly)) + fabs(lz) <= e.w) im |= 1u << s;
      }
    }
  }
  unsigned rest = im;
  while (rest) {
    const int s = __builtin_ctz(rest);
    rest &= rest - 1;
    int di, dj, dk;
    slot_delta(s, di, dj, dk);
    atomicMin(&B.claim[claim_index(P, ci + di, cj + dj, ck + dk)], g);
    const int wd = brick_word(P, ci + di, cj + dj, ck + dk) + 1;
    const unsigned long long bt = brick_bit(ci + di, cj + dj, ck + dk);
    if (!st || !stage_add(*st, wd, bt)) atomicOr(&B.bmap[wd], bt);
  }
  return im;
}

// The active slots in brick order.  k_count leaves the active-cell count of
// each block of bricks; k_compact block b starts from the sum of the counts of
// the blocks before it, and each thread lists its own brick's active slots
// (bricks in 256-brick tiles, one block scan per tile).  The list's length is
// the active count.
__device__ __forceinline__ void cmp_range(const DevParams& P, int blk, int nblk, long long& lo,
                                          long long& hi) {
  const long long nB = (long long)P.nbX * P.nbY * P.nbZ;
  const long long per = ((nB + nblk - 1) / nblk + kBlock - 1) / kBlock * kBlock;
  lo = (long long)blk * per < nB ? (long long)blk * per : nB;
  hi = lo + per < nB ? lo + per : nB;
}

__device__ __forceinline__ void count_body(const DevParams& P, const DevBuffers& B, int blk,
                                           int nblk) {
  if (B.ctr->skip) return;
  __shared__ unsigned long long s_sm[kBlock / 64];
  long long lo, hi;
  cmp_range(P, blk, nblk, lo, hi);
  unsigned long long mine = 0;
  for (long long b = lo + threadIdx.x; b < hi; b += kBlock) mine += __popcll(B.amap[b]);
  unsigned long long tot;
  block_scan_excl(mine, &tot, s_sm);
  if (threadIdx.x == 0) B.cmptot[blk] = (unsigned)tot;
}

__device__ __forceinline__ void compact_body(const DevParams& P, const DevBuffers& B, int blk,
                                             int nblk) {
  if (B.ctr->skip) return;
  __shared__ unsigned long long s_sm[kBlock / 64];
  long long lo, hi;
  cmp_range(P, blk, nblk, lo, hi);
  unsigned long long pre = 0;
  for (int q = threadIdx.x; q < blk; q += kBlock) pre += B.cmptot[q];
  unsigned long long run;
  block_scan_excl(pre, &run, s_sm);
  if (B.cmptot[blk] == 0u) return;
  // Per tile of 256 bricks, one brick per thread: each thread puts its brick's
  // active slots, lowest first, at the brick's position in its wave's stretch
  // of the output (a wave scan of the active counts; the totals of the waves
  // before come from LDS, double-buffered so that one barrier per tile
  // suffices).  A wave stages its stretch in LDS and writes it out coalesced;
  // a stretch too long for the stage is written directly.
  constexpr int kWaveStage = 2048;
  __shared__ int s_stage[kBlock / 64][kWaveStage];
  __shared__ unsigned s_wt[2][kBlock / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int par = 0;
  for (long long base = lo; base < hi; base += kBlock, par ^= 1) {
    const long long b = base + threadIdx.x;
    unsigned long long w = b < hi ? B.amap[b] : 0ull;
    const unsigned pop = (unsigned)__popcll(w);
    unsigned inc = pop;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const unsigned u = __shfl_up(inc, d, 64);
      if (lane >= d) inc += u;
    }
    if (lane == 63) s_wt[par][wv] = inc;
    __syncthreads();
    unsigned wbase = 0, tot = 0;
#pragma unroll
    for (int q = 0; q < kBlock / 64; q++) {
      const unsigned t = s_wt[par][q];
      if (q < wv) wbase += t;
      tot += t;
    }
    const unsigned wtot = s_wt[par][wv];
    int* __restrict__ gout = B.alist + run + wbase;
    const int q0 = (int)(64 * b);
    if (wtot <= (unsigned)kWaveStage) {
      int* st = s_stage[wv];
      unsigned pos = inc - pop;
      while (w) {
        st[pos++] = q0 + __builtin_ctzll(w);
        w &= w - 1;
      }
      // the wave's LDS operations complete in order: its reads below see
      // every lane's writes above
      __builtin_amdgcn_wave_barrier();
      for (unsigned i = lane; i < wtot; i += 64) gout[i] = st[i];
      __builtin_amdgcn_wave_barrier();
    } else {
      int* out = gout + (inc - pop);
      while (w) {
        *out++ = q0 + __builtin_ctzll(w);
        w &= w - 1;
      }
    }
    run += tot;
  }
}

// Cell of a slot.
__device__ __forceinline__ void slot_ijk(const DevParams& P, long long q, int& i, int& j, int& k) {
  int bx, by, bz;
  brick_xyz(P, q >> 6, bx, by, bz);
  const int l = (int)(q & 63);
  i = 4 * bx + (l & 3);
  j = 4 * by + ((l >> 2) & 3);
  k = 4 * bz + (l >> 4);
}

// The id scan of a list wave covers the 16-id groups [g0, g0 + n4) that hold
// the ids of the grains that can capture in it: for wave 0 the range of the
// grains that claimed a cell (a grain captures only cells it claimed), which
// k_claim leaves in 64 (min, max) slots; for wave 1 wave 0's children.  Block b
// takes the groups g0 + [b per, (b + 1) per), per a whole number of 256-thread
// tiles.  The counters it reads are written by earlier kernels, so every block
// sees the same range.  Called by every thread of the block.
struct IdRange {
  long long g0, n4, per;
};
__device__ __forceinline__ IdRange id_range(const DevParams& P, const DevCounters* ctr, int wave) {
  long long lo, hi;
  if (wave == 0) {
    __shared__ int s_lo, s_hi;
    if (threadIdx.x < 64) {
      int a = ctr->claim_lo[threadIdx.x], b = ctr->claim_hi[threadIdx.x];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        a = min(a, __shfl_xor(a, o, 64));
        b = max(b, __shfl_xor(b, o, 64));
      }
      if (threadIdx.x == 0) {
        s_lo = a;
        s_hi = b;
      }
    }
    __syncthreads();
    lo = s_lo;
    hi = (long long)s_hi + 1;
  } else {
    lo = ctr->w1_lo;
    hi = lo + ctr->w1_n;
  }
  IdRange r;
  r.g0 = hi > lo ? lo >> 4 : 0;
  r.n4 = hi > lo ? ((hi + 15) >> 4) - r.g0 : 0;
  r.per = ((r.n4 + P.ids_blocks - 1) / P.ids_blocks + kBlock - 1) / kBlock * kBlock;
  if (r.per == 0) r.per = kBlock;
  return r;
}

// Envelope tests of one entry of a wave >= 1 (the grain in slot q).  Its own
// cell's claimed bit (set by the wave that captured it) stays set: the
// claimed word is read only where the liquid word has the cell's bit, and a
// captured cell is never liquid again.
template <bool kCoherent>
__device__ __forceinline__ unsigned long long claim_entry(const DevParams& P, const DevBuffers& B,
                                                          int q, WordStage* st) {
  int ci, cj, ck;
  slot_ijk(P, q, ci, cj, ck);
  const SlotRec r = B.rec[rec_slot(q)];
  const unsigned lm = brick_bits26<0, kCoherent>(P, B.bmap, ci, cj, ck);
  const unsigned im =
      claim_tests(P, B, r.id, r.o, ci, cj, ck, lm, make_double4(r.x, r.y, r.z, B.sL[q]), st);
  return ((unsigned long long)im << 32) | lm;
}

// A list wave's envelope tests, one entry per thread.  Wave 0 first grows
// each active grain (Grain::grow, Grain.h:27, over the whole list as in
// :521-538); wave 1's grains were created in this step and do not grow.
// Neighbouring entries sit in the same bricks, so their map words and
// records share cache lines.
__device__ __forceinline__ void claim_body(const DevParams& P, const DevBuffers& B, int wave,
                                           int blk, int nblk) {
  const DevCounters* ctr = B.ctr;
  if (ctr->skip) return;
  // the id scan's per-block capture totals start from zero (k_resolve adds)
  for (int q = blk * kBlock + threadIdx.x; q < P.ids_blocks; q += nblk * kBlock) B.idtot[q] = 0u;
  const int* list;
  int n;
  list_wave(B, wave, list, n);
  int lo, hi;
  chunk_of(n, blk, nblk, kBlock, lo, hi);
  if (lo >= hi) return;
  // the claimed bits, staged per work-group
  __shared__ WordStage s_st;
  stage_init(s_st);
  __syncthreads();
  if (wave == 0) {
    const double inc = ctr->dt * ctr->v;
    int id_lo = INT_MAX, id_hi = -1;  // this block's claiming grains
    for (int t = lo + threadIdx.x; t < hi; t += kBlock) {
      const long long q = list[t];
      int i, j, k;
      slot_ijk(P, q, i, j, k);
      const SlotRec r = B.rec[rec_slot(q)];
      const double L = B.sL[q] + inc;
      B.sL[q] = L;
      const unsigned lm = brick_bits26<0, false>(P, B.bmap, i, j, k);
      const unsigned im =
          claim_tests(P, B, r.id, r.o, i, j, k, lm, make_double4(r.x, r.y, r.z, L), &s_st);
      B.masks[t] = ((unsigned long long)im << 32) | lm;
      if (im) {
        id_lo = min(id_lo, r.id);
        id_hi = max(id_hi, r.id);
      }
    }
    // the block's claimer id range into one of the 64 slots
    __shared__ int s_lo, s_hi;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      id_lo = min(id_lo, __shfl_xor(id_lo, o, 64));
      id_hi = max(id_hi, __shfl_xor(id_hi, o, 64));
    }
    if (threadIdx.x == 0) {
      s_lo = INT_MAX;
      s_hi = -1;
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0 && id_hi >= 0) {
      atomicMin(&s_lo, id_lo);
      atomicMax(&s_hi, id_hi);
    }
    __syncthreads();
    if (threadIdx.x == 0 && s_hi >= 0) {
      atomicMin(&B.ctr->claim_lo[blk & 63], s_lo);
      atomicMax(&B.ctr->claim_hi[blk & 63], s_hi);
    }
  } else {
    for (int t = lo + threadIdx.x; t < hi; t += kBlock)
      B.masks[t] = claim_entry<false>(P, B, list[t], &s_st);
  }
  __syncthreads();
  stage_flush_or(s_st, B.bmap);
}

// Captured slots and the keep bit of the grain in slot q at (ci, cj, ck) from
// its masks m and the claimed bits cmask of its neighbours.  A liquid slot
// nobody claimed in this wave needs no claim word (its claim is "none", larger
// than any id); only claimed slots read theirs, all loads issued before any is
// used.  The grain's id is read (into g) only when a claim word is compared
// with it, which every capture needs.
template <bool kCoherent>
__device__ __forceinline__ unsigned resolve_bits(const DevParams& P, const DevBuffers& B,
                                                 long long q, unsigned long long m, unsigned cmask,
                                                 int ci, int cj, int ck, int& g) {
  const unsigned lm = (unsigned)m, im = (unsigned)(m >> 32);
  if (lm == 0) return 0u;
  const unsigned cm = cmask & lm;
  // an unclaimed liquid neighbour outside the envelope keeps the grain active
  unsigned keep = (lm & ~im & ~cm) ? 1u : 0u;
  const unsigned need = keep ? (cm & im) : cm;
  if (need == 0) return keep << 31;
  g = B.skey[q];
  // claim_index of neighbour (di, dj, dk) = cx[di + 1] + cy[dj + 1] + cz[dk + 1]
  unsigned cx[3], cy[3], cz[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const unsigned x = (unsigned)(ci + d - 1), y = (unsigned)(cj + d - 1), z = (unsigned)(ck + d - 1);
    cx[d] = 64u * (x >> 2) + (x & 3u);
    cy[d] = 64u * (unsigned)P.nbX * (y >> 2) + ((y & 3u) << 2);
    cz[d] = 64u * (unsigned)P.nbX * (unsigned)P.nbY * (z >> 2) + ((z & 3u) << 4);
  }
  int cl[26];
#pragma unroll
  for (int s = 0; s < 26; s++) {
    cl[s] = INT_MAX;
    if ((need >> s) & 1u) {
      int di, dj, dk;
      slot_delta(s, di, dj, dk);
      const int* w = B.claim + (cx[di + 1] + cy[dj + 1] + cz[dk + 1]);
      cl[s] = kCoherent ? ld_coherent(w) : *w;
    }
  }
  unsigned cap = 0;
#pragma unroll
  for (int s = 0; s < 26; s++) {
    if ((need >> s) & 1u) {
      if ((im >> s) & 1u) {
        if (cl[s] == g) cap |= 1u << s;
      } else if (cl[s] > g) {
        keep = 1u;
      }
    }
  }
  return cap | (keep << 31);
}

// One entry of a wave >= 1 (slot q); a grain that does not stay active leaves
// the active map.
template <bool kCoherent>
__device__ __forceinline__ unsigned resolve_entry(const DevParams& P, const DevBuffers& B, int q,
                                                  unsigned long long m) {
  unsigned r = 0u;
  if ((unsigned)m != 0u) {
    int ci, cj, ck;
    slot_ijk(P, q, ci, cj, ck);
    int g;
    r = resolve_bits<kCoherent>(P, B, q, m, brick_bits26<1, kCoherent>(P, B.bmap, ci, cj, ck), ci,
                                cj, ck, g);
  }
  if (!(r >> 31)) atomicAnd(&B.amap[q >> 6], ~(1ull << (q & 63)));
  return r;
}

// Inclusive prefix sum over the 64 lanes of a wave (every lane must call it).
__device__ __forceinline__ unsigned wave_incl(unsigned v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned u = __shfl_up(v, d, 64);
    if (lane >= d) v += u;
  }
  return v;
}

// A list wave's decisions, over the same list chunks as its k_emit.  A grain
// that does not stay active leaves the active map; a grain that captured
// writes (id, slot, captured slots, first child in the block) to its block's
// segment of the parent list, which starts at the block's first list entry,
// and its capture count to the id-indexed count array.  Each block leaves
// (parents << 32 | survivors) and its child count in its statistics words
// and adds its captures per id-scan block into idtot; no other counter is
// shared between blocks.
__device__ __forceinline__ void resolve_body(const DevParams& P, const DevBuffers& B, int wave,
                                             int blk, int nblk) {
  const DevCounters* ctr = B.ctr;
  if (ctr->skip) return;
  __shared__ unsigned long long s_pc;         // parents << 32 | children
  __shared__ int s_surv;
  __shared__ unsigned s_hist[kMaxIdsBlocks];  // captures per id-scan block
  if (threadIdx.x == 0) {
    s_pc = 0ull;
    s_surv = 0;
  }
  for (int q = threadIdx.x; q < P.ids_blocks; q += kBlock) s_hist[q] = 0u;
  __syncthreads();
  const int* list;
  int n;
  list_wave(B, wave, list, n);
  int lo, hi;
  chunk_of(n, blk, nblk, kBlock, lo, hi);
  int4* __restrict__ seg = B.plist + lo;
  const IdRange ir = id_range(P, ctr, wave);
  const int lane = threadIdx.x & 63;
  for (int base = lo; base < hi; base += kBlock) {
    const int t = base + threadIdx.x;
    unsigned r = 0u;
    int g = 0;
    long long q = 0;
    if (t < hi) {
      q = list[t];
      const unsigned long long m = B.masks[t];
      if ((unsigned)m != 0u) {
        int i, j, k;
        slot_ijk(P, q, i, j, k);
        r = resolve_bits<false>(P, B, q, m, brick_bits26<1, false>(P, B.bmap, i, j, k), i, j, k, g);
      }
      if (!(r >> 31)) atomicAnd(&B.amap[q >> 6], ~(1ull << (q & 63)));
    }
    const unsigned long long keepm = __ballot(t < hi && (r >> 31));
    if (lane == 0 && keepm) atomicAdd(&s_surv, (int)__popcll(keepm));
    const unsigned cap = r & 0x3ffffffu;
    const unsigned long long capm = __ballot(cap != 0u);
    if (capm) {
      const unsigned pc = (unsigned)__builtin_popcount(cap);
      const unsigned inc = wave_incl(pc);
      const unsigned wtot = __shfl(inc, 63, 64);  // the wave's children
      unsigned long long wb = 0ull;
      if (lane == 0) wb = atomicAdd(&s_pc, ((unsigned long long)__popcll(capm) << 32) | wtot);
      wb = __shfl(wb, 0, 64);
      if (cap) {
        const int pp = (int)(wb >> 32) + (int)__popcll(capm & ((1ull << lane) - 1ull));
        const int cp = (int)(unsigned)wb + (int)(inc - pc);
        seg[pp] = make_int4(g, (int)q, (int)cap, cp);
        B.cnt[g] = (unsigned char)pc;
        atomicAdd(&s_hist[((g >> 4) - ir.g0) / ir.per], pc);
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    B.lb_agg[kStatBase + blk] = (s_pc & 0xffffffff00000000ull) | (unsigned)s_surv;
    B.lb_pref[kStatBase + blk] = s_pc & 0xffffffffull;
  }
  for (int q = threadIdx.x; q < P.ids_blocks; q += kBlock)
    if (s_hist[q]) atomicAdd(&B.idtot[q], s_hist[q]);
}

// A list wave's children are numbered in (parent id, slot) order, as the
// reference's sweep appends them: the first child of a capturing grain is the
// sum of the capture counts of the capturing grains with smaller ids.  The
// counts sit in an id-indexed byte array, and k_resolve also added them up per
// block of this scan (idtot).  Block b starts from the sum of the totals of
// the blocks before it, scans its id groups in tiles of 4096 ids (16 per
// thread, one uint4 of counts), writes each capturing grain's first child
// (wave-relative) and zeroes the counts it has read.  The last block also
// turns the `nres` k_resolve blocks' child counts into each block's first
// child, adds the wave's survivors to the active count, and publishes the
// wave's children (the next id range and list length).
__device__ __forceinline__ void idscan_body(const DevParams& P, const DevBuffers& B, int wave,
                                            int blk, int nres) {
  DevCounters* ctr = B.ctr;
  if (ctr->skip) return;
  __shared__ unsigned long long s_sm[kBlock / 64];
  const int nblk = P.ids_blocks;
  const IdRange ir = id_range(P, ctr, wave);
  const long long lo = (long long)blk * ir.per < ir.n4 ? (long long)blk * ir.per : ir.n4;
  const long long hi = lo + ir.per < ir.n4 ? lo + ir.per : ir.n4;
  // a block past the wave's ids has nothing to scan (the last one still sums
  // the totals and publishes the wave)
  if (lo >= hi && blk != nblk - 1) return;
  unsigned long long pre = 0;
  for (int q = threadIdx.x; q < blk; q += kBlock) pre += B.idtot[q];
  unsigned long long start;
  block_scan_excl(pre, &start, s_sm);
  const unsigned mine = B.idtot[blk];
  uint4* __restrict__ c4 = reinterpret_cast<uint4*>(B.cnt) + ir.g0;
  auto bytesum = [](unsigned v) {
    return (v & 0xffu) + ((v >> 8) & 0xffu) + ((v >> 16) & 0xffu) + (v >> 24);
  };
  unsigned long long run = start;
  if (mine) {
    // each tile's counts are loaded while the tile before is scanned
    uint4 vn = make_uint4(0u, 0u, 0u, 0u);
    if (lo + threadIdx.x < hi) vn = c4[lo + threadIdx.x];
    for (long long base = lo; base < hi; base += kBlock) {
      const long long u = base + threadIdx.x;
      const uint4 v = vn;
      vn = make_uint4(0u, 0u, 0u, 0u);
      if (u + kBlock < hi) vn = c4[u + kBlock];
      const bool any = (v.x | v.y | v.z | v.w) != 0u;
      const unsigned sum = any ? bytesum(v.x) + bytesum(v.y) + bytesum(v.z) + bytesum(v.w) : 0u;
      unsigned long long tot;
      unsigned long long off = run + block_scan_excl(sum, &tot, s_sm);
      if (any) {
        const unsigned w[4] = {v.x, v.y, v.z, v.w};
        const int id = (int)(16 * (ir.g0 + u));
#pragma unroll
        for (int t = 0; t < 16; t++) {
          const unsigned c = (w[t >> 2] >> (8 * (t & 3))) & 0xffu;
          if (c) {
            B.coff[id + t] = (int)off;
            off += c;
          }
        }
        c4[u] = make_uint4(0u, 0u, 0u, 0u);
      }
      run += tot;
    }
  }
  if (blk == nblk - 1) {
    // survivors of the wave and each k_resolve block's first child: thread t
    // takes k_resolve blocks [t per, (t + 1) per) (nres <= kMaxBlocks), loads
    // them all, and one scan of (survivors << 32 | children) numbers them
    // (both totals are below 2^32: at most one per cell)
    constexpr int kPer = kMaxBlocks / kBlock;
    const int per = (nres + kBlock - 1) / kBlock;
    const int q0 = threadIdx.x * per;
    unsigned long long c[kPer];
    unsigned long long both = 0;
#pragma unroll
    for (int x = 0; x < kPer; x++) {
      c[x] = 0;
      if (x < per && q0 + x < nres) {
        c[x] = B.lb_pref[kStatBase + q0 + x];
        both += ((unsigned long long)(unsigned)B.lb_agg[kStatBase + q0 + x] << 32) + c[x];
      }
    }
    unsigned long long tot;
    unsigned long long off = block_scan_excl(both, &tot, s_sm) & 0xffffffffull;
#pragma unroll
    for (int x = 0; x < kPer; x++) {
      if (x < per && q0 + x < nres) {
        B.lb_pref[kStatBase + q0 + x] = off;
        off += c[x];
      }
    }
    const unsigned long long svt = tot >> 32;
    if (threadIdx.x == 0) {
      const unsigned long long nc = start + mine;
      const int gbase = ctr->n_grains;
      ctr->wave_gbase = gbase;
      ctr->n_grains = gbase + (int)nc;
      ctr->nW[wave] = (int)nc;
      ctr->step_cap += (int)nc;
      ctr->nA_out += (int)svt;
      if (wave == 0) {
        ctr->w1_lo = gbase;
        ctr->w1_n = (int)nc;
      }
      if (nc > 0 && ctr->max_wave < wave + 1) ctr->max_wave = wave + 1;
      if (nc > 0 && ctr->step_wave < wave + 1) ctr->step_wave = wave + 1;
    }
  }
}

// A captured cell's new grain: record, L and id in the cell's slot, liquid bit
// cleared, active bit set (CellularAutomataManager_3D::capture_cell,
// :676-707; Grain.cpp:13-18).  Returns the slot.
__device__ __forceinline__ int write_child(const DevParams& P, const DevBuffers& B, int id, int ni,
                                           int nj, int nk, int o, const double4& env) {
  const long long q = claim_index(P, ni, nj, nk);
  SlotRec r;
  r.x = env.x;
  r.y = env.y;
  r.z = env.z;
  r.id = id;
  r.o = o;
  B.rec[rec_slot(q)] = r;
  B.sL[q] = env.w;
  B.skey[q] = id;
  brick_clear(P, B.bmap, 0, ni, nj, nk);
  atomicOr(&B.amap[q >> 6], 1ull << (q & 63));
  return (int)q;
}

// Octahedron::capture_cell (Octahedron.cpp:133-275): the child envelope of
// parent envelope e (orientation row r) for its neighbour slot s.
__device__ __forceinline__ double4 child_of(const DevParams& P, const double* r, const double4& e,
                                            int ci, int cj, int ck, int di, int dj, int dk) {
  const double gx = centre_coord(P.x0, P.h, ci + di) - e.x;
  const double gy = centre_coord(P.y0, P.h, cj + dj) - e.y;
  const double gz = centre_coord(P.z0, P.h, ck + dk) - e.z;
  double out[4];
  child_envelope(r, e.w, P.h, P.trunc, P.sqrt2f, gx, gy, gz, out);
  return make_double4(out[0] + e.x, out[1] + e.y, out[2] + e.z, out[3]);
}

// The children of a parent at (ci, cj, ck) with orientation o and envelope e:
// each captured cell (slots ascending) gets a new grain, ids id0, id0+1, ...,
// with the parent orientation; their slots are listed at kids[pos0 ...].
__device__ __forceinline__ void emit_from(const DevParams& P, const DevBuffers& B, int o,
                                          const double4& e, int ci, int cj, int ck, unsigned cap,
                                          int id0, int* __restrict__ kids, int pos0) {
  double r[9];
  const double* rp = B.l2g + 9 * (size_t)o;
#pragma unroll
  for (int q = 0; q < 9; q++) r[q] = rp[q];
  int j = 0;
  while (cap) {
    const int s = __builtin_ctz(cap);
    cap &= cap - 1;
    int di, dj, dk;
    slot_delta(s, di, dj, dk);
    kids[pos0 + j] =
        write_child(P, B, id0 + j, ci + di, cj + dj, ck + dk, o, child_of(P, r, e, ci, cj, ck, di, dj, dk));
    j++;
  }
}

// The envelope of the grain in slot q.
__device__ __forceinline__ double4 slot_env(const DevBuffers& B, const SlotRec& r, long long q) {
  return make_double4(r.x, r.y, r.z, B.sL[q]);
}

// The children of the parent in slot qp.
__device__ __forceinline__ void emit_children(const DevParams& P, const DevBuffers& B, int qp,
                                              unsigned cap, int id0, int* __restrict__ kids,
                                              int pos0) {
  int ci, cj, ck;
  slot_ijk(P, qp, ci, cj, ck);
  const SlotRec pr = B.rec[rec_slot(qp)];
  emit_from(P, B, pr.o, slot_env(B, pr, qp), ci, cj, ck, cap, id0, kids, pos0);
}

// A list wave's children, one per thread.  A capturing grain's children take
// the ids wave_gbase + first child + 0, 1, ... in slot order.  The block's
// parent-list segment (the same list chunk as its k_resolve block) holds each
// parent's first child within the block, increasing along the segment; per
// tile of 256 parents, every parent stages (parent, neighbour slot) for its
// children in LDS at their block positions, kStage at a time, and the block
// then builds one child per thread, so that a parent with many children does
// not hold up the other lanes of its wave.  Wave 0 lists the children's slots
// for wave 1 by parent block and position in the block (parents in brick
// order); wave 1 lists them for the tail in id order.
__device__ __forceinline__ void emit_body(const DevParams& P, const DevBuffers& B, int wave,
                                          int blk, int nblk) {
  const DevCounters* ctr = B.ctr;
  if (ctr->skip) return;
  constexpr int kStage = 1024;
  __shared__ int2 s_stage[kStage];
  const int* list;
  int n;
  list_wave(B, wave, list, n);
  int lo, hi;
  chunk_of(n, blk, nblk, kBlock, lo, hi);
  const int np = (int)(B.lb_agg[kStatBase + blk] >> 32);
  if (np == 0) return;
  const int cb = (int)B.lb_pref[kStatBase + blk];
  // children of this block: up to the next block's first child, or the wave's
  // total for the last block
  const int cend = (blk + 1 < nblk ? (int)B.lb_pref[kStatBase + blk + 1] : ctr->nW[wave]) - cb;
  const int4* __restrict__ seg = B.plist + lo;
  const int gbase = ctr->wave_gbase;
  int* __restrict__ kids = B.listW[wave];
  for (int pbase = 0; pbase < np; pbase += kBlock) {
    const int pt = pbase + threadIdx.x;
    int4 pe = make_int4(0, 0, 0, 0);
    if (pt < np) pe = seg[pt];
    const int c0 = seg[pbase].w;
    const int c1 = pbase + kBlock < np ? seg[pbase + kBlock].w : cend;
    for (int w0 = c0; w0 < c1; w0 += kStage) {
      if (pt < np) {
        unsigned cap = (unsigned)pe.z;
        int c = pe.w;
        while (cap) {
          const int sl = __builtin_ctz(cap);
          cap &= cap - 1;
          if (c >= w0 && c < w0 + kStage) s_stage[c - w0] = make_int2(pt, sl);
          c++;
        }
      }
      __syncthreads();
      const int m = min(kStage, c1 - w0);
      for (int u = threadIdx.x; u < m; u += kBlock) {
        const int2 it = s_stage[u];
        const int4 pp = seg[it.x];
        const int q = pp.y;
        int ci, cj, ck;
        slot_ijk(P, q, ci, cj, ck);
        const SlotRec pr = B.rec[rec_slot(q)];
        const double4 e = slot_env(B, pr, q);
        double r[9];
        const double* rp = B.l2g + 9 * (size_t)pr.o;
#pragma unroll
        for (int x = 0; x < 9; x++) r[x] = rp[x];
        int di, dj, dk;
        slot_delta(it.y, di, dj, dk);
        const int c = w0 + u;         // block position of the child
        const int rank = c - pp.w;    // its rank among its parent's children
        const int p0 = B.coff[pp.x];  // the parent's first child in the wave
        const int pos = wave == 0 ? cb + c : p0 + rank;
        kids[pos] = write_child(P, B, gbase + p0 + rank, ci + di, cj + dj, ck + dk, pr.o,
                                child_of(P, r, e, ci, cj, ck, di, dj, dk));
      }
      __syncthreads();
    }
  }
}

// Waves from 2 on, in one work-group (these are rare and short), then the end
// of the step: time_ += dT_, iStep_++, liquid count (:650-659).
constexpr int kTailBlock = 1024;

__device__ __forceinline__ void tail_body(const DevParams& P, const DevBuffers& B) {
  DevCounters* ctr = B.ctr;
  if (ctr->skip) return;
  __shared__ unsigned long long s_sm[kTailBlock / 64];
  __shared__ int s_n, s_gbase, s_sbase;
  const int n_wave1 = ctr->nW[0];  // wave-0 captures, before wave 2 reuses the slot
  int wave = 2;
  while (true) {
    const int in = (wave - 1) & 1;
    if (threadIdx.x == 0) {
      s_n = ctr->nW[in];
      s_gbase = ctr->n_grains;
      s_sbase = ctr->nA_out;
    }
    __syncthreads();
    const int n = s_n;
    if (n == 0) break;
    const int* list = B.listW[in];
    int* kids = B.listW[wave & 1];
    for (int i = threadIdx.x; i < n; i += kTailBlock)
      B.masks[i] = claim_entry<true>(P, B, list[i], nullptr);
    // one work-group: a work-group-scope fence orders its phases (the claim
    // and map words are read with agent-scope loads); an agent-scope fence
    // would also write back and invalidate the L2
    __threadfence_block();
    __syncthreads();
    unsigned long long run = 0;
    for (int base = 0; base < n; base += kTailBlock) {
      const int i = base + threadIdx.x;
      unsigned r = 0;
      int q = 0;
      if (i < n) {
        q = list[i];
        r = resolve_entry<true>(P, B, q, B.masks[i]);
      }
      const unsigned cap = r & 0x3ffffffu;
      unsigned long long tot;
      unsigned long long off =
          run + block_scan_excl(((unsigned long long)__builtin_popcount(cap) << 32) | (r >> 31),
                                &tot, s_sm);
      if (cap) {
        const int pos = (int)(off >> 32);
        emit_children(P, B, q, cap, s_gbase + pos, kids, pos);
      }
      run += tot;
    }
    __threadfence_block();
    __syncthreads();
    if (threadIdx.x == 0) {
      const int nc = (int)(run >> 32);
      ctr->n_grains = s_gbase + nc;
      ctr->nA_out = s_sbase + (int)(run & 0xffffffffull);
      ctr->nW[wave & 1] = nc;
      ctr->step_cap += nc;
      if (nc > 0 && ctr->max_wave < wave + 1) ctr->max_wave = wave + 1;
      if (nc > 0 && ctr->step_wave < wave + 1) ctr->step_wave = wave + 1;
    }
    __syncthreads();
    wave++;
  }
  if (threadIdx.x == 0) {
    const double t0 = ctr->t;
    const double dt = ctr->dt;
    ctr->t = t0 + dt;
    const long long st = ctr->step;
    ctr->step = st + 1;
    const long long solid = ctr->n_solid + ctr->step_nuc + ctr->step_cap;
    ctr->n_solid = solid;
    const long long liquid = P.N - solid;
    if (liquid == 0) ctr->done = 1;
    if (B.log && st < ctr->log_cap) {
      StepLogRow row;
      row.time_before = t0;
      row.dt = dt;
      row.n_nucleated = ctr->step_nuc;
      row.n_active_before = ctr->step_nA0;
      row.n_captured = ctr->step_cap;
      row.n_active_after = ctr->nA_out;
      row.n_liquid = liquid;
      row.n_wave1 = n_wave1;
      row.max_wave = ctr->step_wave;
      row.n_sites = ctr->step_sites;
      B.log[st] = row;
    }
    ctr->parity ^= 1;
    ctr->nA_in = ctr->nA_out;
    ctr->nA_out = 0;
  }
}

// ---------------------------------------------------------------------------

__global__ void k_envelope_selftest(int n, const double* env4, const double* l2g9,
                                    const double* cell3, double h, double trunc, double sqrt2f,
                                    int* inside, double* child4) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* e = env4 + 4 * (size_t)i;
  const double* r = l2g9 + 9 * (size_t)i;
  const double* x = cell3 + 3 * (size_t)i;
  double gx = x[0] - e[0], gy = x[1] - e[1], gz = x[2] - e[2];
  bool in = inside_envelope(r, e[3], gx, gy, gz);
  inside[i] = in ? 1 : 0;
  double out[4] = {0.0, 0.0, 0.0, 0.0};
  if (in) child_envelope(r, e[3], h, trunc, sqrt2f, gx, gy, gz, out);
  child4[4 * (size_t)i + 0] = in ? out[0] + e[0] : 0.0;
  child4[4 * (size_t)i + 1] = in ? out[1] + e[1] : 0.0;
  child4[4 * (size_t)i + 2] = in ? out[2] + e[2] : 0.0;
  child4[4 * (size_t)i + 3] = out[3];
}

// ---------------------------------------------------------------------------
// launchers

// ---------------------------------------------------------------------------
// single-domain kernels: the whole grid serves one domain; list kernels give
// each XCD a contiguous eighth of the list

__global__ void __launch_bounds__(kBlock) k_nucleate(DevParams P, DevBuffers B) {
  nucleate_body(P, B, blockIdx.x, gridDim.x);
}
__global__ void __launch_bounds__(kBlock) k_claim(DevParams P, DevBuffers B, int wave) {
  claim_body(P, B, wave, xcd_block(), gridDim.x);
}
__global__ void __launch_bounds__(kBlock) k_resolve(DevParams P, DevBuffers B, int wave) {
  resolve_body(P, B, wave, xcd_block(), gridDim.x);
}
__global__ void __launch_bounds__(kBlock) k_emit(DevParams P, DevBuffers B, int wave) {
  emit_body(P, B, wave, xcd_block(), gridDim.x);
}
__global__ void __launch_bounds__(kTailBlock) k_tail(DevParams P, DevBuffers B) { tail_body(P, B); }
__global__ void __launch_bounds__(kBlock) k_count(DevParams P, DevBuffers B) {
  count_body(P, B, blockIdx.x, gridDim.x);
}
__global__ void __launch_bounds__(kBlock) k_compact(DevParams P, DevBuffers B) {
  compact_body(P, B, blockIdx.x, gridDim.x);
}
__global__ void __launch_bounds__(kBlock) k_idscan(DevParams P, DevBuffers B, int wave, int nres) {
  idscan_body(P, B, wave, blockIdx.x, nres);
}
// The reference's layouts from the slots, for downloads: per interior cell,
// cellGrainID_ / cellOrientationID_ (-1 when liquid), and per solid cell's
// grain its envelope and (cell, orientation) by id.  Ghost cells keep the 0
// that k_init_grid wrote.  In this path a cell never melts again, so every
// solid cell holds the grain created there.
__global__ void k_materialize(DevParams P, DevBuffers B) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < P.NG; c += stride) {
    int i, j, k;
    cell_ijk(P, (int)c, i, j, k);
    if (i == 0 || j == 0 || k == 0 || i > P.nx || j > P.ny || k > P.nz) continue;
    if (B.bmap[brick_word(P, i, j, k)] & brick_bit(i, j, k)) {
      B.gid[c] = -1;
      B.oid[c] = -1;
      continue;
    }
    const long long q = claim_index(P, i, j, k);
    const SlotRec r = B.rec[rec_slot(q)];
    B.gid[c] = r.id;
    B.oid[c] = r.o;
    B.genv[r.id] = make_double4(r.x, r.y, r.z, B.sL[q]);
    B.gco[r.id] = make_int2((int)c, r.o);
  }
}

// ---------------------------------------------------------------------------
// batched kernels: block b serves domain map[b].x as its block map[b].y of
// map[b].z (the per-domain launch geometry of the single-domain kernels)

__global__ void __launch_bounds__(kBlock) kb_nucleate(const DevParams* Ps, const DevBuffers* Bs,
                                                      const int4* map) {
  const int4 m = map[blockIdx.x];
  nucleate_body(Ps[m.x], Bs[m.x], m.y, m.z);
}
__global__ void __launch_bounds__(kBlock) kb_claim(const DevParams* Ps, const DevBuffers* Bs,
                                                   const int4* map, int wave) {
  const int4 m = map[blockIdx.x];
  claim_body(Ps[m.x], Bs[m.x], wave, m.y, m.z);
}
__global__ void __launch_bounds__(kBlock) kb_resolve(const DevParams* Ps, const DevBuffers* Bs,
                                                     const int4* map, int wave) {
  const int4 m = map[blockIdx.x];
  resolve_body(Ps[m.x], Bs[m.x], wave, m.y, m.z);
}
__global__ void __launch_bounds__(kBlock) kb_emit(const DevParams* Ps, const DevBuffers* Bs,
                                                  const int4* map, int wave) {
  const int4 m = map[blockIdx.x];
  emit_body(Ps[m.x], Bs[m.x], wave, m.y, m.z);
}
__global__ void __launch_bounds__(kTailBlock) kb_tail(const DevParams* Ps, const DevBuffers* Bs,
                                                      const int4* map) {
  const int4 m = map[blockIdx.x];
  tail_body(Ps[m.x], Bs[m.x]);
}
__global__ void __launch_bounds__(kBlock) kb_compact(const DevParams* Ps, const DevBuffers* Bs,
                                                     const int4* map) {
  const int4 m = map[blockIdx.x];
  compact_body(Ps[m.x], Bs[m.x], m.y, m.z);
}
__global__ void __launch_bounds__(kBlock) kb_count(const DevParams* Ps, const DevBuffers* Bs,
                                                   const int4* map) {
  const int4 m = map[blockIdx.x];
  count_body(Ps[m.x], Bs[m.x], m.y, m.z);
}
__global__ void __launch_bounds__(kBlock) kb_idscan(const DevParams* Ps, const DevBuffers* Bs,
                                                    const int4* map, int wave) {
  const int4 m = map[blockIdx.x];
  idscan_body(Ps[m.x], Bs[m.x], wave, m.y, m.w);
}

// per-domain status after a batch of steps: (finished, error, steps low, steps high)
__global__ void kb_status(const DevParams* Ps, const DevBuffers* Bs, int n, int4* out) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= n) return;
  const DevCounters* c = Bs[d].ctr;
  const int fin = (c->done || !(c->t < Ps[d].final_time)) ? 1 : 0;
  const unsigned long long st = (unsigned long long)c->step;
  out[d] = make_int4(fin, c->err, (int)(st & 0x7fffffffull), (int)(st >> 31));
}

hipError_t launch_batch_status(const BatchLaunch& L, int n, int4* out, hipStream_t s) {
  hipLaunchKernelGGL(kb_status, dim3((n + 63) / 64), dim3(64), 0, s, L.P, L.B, n, out);
  return hipGetLastError();
}

hipError_t launch_batch_step(const BatchLaunch& L, hipStream_t s) {
  hipLaunchKernelGGL(kb_nucleate, dim3(L.n_nuc), dim3(kBlock), 0, s, L.P, L.B, L.map_nuc);
  hipLaunchKernelGGL(kb_count, dim3(L.n_cmp), dim3(kBlock), 0, s, L.P, L.B, L.map_cmp);
  hipLaunchKernelGGL(kb_compact, dim3(L.n_cmp), dim3(kBlock), 0, s, L.P, L.B, L.map_cmp);
  for (int wave = 0; wave < 2; wave++) {
    hipLaunchKernelGGL(kb_claim, dim3(L.n_list), dim3(kBlock), 0, s, L.P, L.B, L.map_list, wave);
    hipLaunchKernelGGL(kb_resolve, dim3(L.n_emit), dim3(kBlock), 0, s, L.P, L.B, L.map_emit, wave);
    hipLaunchKernelGGL(kb_idscan, dim3(L.n_ids), dim3(kBlock), 0, s, L.P, L.B, L.map_ids, wave);
    hipLaunchKernelGGL(kb_emit, dim3(L.n_emit), dim3(kBlock), 0, s, L.P, L.B, L.map_emit, wave);
  }
  hipLaunchKernelGGL(kb_tail, dim3(L.n_one), dim3(kTailBlock), 0, s, L.P, L.B, L.map_one);
  return hipGetLastError();
}

hipError_t launch_init_grid(const DevParams& P, const DevBuffers& B, hipStream_t s) {
  long long blocks = (P.NG + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_init_grid, dim3((unsigned)blocks), dim3(256), 0, s, P, B);
  return hipGetLastError();
}

hipError_t launch_nucleate(const DevParams& P, const DevBuffers& B, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_nucleate, dim3(blocks), dim3(kBlock), 0, s, P, B);
  return hipGetLastError();
}

hipError_t launch_claim(const DevParams& P, const DevBuffers& B, int wave, int blocks,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_claim, dim3(blocks), dim3(kBlock), 0, s, P, B, wave);
  return hipGetLastError();
}

hipError_t launch_resolve(const DevParams& P, const DevBuffers& B, int wave, int blocks,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_resolve, dim3(blocks), dim3(kBlock), 0, s, P, B, wave);
  return hipGetLastError();
}

hipError_t launch_emit(const DevParams& P, const DevBuffers& B, int wave, int blocks,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_emit, dim3(blocks), dim3(kBlock), 0, s, P, B, wave);
  return hipGetLastError();
}

hipError_t launch_count(const DevParams& P, const DevBuffers& B, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_count, dim3(blocks), dim3(kBlock), 0, s, P, B);
  return hipGetLastError();
}

hipError_t launch_compact(const DevParams& P, const DevBuffers& B, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_compact, dim3(blocks), dim3(kBlock), 0, s, P, B);
  return hipGetLastError();
}

hipError_t launch_idscan(const DevParams& P, const DevBuffers& B, int wave, int nres,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_idscan, dim3(P.ids_blocks), dim3(kBlock), 0, s, P, B, wave, nres);
  return hipGetLastError();
}

hipError_t launch_materialize(const DevParams& P, const DevBuffers& B, hipStream_t s) {
  long long blocks = ((long long)P.NG + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_materialize, dim3((unsigned)blocks), dim3(256), 0, s, P, B);
  return hipGetLastError();
}

hipError_t launch_tail(const DevParams& P, const DevBuffers& B, hipStream_t s) {
  hipLaunchKernelGGL(k_tail, dim3(1), dim3(kTailBlock), 0, s, P, B);
  return hipGetLastError();
}

hipError_t launch_envelope_selftest(int n, const double* env4, const double* l2g9,
                                    const double* cell3, double h, double trunc, double sqrt2f,
                                    int* inside, double* child4, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_envelope_selftest, dim3((n + 255) / 256), dim3(256), 0, s, n, env4, l2g9,
                     cell3, h, trunc, sqrt2f, inside, child4);
  return hipGetLastError();
}

}  // namespace amca
