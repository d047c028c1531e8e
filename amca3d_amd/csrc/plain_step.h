// Device-side state of one CA domain and the kernels of the plain step.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace amca {

// Division by a fixed 32-bit divisor d as a multiply and shifts (the
// round-up method of Granlund and Montgomery, exact for every 32-bit n):
// t = mulhi(m, n), n / d = (t + ((n - t) >> s1)) >> s2.
struct FastDiv {
  unsigned m;
  int s1, s2;
};

// Constant description of a domain, passed by value to every kernel.
struct DevParams {
  int nx, ny, nz;
  int nxG, nyG, nzG;
  int nxyG;
  int NG;                 // ghosted cell count
  long long N;            // interior cell count
  double h, x0, y0, z0;
  double T0, Tm, Tdot, a1, a2, a3;
  double trunc, tsf, max_dt, final_time;
  double sqrt2f;          // (double)sqrtf(2), Octahedron.cpp:260
  int off[26];            // 26-neighbour offsets in the reference's k,j,i order
  int orient_cap;
  int nbX, nbY, nbZ;      // 4x4x4-cell bricks per axis of the liquid / claimed map
  FastDiv fdX, fdY;       // division by nbX and by nbY
  int ids_blocks;         // blocks of the id scan (k_idscan)
};

// Mutable scalars that live on the device, so that a whole step (or many
// steps) runs without a host round trip.
struct DevCounters {
  double t;               // time_
  double dt;              // dT_ of the current step
  double v;               // growth velocity of the current step
  long long step;         // iStep_
  long long n_solid;      // solid interior cells
  int skip;               // 1: this step is past termination, every kernel returns
  int done;               // no liquid cell left
  int parity;             // which active-list buffer is the input of this step
  int site_sel;           // which site buffer holds the live sites
  int n_grains, n_orient;
  int nA_in, nA_out;      // active list lengths (input / survivors)
  int nW[2];              // capture-wave list lengths
  int n_sites;
  int step_nuc, step_cap, step_nA0;
  int step_sites;         // live sites at the running step's start
  int max_wave;
  int step_wave;          // deepest capturing wave of the running step
  int wave_gbase;         // first new grain id of the running wave
  int w1_lo, w1_n;        // ids of wave 1's entries (wave 0's children)
  int err;                // 1: orientation table exhausted
  int log_cap;
  unsigned tickets[8];    // block tickets of the look-back kernels
  // ids of wave 0's claiming grains (min, max), spread over 64 slots by k_claim
  // block so that blocks rarely meet at one address
  int claim_lo[64], claim_hi[64];
};

struct StepLogRow {
  double time_before, dt;
  int n_nucleated, n_active_before, n_captured, n_active_after;
  long long n_liquid;
  int n_wave1, max_wave;
  int n_sites;
};

// The grain in a cell's slot: its centre and (id, orientation).  The
// half-diagonal L sits in an array of its own, since every step rewrites it
// for each active grain while the rest stays as written at capture.
struct alignas(32) SlotRec {
  double x, y, z;
  int id, o;
};

// Device buffers of one domain (all owned by the handle).
//
// The grid is cut into 4x4x4-cell bricks (x fastest); cell (i, j, k) is lane
// ((k&3)*16 + (j&3)*4 + (i&3)) of brick b, and its "slot" is 64 b + lane
// (claim_index).  Per brick: two 64-bit words of the 1-bit maps, [2b] liquid
// and [2b+1] claimed, and the active word (a grain whose cell has the bit set
// is on the reference's active list).  Per slot: the grain occupying the cell
// (record, L, id) and the claim word of the capture waves.  Steps read and
// write only these: wave 0 walks the active slots in brick order, so
// neighbouring entries share map words and record lines, and the lists of
// waves >= 1 hold slots in grain-id order.  The reference's layouts
// (cellGrainID_ / cellOrientationID_ on the ghosted grid, x fastest; records
// by grain id) are materialized from the slots when a download asks for them.
struct DevBuffers {
  int* gid;               // cellGrainID_ [NG], materialized for downloads
  int* oid;               // cellOrientationID_ [NG], materialized for downloads
  int* claim;             // smallest claiming grain id per slot, INT_MAX when free [64 nB]
  unsigned long long* bmap;  // brick words: liquid, claimed [2 nB]
  unsigned long long* amap;  // brick words: active grains [nB]
  SlotRec* rec;           // centre, id and orientation of the grain in each slot [64 nB]
  double* sL;             // half-diagonal L of the grain in each slot [64 nB]
  int* skey;              // grain id of each slot (the id of rec, packed alone) [64 nB]
  double4* genv;          // envelope by grain id, materialized for downloads [N]
  int2* gco;              // grain -> (cell, orientation), materialized for downloads [N]
  int* alist;             // wave-0 list: active slots in brick order [N]
  int4* plist;            // wave-0 capturing grains: (id, slot, captured slots), one segment
                          // per k_resolve block from its first list entry [N]
  unsigned char* cnt;     // captures per grain id, nonzero only inside a wave-0 step [N]
  int* coff;              // first child (wave-relative) per capturing grain id [N]
  unsigned* idtot;        // wave-0 captures per id-scan block [kMaxIdsBlocks]
  unsigned* cmptot;       // active cells per compaction block [kMaxBlocks]
  int* listW[2];          // capture-wave lists (slots, in grain-id order) [N]
  unsigned long long* masks;  // per wave entry: liquid bits | inside bits << 32 [N]
  int* site_cell[2];      // live nucleation sites, ascending cell, ping-pong
  double* site_dt[2];
  double* l2g;            // orientation table, 9 doubles per row
  unsigned* lb_flag;      // look-back state: epoch<<2 | state   [kLbSites * kMaxBlocks]
  unsigned long long* lb_agg;
  unsigned long long* lb_pref;
  StepLogRow* log;
  DevCounters* ctr;
};

constexpr int kBlock = 256;       // threads per block of the list kernels
constexpr int kItems = 4;         // consecutive list entries per thread in scan passes
constexpr int kMaxBlocks = 4096;  // blocks of a look-back kernel
constexpr int kLbSites = 2;       // look-back arrays: [0] k_nucleate's look-back, [1] k_resolve's
                                  // per-block statistics (lb_agg, lb_pref at kStatBase + block)
constexpr int kStatBase = kMaxBlocks;
constexpr int kMaxIdsBlocks = 2048;  // blocks of the id scan (a k_resolve block keeps one count each)

inline FastDiv make_fastdiv(unsigned d) {
  int l = 0;
  while ((1ull << l) < d) l++;  // ceil(log2 d)
  FastDiv f;
  f.m = (unsigned)(((1ull << 32) * ((1ull << l) - d)) / d + 1);
  f.s1 = l < 1 ? l : 1;
  f.s2 = l > 1 ? l - 1 : 0;
  return f;
}

// Kernel launchers (plain_step.hip).  Each returns hipGetLastError().
hipError_t launch_init_grid(const DevParams& P, const DevBuffers& B, hipStream_t s);
hipError_t launch_nucleate(const DevParams& P, const DevBuffers& B, int blocks, hipStream_t s);
hipError_t launch_claim(const DevParams& P, const DevBuffers& B, int wave, int blocks,
                        hipStream_t s);
// Waves 0 and 1: k_claim, k_resolve, k_idscan and k_emit in that order;
// k_resolve and k_emit with the same `blocks` (<= kMaxBlocks list chunks)
hipError_t launch_resolve(const DevParams& P, const DevBuffers& B, int wave, int blocks,
                          hipStream_t s);
// wave 0's list: the active slots in brick order (k_count: active cells per
// block of bricks, then k_compact over the same blocks)
hipError_t launch_count(const DevParams& P, const DevBuffers& B, int blocks, hipStream_t s);
hipError_t launch_compact(const DevParams& P, const DevBuffers& B, int blocks, hipStream_t s);
// child offsets of a wave's capturing grains, an exclusive scan over their
// ids (P.ids_blocks blocks; `nres` = blocks of the wave's k_resolve, whose
// statistics words it reads)
hipError_t launch_idscan(const DevParams& P, const DevBuffers& B, int wave, int nres,
                         hipStream_t s);
// gid / oid on the ghosted grid, genv / gco by grain id, from the slots
hipError_t launch_materialize(const DevParams& P, const DevBuffers& B, hipStream_t s);
hipError_t launch_emit(const DevParams& P, const DevBuffers& B, int wave, int blocks,
                       hipStream_t s);
hipError_t launch_tail(const DevParams& P, const DevBuffers& B, hipStream_t s);

// One step of many domains in one launch per kernel (device arrays of the
// domains' parameters and buffers; per-block (domain, block, count) maps).
struct BatchLaunch {
  const DevParams* P;
  const DevBuffers* B;
  const int4* map_nuc;   // nucleation geometry per domain
  const int4* map_list;  // claim geometry
  const int4* map_emit;  // resolve / emit geometry
  const int4* map_one;   // one block per domain (tail)
  const int4* map_ids;   // id-scan geometry
  const int4* map_cmp;   // compaction geometry (k_count and k_compact)
  int n_nuc, n_list, n_emit, n_one, n_ids, n_cmp;
};
hipError_t launch_batch_step(const BatchLaunch& L, hipStream_t s);
// out[d] = (finished, error, step & 0x7fffffff, step >> 31) for each of n domains
hipError_t launch_batch_status(const BatchLaunch& L, int n, int4* out, hipStream_t s);
hipError_t launch_envelope_selftest(int n, const double* env4, const double* l2g9,
                                    const double* cell3, double h, double trunc, double sqrt2f,
                                    int* inside, double* child4, hipStream_t s);

}  // namespace amca
