// Device-side state of one CA domain and the kernels of the plain step.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace amca {

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
  int err;                // 1: orientation table exhausted
  int log_cap;
  unsigned tickets[8];    // block tickets of the look-back kernels
};

struct StepLogRow {
  double time_before, dt;
  int n_nucleated, n_active_before, n_captured, n_active_after;
  long long n_liquid;
  int n_wave1, max_wave;
  int n_sites;
};

// Device buffers of one domain (all owned by the handle).
//
// Cell-indexed (ghosted grid, NG cells): the two grain-ID fields and the claim
// word of the capture waves.  The liquid and claimed states are 1 bit per cell
// in 4x4x4-cell bricks: brick b holds two 64-bit words, [2b] liquid and [2b+1]
// claimed, bit ((k&3)*16 + (j&3)*4 + (i&3)); bricks run x fastest.  A cell's 26
// neighbours then lie in at most 8 bricks and usually in two or three cache
// lines, against nine rows of a cell-linear map.  Grain-indexed (grain id =
// creation order, capacity N): envelope, cell and orientation.  Lists hold
// grain ids, so list order is id order and grain records are read in
// ascending address order.
struct DevBuffers {
  int* gid;               // cellGrainID_ [NG]
  int* oid;               // cellOrientationID_ [NG]
  int* claim;             // smallest claiming grain id per cell, INT_MAX when free, in
                          // 4x4x4 bricks (claim_index) [64 * nbX*nbY*nbZ]
  unsigned long long* bmap;  // brick words: liquid, claimed [2 * nbX*nbY*nbZ]
  double4* genv;          // grain envelope: xc, yc, zc, L [N]
  int2* gco;              // grain -> (cell, orientation) [N]
  int* listA[2];          // active lists (grain ids), ping-pong by step parity [N]
  int* listW[2];          // capture-wave lists (grain ids) [N]
  unsigned long long* masks;  // per wave entry: liquid bits | inside bits << 32 [N]
  unsigned* res;          // per wave entry: captured bits | keep << 31 [N]
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
constexpr int kLbSites = 4;       // look-back launch sites per step

// Kernel launchers (plain_step.hip).  Each returns hipGetLastError().
hipError_t launch_init_grid(const DevParams& P, const DevBuffers& B, hipStream_t s);
hipError_t launch_nucleate(const DevParams& P, const DevBuffers& B, int blocks, hipStream_t s);
hipError_t launch_claim(const DevParams& P, const DevBuffers& B, int wave, int blocks,
                        hipStream_t s);
// k_resolve, k_scan and k_emit of one wave, launched in that order with the
// same `blocks` (<= kMaxBlocks chunks)
hipError_t launch_resolve(const DevParams& P, const DevBuffers& B, int wave, int blocks,
                          hipStream_t s);
hipError_t launch_scan(const DevParams& P, const DevBuffers& B, int wave, int blocks,
                       hipStream_t s);
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
  const int4* map_scan;  // one block per domain (scan and tail), z = emit block count
  int n_nuc, n_list, n_emit, n_one;
};
hipError_t launch_batch_step(const BatchLaunch& L, hipStream_t s);
// out[d] = (finished, error, step & 0x7fffffff, step >> 31) for each of n domains
hipError_t launch_batch_status(const BatchLaunch& L, int n, int4* out, hipStream_t s);
hipError_t launch_envelope_selftest(int n, const double* env4, const double* l2g9,
                                    const double* cell3, double h, double trunc, double sqrt2f,
                                    int* inside, double* child4, hipStream_t s);

}  // namespace amca
