// Device-side state of a coupled CA-FE (remelting) domain and its kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace amca {

struct RmParams {
  int nx, ny, nz;
  int nxG, nyG, nzG;
  int nxyG;
  int NG;
  long long N;
  double h, x0, y0, z0;
  double Tm, Ts, Tmin_nuc, a1, a2, a3;   // kinetics: v = a1 dT^b1 + a2 dT^b2 + a3 dT^b3
  double b1, b2, b3;
  double trunc, tsf, sqrt2f, cutoff, final_time, void_temp;
  double vts;  // growth velocity at Ts + 0.01 K (cells below Ts grow with it), set by rm_launch_vts
  int off[26];
  int fnx, fny, fnz, nodes, ncx, ncy;    // structured FE mesh
  int n_frames;
  int orient_cap;
  // temperature surrogate (sur = 1): moving-frame tables on a xi lattice of
  // spacing sur_dxi from sur_xi0; FE fallback within sur_band of Tm and in
  // elements whose nodes disagree with the surrogate by more than sur_resid
  int sur;
  double sur_xi0, sur_dxi, sur_x_start, sur_v, sur_band, sur_resid;
  int sur_nxi, sur_nxi_nodes;
  double fx0, fy0, fz0, fdx;             // uniform node lattice of the structured mesh
  int sur_hot[6];        // xi-index, j, k ranges of the cell table at or above Tm - sur_band
  // active box: each step runs over the box of cells that can change (see k_rm_field)
  int window;
  int void_box[6];       // cells of elements whose nodes are void in some frame and not in another
  double hot_margin;     // K below Tm at which a node counts as hot
};

// The nodal field of one step (thetaArray_), as a function of the resident
// frames: frame 0 itself at the start, else the time interpolation between
// frames `frame` - 1 and `frame` with weight `ratio`.
struct RmField {
  int frame, init;
  double ratio;
};

// inclusive cell-index box (ghost layer included); empty when lo > hi
struct RmBox {
  int lo[3], hi[3];
};

struct RmCounters {
  double t;          // CA time_
  double tcafe;      // CafeSolver time
  double dt;         // dT_ of the running step
  double max_dt;     // maxTimestep_ = dT_FE
  double time_fe;    // FiniteElementManager time_
  double in_t0, in_t1;  // inTimeSeries_[0], [1]
  double ratio;      // nodal time-interpolation ratio
  long long step;
  // ordered keys of the minimum flagged temperature and flagged-cell counts,
  // spread over 64 words by block so that blocks rarely meet at one address
  unsigned long long min_t_keys[64];
  int flagged_parts[64];
  int frame;         // frame held as theta_[0]; frame-1 is theta_[1]
  int skip, exhausted;
  int n_grains, n_orient, n_sites;
  int step_nuc, step_cap, step_fresh;
  int err;
  int log_cap;
  int n_fallback;      // cells of the running step that took the FE path
  unsigned tickets[4];
  // active box of step n: box[n & 1], from acc[n & 1] = {hot elements at step
  // n, liquid cells with material after the phase of step n-1, cells
  // captured in step n-1}, each widened by two cells
  RmBox box[2];
  RmBox acc[2][3];
  int force_full;      // the next step runs over the whole grid
  int sur_used[2];     // step n used the surrogate: sur_used[n & 1]
  RmField fs[3];       // the field of step n: fs[n % 3] (set by the end of step n - 1)
  // Orientation rows: a step can nucleate at most at the sites whose cell is
  // liquid when it starts (armed[step & 1], counted by k_rm_field); a step
  // that could need more rows than the host has pushed (rows_avail) is not
  // run: need_rows stops it and every later step until the host pushes rows.
  int rows_avail, need_rows;
  int armed[2];
  double time_fe_prev; // FE time of the last step
  // the last phase, for the completion pass: step, FE times and the dt of its rate
  long long last_step;
  double last_time_fe, last_time_fe_prev, last_dt;
};

struct RmLogRow {
  double time_before, dt;
  int n_nucleated, n_captured, n_fresh, n_flagged;
  int n_fallback, n_box;
};

struct RmBuffers {
  int* gid;               // cellGrainID_ [NG]
  int* oid;               // cellOrientationID_ [NG]
  double* T;              // cellTemperature_ [NG]
  unsigned char* cooling; // cooling_ [NG]
  unsigned char* melted;  // hasBeenMelted_ [NG]
  unsigned char* flag;    // cellLiquidNeigh [NG]
  int* tag;               // element tag [NG]
  double* par;            // parametric coordinates, 3 per cell [3*NG]
  int* claim;             // smallest claiming sweep key, INT_MAX when free [NG]
  unsigned* msk;          // per flagged cell: enclosed / captured neighbour slots [NG]
  double4* genv;          // grain envelope by grain id [N]
  int2* gco;              // grain -> (cell, orientation) [N]
  const double* frames;   // nodal frames [n_frames * nodes]
  const double* frame_time;  // [n_frames], relative to the first
  double* theta;          // thetaArray_ of the coming step [nodes] (written with the surrogate
                          // each step, else at the end of a run)
  const int* frame_hot;   // cell box of the hot nodes of each frame [6 * n_frames]
  int* site_cell;         // nucleation sites, ascending cell (never erased here)
  double* site_dt;
  double* l2g;            // orientation rows, 9 doubles each
  int* rowcnt;            // fresh captures per (i, j) row, x-major row order [nx*ny]
  int* rowoff;
  unsigned* lb_flag;
  unsigned long long* lb_agg;
  unsigned long long* lb_pref;
  RmLogRow* log;
  RmCounters* ctr;
  const float* sur_cell;       // [nzG][nyG][sur_nxi] surrogate temperature at cell (y, z)
  const float* sur_node;       // [fnz][fny][sur_nxi_nodes] at node (y, z)
  unsigned char* node_bad;     // [nodes]
  unsigned char* elem_fb;      // [2][elements], step parity
  double* sur_stats;           // [8] error statistics
  int* fb_slots;               // [256] per-step fallback cell counts by block slot
  int* blk_box;                // [6 * 16384] per-block boxes of k_rm_liquid
};

hipError_t rm_launch_init(const RmParams& P, const RmBuffers& B, hipStream_t s);
// writes the growth velocity at Ts + 0.01 K to *out (device memory)
hipError_t rm_launch_vts(const RmParams& P, double* out, hipStream_t s);
// One step: kernels 0..kRmKernels-1 in order on stream s.
constexpr int kRmKernels = 10;
extern const char* const kRmKernelNames[kRmKernels];
hipError_t rm_launch_kernel(int q, const RmParams& P, const RmBuffers& B, int site_blocks,
                            hipStream_t s);
hipError_t rm_launch_step(const RmParams& P, const RmBuffers& B, int site_blocks, hipStream_t s);
// surrogate error statistics into B.sur_stats (zeroed by the caller)
hipError_t rm_launch_sur_error(const RmParams& P, const RmBuffers& B, hipStream_t s);
// T and cooling of the last step for every cell outside its active box, and
// the coming step's nodal field into B.theta
hipError_t rm_launch_complete(const RmParams& P, const RmBuffers& B, hipStream_t s);
// out[i] = pow(x[i], y) as glibc computes it (x[i] > 0, y supported), n >= 1
hipError_t launch_pow_selftest(long long n, const double* x, double y, double* out,
                               hipStream_t s);
// P.sur_hot from the cell table (B.sur_cell), written to *out (6 ints, device)
hipError_t rm_launch_sur_hot(const RmParams& P, const RmBuffers& B, int* out, hipStream_t s);

}  // namespace amca
