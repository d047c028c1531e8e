/*
 * amca3d.h -- C ABI of the MI355X (gfx950) cellular-automaton library.
 *
 * The reference (AMCA3D) has no FFI: its hot path sits behind the C++ virtual
 * interface `CellularAutomataManager` (CA/include/CellularAutomataManager.h:20-246)
 * and is driven by `CellularAutomataSolver` / `CafeSolver`.  The entry points
 * below are what a C++ subclass of that interface (or any other host) binds to
 * replace the per-cell loops; each cites the reference member it replaces.
 *
 * Conventions
 *   - Every call returns an amca_status; no C++ exception crosses the ABI.
 *     amca_last_error() gives a text for the last failure on this thread.
 *   - Device memory is owned by the handle.  Host pointers passed in are only
 *     borrowed for the duration of the call.
 *   - One host thread per handle; work is ordered on the handle's own HIP
 *     stream, handles are independent.
 *   - Cell indices are "local-ghost" indices of the reference:
 *     c = k*(nx+2)*(ny+2) + j*(nx+2) + i with i,j,k in [0, n+1], ghosts on the
 *     faces (CellularAutomataManager_3D.cpp:876-888 at one rank).
 */
#ifndef AMCA3D_H
#define AMCA3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  AMCA_OK = 0,
  AMCA_ERR_INVALID = 1,     /* bad argument or state */
  AMCA_ERR_HIP = 2,         /* a HIP runtime call failed */
  AMCA_ERR_CAPACITY = 3,    /* a fixed-capacity table ran out */
  AMCA_ERR_NO_DEVICE = 4,   /* no GPU visible */
  AMCA_ERR_IO = 5,          /* file or parse error (host layer) */
} amca_status;

typedef struct amca_domain amca_domain;

/* Domain description: geometry, kinetics and stepping constants of one realm.
 * Fields mirror the reference members they come from. */
typedef struct {
  int32_t nx, ny, nz;         /* interior cells (nx_, ny_, nz_) */
  double h;                   /* cell size (h_) */
  double x0, y0, z0;          /* original point (x0_, y0_, z0_) */
  double T0, Tm, Tdot;        /* initial, melting temperature, cooling rate */
  double a1, a2, a3;          /* RappazGandin1993 kinetics (ProblemPhysics.cpp:117-135) */
  double trunc;               /* envelope truncation coefficient (Octahedron.cpp:25,319) */
  double time_step_factor;    /* timeStepFactor_ (CellularAutomataManager.cpp:44) */
  double max_dt;              /* maxTimestep_ (time_step key) */
  double final_time;          /* termination_time */
} amca_domain_desc;

/* Counters after a step or a run (replace the per-step log lines of
 * CellularAutomataManager_3D.cpp:658-667). */
typedef struct {
  int64_t step;               /* iStep_ */
  double time;                /* time_ */
  double dt;                  /* dT_ of the last step taken */
  int64_t n_liquid;           /* get_amount_of_liquid_cells (:2039-2061) */
  int32_t n_grains;           /* grainVector_.size() */
  int32_t n_orient;           /* orientationVector_.size() */
  int32_t n_active;           /* activeGrainList_.size() */
  int32_t n_sites;            /* live nucleation sites */
  int32_t done;               /* solidificationDone_ or time >= termination_time */
  int32_t max_wave;           /* deepest same-step capture cascade seen */
} amca_counters;

/* One row of the optional per-step log. */
typedef struct {
  double time_before, dt;
  int32_t n_nucleated, n_active_before, n_captured, n_active_after;
  int64_t n_liquid;
  int32_t n_wave1;    /* grains captured by the grains active at the step's start */
  int32_t max_wave;   /* deepest capture wave that captured a cell (0: none) */
  int32_t n_sites;    /* live nucleation sites at the step's start */
} amca_step_record;

/* ---- domain lifetime ---------------------------------------------------- */

/* Allocates the ghosted grid on the current device and fills it like
 * CellularAutomataManager_3D::setup_cells + initialize_boundaries
 * (CellularAutomataManager_3D.cpp:404-435, 735-833; CellularAutomataManager.cpp:303-326):
 * every cell liquid (-1), ghost faces grain/orientation 0.
 * `orient_capacity` bounds the orientation table (nucleation events). */
amca_status amca_domain_create(const amca_domain_desc* desc, int32_t orient_capacity,
                               amca_domain** out);

/* Frees the device memory (CellularAutomataManager.cpp:73-102). */
amca_status amca_domain_destroy(amca_domain* d);

/* ---- inputs drawn on the host -------------------------------------------- */

/* Nucleation sites, sorted ascending by cell, with their critical
 * undercooling: the output of create_nucleation_sites
 * (CellularAutomataManager.cpp:153-213). */
amca_status amca_domain_set_sites(amca_domain* d, int32_t n, const int32_t* cell,
                                  const double* dT_crit);

/* Orientation rows in nucleation-event order: Euler angles and the row-major
 * local-to-global rotation of Octahedron::set_coordinate_transformation_matrix
 * (Octahedron.cpp:281-310), drawn as in get_random_orientation_angle
 * (CellularAutomataManager_3D.cpp:838-850).  Appends after rows already set. */
amca_status amca_domain_push_orientations(amca_domain* d, int32_t n, const double* euler3,
                                          const double* l2g9);

/* ---- stepping ------------------------------------------------------------ */

/* One calculate_time_step + single_integration_step
 * (CellularAutomataManager_3D.cpp:440-669): nucleation, dt, growth, capture
 * cascade, time advance, liquid count.  A step after termination is a no-op. */
amca_status amca_domain_step(amca_domain* d, amca_counters* out);

/* CellularAutomataManager::time_integration (CellularAutomataManager.cpp:134-147):
 * steps until time >= termination_time or no liquid cell remains, or until
 * `max_steps` steps were taken in this call (max_steps < 0: no limit). */
amca_status amca_domain_run(amca_domain* d, int64_t max_steps, amca_counters* out);

/* Same as amca_domain_run but takes exactly `n_steps` step launches without
 * reading anything back in between (steps after termination are no-ops).
 * Used for timing. */
amca_status amca_domain_run_fixed(amca_domain* d, int64_t n_steps);

amca_status amca_domain_counters(amca_domain* d, amca_counters* out);

/* Per-step log: enable before stepping (capacity rows, 0 disables), then read. */
amca_status amca_domain_enable_log(amca_domain* d, int32_t capacity);
amca_status amca_domain_read_log(amca_domain* d, int32_t cap, amca_step_record* rows,
                                 int32_t* n_rows);

/* ---- outputs ------------------------------------------------------------- */

/* Copies the state to host buffers (any pointer may be NULL), the data of
 * output_microstructure_information (CellularAutomataManager_3D.cpp:1835-1913):
 *   orient[ng], grain[ng]   cellOrientationID_, cellGrainID_  (ng = (nx+2)(ny+2)(nz+2))
 *   env[4*ng]               per-cell grain envelope xc, yc, zc, L (FP64)
 *   site_cell[n], site_dt[n] live nucleation sites (n <= site_cap) */
amca_status amca_domain_download(amca_domain* d, int32_t* orient, int32_t* grain, double* env,
                                 int32_t site_cap, int32_t* site_cell, double* site_dt,
                                 int32_t* n_sites);

/* Grain records in grain-id order (grainVector_ of the reference): envelope
 * xc, yc, zc, L, owning cell and orientation; copies min(cap, n) rows. */
amca_status amca_domain_download_grains(amca_domain* d, int32_t cap, double* env4, int32_t* cell,
                                        int32_t* orient, int32_t* n_grains);

/* Device pointers of the grain-ID fields (int32 per ghosted cell) and of the
 * grain envelopes (double4 per grain id), and the stream, for callers that
 * keep the data on the GPU.  The contents are as of this call: the step keeps
 * grains by cell in its own layout and this call writes these arrays from it
 * (later steps do not update them; call again after stepping).  Valid until
 * amca_domain_destroy. */
amca_status amca_domain_device_fields(amca_domain* d, void** orient, void** grain, void** env,
                                      void** stream);

/* ---- kernel timing (HIP events on the handle's stream) -------------------- */

/* Enables accumulation of per-kernel durations (milliseconds) for the named
 * kernels of the step; amca_domain_kernel_times returns totals and launch
 * counts in a fixed order given by amca_kernel_name(i). */
amca_status amca_domain_enable_timing(amca_domain* d, int32_t on);
int32_t amca_kernel_count(void);
const char* amca_kernel_name(int32_t i);
amca_status amca_domain_kernel_times(amca_domain* d, double* ms, int64_t* launches);

/* ---- batches of plain domains ------------------------------------------------
 * Many independent domains (own sizes, seeds, cooling rates, site densities)
 * stepped together: one launch per kernel for the whole batch, captured in a
 * HIP graph of `steps_per_graph` steps and replayed until every domain has
 * finished.  No counterpart in the reference, which runs one realm per
 * process; each domain's results equal its own amca_domain_run.  Domains stay
 * owned by the caller (create, sites and orientation rows as usual, then
 * batch them; read them back with amca_domain_download afterwards). */
typedef struct amca_batch amca_batch;

typedef struct {
  int32_t list_blocks;      /* claim blocks, shared in proportion to domain cell counts */
  int32_t emit_blocks;      /* resolve / emit blocks, same sharing */
  int32_t min_blocks;       /* at least this many per domain (1..4096) */
  int32_t interleave;       /* 0: blocks domain by domain (largest first), 1: round robin */
  int32_t steps_per_graph;  /* steps captured in one graph replay */
} amca_batch_schedule;

/* AMCA_BATCH_GRAPHS=0 in the environment: replays launch their kernels one by
 * one instead of from a captured graph (any other value than 0 or 1 fails). */
amca_status amca_batch_create(amca_domain* const* domains, int32_t n, amca_batch** out);
amca_status amca_batch_destroy(amca_batch* b);
amca_status amca_batch_set_schedule(amca_batch* b, const amca_batch_schedule* s);
amca_status amca_batch_get_schedule(amca_batch* b, amca_batch_schedule* s);
/* Replays until all domains finished (max_graphs < 0) or max_graphs replays;
 * returns the finished-domain count, device milliseconds and replays.  Each
 * time max(4, n/8) more domains have finished, the block budgets are shared
 * again over the running domains and the graph is captured anew (results do
 * not change; finished domains get no blocks). */
amca_status amca_batch_run(amca_batch* b, int64_t max_graphs, int32_t* n_finished, double* ms,
                           int64_t* replays);
/* Tries candidate schedules on the running batch (trial_graphs replays each of
 * a graph of min(4, steps_per_graph) steps, `rounds` rounds, odd rounds in
 * reverse candidate order) and keeps the fastest per step: first one block
 * budget for every list kernel (4, 16 or 64 blocks per CU) in both block
 * orders, then the claim grid and the resolve / emit grid sized apart around
 * that choice.  The trial steps are ordinary steps of the run; the schedule's
 * steps_per_graph is kept. */
amca_status amca_batch_autotune(amca_batch* b, int32_t trial_graphs, int32_t rounds,
                                amca_batch_schedule* chosen);

/* ---- temperature surrogate (MLP) ----------------------------------------------
 * A 5-128-128-128-128-1 network with SiLU on the hidden layers: inputs
 * (x - v t, y, z, P, v) normalised by the caller, output the normalised
 * temperature.  BF16 operands on the matrix cores, FP32 accumulation, FP32
 * master weights trained with Adam on the mean squared error.  No counterpart
 * in the reference, whose temperatures come only from FE frames.  The
 * parameter vector is W1[5][128] b1[128] W2..W4[128][128] b2..b4[128]
 * W5[128] b5, weights stored [fan_in][fan_out]. */
typedef struct amca_mlp amca_mlp;
int32_t amca_mlp_param_count(void);
amca_status amca_mlp_create(uint64_t seed, amca_mlp** out);
amca_status amca_mlp_destroy(amca_mlp* m);
amca_status amca_mlp_set_params(amca_mlp* m, const float* params);
amca_status amca_mlp_get_params(amca_mlp* m, float* params);
/* Host rows x5[n][5] -> t[n]. */
amca_status amca_mlp_forward(amca_mlp* m, int32_t n, const float* x5, float* t);
/* Gradient of mean((mlp(x) - y)^2) over n host rows; loss = that mean. */
amca_status amca_mlp_gradients(amca_mlp* m, int32_t n, const float* x5, const float* y,
                               float* grads, float* loss);
/* `steps` Adam steps over random contiguous batches of device-resident samples
 * (dX[n][5], dY[n], pre-shuffled); optional per-step loss trace (host). */
amca_status amca_mlp_train_device(amca_mlp* m, int64_t n, const float* dX, const float* dY,
                                  int32_t batch, int32_t steps, float lr, uint64_t seed,
                                  float* loss_trace);
/* The same forward pass as one fused kernel (weights in LDS, activations in
 * registers, the whole network per 64 rows in one wave); host pointers, n rows.
 * Equal to amca_mlp_forward up to FP32 accumulation order and the BF16
 * rounding of the hidden activations, which this kernel rounds scaled by
 * -log2(e) (the scaling cancels in the next layer). */
amca_status amca_mlp_forward_fused(amca_mlp* h, int32_t n, const float* x5, float* t);
/* Device rows dX[n][5] -> dT[n]. */
amca_status amca_mlp_forward_device(amca_mlp* m, int64_t n, const float* dX, float* dT);
/* amca_mlp_forward_device without any allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The same kernels and
 * results; the activations go through a file-scope device array of 16 MiB per
 * GPU (held until the process ends, shared by every call on that GPU, so
 * calls that overlap on one GPU race).  n <= AMCA_MLP_NOALLOC_MAX_ROWS per call
 * (AMCA_ERR_INVALID beyond it); the network's device must be current. */
#define AMCA_MLP_NOALLOC_MAX_ROWS 32768
amca_status amca_mlp_forward_device_noalloc(amca_mlp* m, int64_t n, const float* dX, float* dT);
/* Device time of the fused kernel's launches (amca_mlp_forward_fused and the
 * lattice evaluations of amca_remelt_use_surrogate), from HIP events on the
 * network's stream: off by default; enabling resets the totals (milliseconds,
 * launches, rows evaluated). */
amca_status amca_mlp_enable_timing(amca_mlp* m, int32_t on);
amca_status amca_mlp_fused_times(amca_mlp* m, double* ms, int64_t* launches, int64_t* rows);
/* The temperature field the surrogate learns: a moving point source on a
 * structured FE mesh (uniform node spacing), Rosenthal nodal values at frames
 * frame_dt apart, interpolated linearly in time and by hex8 in space (the FE
 * path of the coupled domain).  Inputs are normalised as
 * ((x - x_start - v t)/xi_scale, (y - y_c)/y_scale, (z - z_top)/z_scale,
 * P/P_scale, v/v_scale), the target as (min(T, T_clip) - T_ref)/T_scale, or
 * its log1p (log_target), which weighs the far field's small rises evenly. */
typedef struct {
  int32_t fnx, fny, fnz;
  double fx0, fy0, fz0, fdx;
  int32_t n_frames;
  double frame_dt;
  double eta, k, alpha, T0, z_top, r_min, void_temperature;
  double xi_scale, y_scale, z_scale, P_scale, v_scale;
  double T_ref, T_scale, T_clip;
  int32_t log_target;  /* 1: target log1p(max(min(T, T_clip) - T_ref, 0) / T_scale) */
} amca_surrogate_desc;

typedef struct {
  double P, v, x_start, y_c;    /* power, speed, source x at t = 0, track line y */
} amca_scan;

/* Draws n training samples of the FE field over the scans (a fraction
 * `near_fraction` close behind the source, the rest uniform in the mesh) into
 * the network's device-resident training set. */
amca_status amca_surrogate_sample(amca_mlp* m, const amca_surrogate_desc* d, const amca_scan* scans,
                                  int32_t n_scans, int64_t n, double near_fraction, uint64_t seed);
/* Adam steps over the training set (amca_mlp_train_device on it). */
amca_status amca_surrogate_train(amca_mlp* m, int32_t batch, int32_t steps, float lr, uint64_t seed,
                                 float* loss_trace);
/* The fused kernel over a lattice of the moving frame, as amca_remelt_use_surrogate
 * builds its tables: out[(k*ny + j)*nxi + i] = surrogate temperature (kelvin) at
 * (x_start + xi0 + i*dxi, y0 + j*dy, z0 + k*dz, t = 0) of scan s.  Host output. */
amca_status amca_surrogate_eval_lattice(amca_mlp* m, const amca_surrogate_desc* d,
                                        const amca_scan* s, int32_t nxi, double xi0, double dxi,
                                        int32_t ny, double y0, double dy, int32_t nz, double z0,
                                        double dz, float* out);
/* Surrogate and FE-path temperatures (kelvin) at n host points xyzt[n][4] of one scan. */
amca_status amca_surrogate_eval(amca_mlp* m, const amca_surrogate_desc* d, const amca_scan* s,
                                int32_t n, const double* xyzt, double* T_surrogate, double* T_fe);

/* The tiled BF16 GEMM alone: C[M][N] = opA(A) opB(B) with FP32 host matrices
 * (row-major; trans_a: A is K x M, trans_b: B is N x K). */
amca_status amca_mlp_gemm_selftest(int32_t M, int32_t N, int32_t K, int32_t trans_a,
                                   int32_t trans_b, const float* A, const float* B, float* C);
const char* amca_mlp_last_error(void);

/* ---- device arithmetic self-check ---------------------------------------- */

/* Runs the device capture test (Octahedron::cell_is_captured, Octahedron.cpp:98-128)
 * and child-envelope construction (Octahedron::capture_cell, Octahedron.cpp:133-275)
 * on n inputs: env4[4n] parent centre+L, l2g9[9n], cell3[3n]; writes
 * inside[n] and child4[4n] (child centre and L, only where inside). */
amca_status amca_envelope_selftest(int32_t n, const double* env4, const double* l2g9,
                                   const double* cell3, double h, double trunc,
                                   int32_t* inside, double* child4);

/* Device stream-copy rate: `reps` copies of a `bytes` buffer by the kernel
 * k_copy_probe (read + write counted), average GB/s and ms per copy.  No
 * counterpart in the reference: it is the measured bandwidth roof and the
 * known-traffic kernel that hardware byte counters are scaled against. */
amca_status amca_copy_probe(int64_t bytes, int32_t reps, double* gbps, double* ms);

const char* amca_last_error(void);
const char* amca_version(void);

/* ---- coupled CA-FE (remelting) domain --------------------------------------
 * The `cellular_automata_remelting` realm driven by CafeSolver
 * (CafeSolver.cpp:159-219, CellularAutomataManager_3DRemelting.cpp) with the
 * FE temperature provider (LoadDataFromFile.cpp, MapVoxelManager.cpp) on the
 * device. */
typedef struct amca_remelt amca_remelt;

typedef struct {
  int32_t nx, ny, nz;
  double h, x0, y0, z0;
  double Tm, Ts, Tmin_nuc;        /* melting, solidus, minimum_temperature_for_nucleation */
  double a1, a2, a3, b1, b2, b3;  /* ProblemPhysics.cpp:142-158; each b 0 or |b| in [2^-65, 2^63) */
  double trunc;                   /* envelope truncation coefficient */
  double time_step_factor;        /* time_step_factor_ca */
  double cutoff_factor;           /* cell_grain_length_cutoff_factor */
  double final_time;              /* termination_time */
  double void_temperature;        /* transfers.void_temperature */
  int32_t bin[3];                 /* transfers.bin_size */
  double tolerance;               /* transfers.tolerance */
  int32_t max_iterations;         /* transfers.max_iterations */
  int32_t initial_melted;         /* phase_state.has_been_melted */
} amca_remelt_desc;

typedef struct {
  int64_t step;
  double time, time_fe, dt, max_dt;
  int32_t n_grains, n_orient, done, frames_exhausted;
} amca_remelt_counters;

typedef struct {
  double time_before, dt;
  int32_t n_nucleated, n_captured, n_fresh, n_flagged;
  int32_t n_fallback;   /* with a surrogate: cells that took the FE path */
  int32_t n_box;        /* cells of the step's active box (ghost layer included) */
} amca_remelt_step_record;

/* Supplies n more orientation rows (Euler angles and row-major local-to-global
 * rotation) continuing the host's rand() stream; returns 0 on success. */
typedef int (*amca_orientation_source)(void* user, int32_t n, double* euler3, double* l2g9);

/* CellularAutomataManager_3DRemelting::initialize (:81-135) storage.
 * AMCA_REMELT_GRAPH_STEPS in the environment (0..1024) sets the graph length,
 * as amca_remelt_set_graph_steps does. */
amca_status amca_remelt_create(const amca_remelt_desc* desc, int32_t orient_capacity,
                               amca_remelt** out);
amca_status amca_remelt_destroy(amca_remelt* d);
/* A structured hex FE mesh and its nodal frames (the data of
 * load_mesh_and_first_two_steps_results_from_txt_file / load_next_step_nodal_data,
 * LoadDataFromFile.cpp:21-191, after offsets and scales); maps every cell to its
 * element and parametric coordinates (MapVoxelManager.cpp:140-699) on the host. */
amca_status amca_remelt_set_fe(amca_remelt* d, int32_t fnx, int32_t fny, int32_t fnz,
                               const double* node_xyz, int32_t n_frames, const double* frame_time,
                               const double* frame_theta);
amca_status amca_remelt_voxel_map(amca_remelt* d, int32_t* tag, double* par);
/* The same voxel mapping without a domain or a GPU: tag[NG], par[3*NG]
 * (MapVoxelManager::execute_with_focus_on_CA_region, MapVoxelManager.cpp:140-482). */
amca_status amca_fe_map_voxels(int32_t nx, int32_t ny, int32_t nz, double h, double x0, double y0,
                               double z0, const int32_t* bin, double tolerance,
                               int32_t max_iterations, int32_t fnx, int32_t fny, int32_t fnz,
                               const double* node_xyz, int32_t* tag, double* par);
/* The same mapping with the per-cell search on the GPU (one thread per cell,
 * the host mapper's arithmetic): tags and parametric coordinates equal
 * amca_fe_map_voxels bit for bit.  amca_remelt_set_fe maps this way. */
amca_status amca_fe_map_voxels_device(int32_t nx, int32_t ny, int32_t nz, double h, double x0,
                                      double y0, double z0, const int32_t* bin, double tolerance,
                                      int32_t max_iterations, int32_t fnx, int32_t fny,
                                      int32_t fnz, const double* node_xyz, int32_t* tag,
                                      double* par);
/* load_microstructure_information + retrieve_grain_information_before_simulation
 * (CellularAutomataManager_3DRemelting.cpp:165-332): per-cell orientation, grain
 * id and melted flag; grain records (cell, orientation), centres reset to the
 * cell centre with zero length.  Orientation rows 0..n_orient-1 must be pushed first. */
amca_status amca_remelt_set_state(amca_remelt* d, const int32_t* orient, const int32_t* grain,
                                  const uint8_t* melted, int32_t n_orient, int32_t n_grains,
                                  const int32_t* grain_cell, const int32_t* grain_orient);
/* Appends n orientation rows (Euler angles; l2g9 may be NULL, then the
 * rotations are computed on the host as Octahedron.cpp:281-310 does). */
amca_status amca_remelt_push_orientations(amca_remelt* d, int32_t n, const double* euler3,
                                          const double* l2g9);
amca_status amca_remelt_set_orientation_source(amca_remelt* d, amca_orientation_source src,
                                               void* user);
/* Nucleation sites (criterion_to_activate_grain.using_nucleation_seed). */
amca_status amca_remelt_set_sites(amca_remelt* d, int32_t n, const int32_t* cell,
                                  const double* dT_crit);
/* The host half of CellularAutomataManager_3DRemelting::initialize: srand(seed),
 * then create_nucleation_sites for the surface (six faces, h^2 per cell) and the
 * bulk (h^3 per cell) populations (CellularAutomataManager.cpp:153-213), and
 * sets the orientation source to the continuing rand() stream
 * (get_random_orientation_angle, CellularAutomataManager_3D.cpp:838-850).
 * The draws merge into the sites already set (amca_remelt_set_sites: the seeds
 * of a loaded microstructure file, CellularAutomataManager_3DRemelting.cpp:244-249):
 * a draw on a listed cell only lowers its undercooling.  Replaces the source. */
amca_status amca_remelt_draw_sites(amca_remelt* d, uint32_t seed, double ns, double dTs_mean,
                                   double dTs_sd, double nv, double dTv_mean, double dTv_sd,
                                   int32_t* n_sites);
/* Same, after discarding `skip_draws` rand() values (the draws a random initial
 * microstructure takes first); returns the new-site counts of each population
 * (numSitesSurface_ / numSitesBulk_). */
amca_status amca_remelt_draw_sites_after(amca_remelt* d, uint32_t seed, int64_t skip_draws,
                                         double ns, double dTs_mean, double dTs_sd, double nv,
                                         double dTv_mean, double dTv_sd, int32_t* n_surface_new,
                                         int32_t* n_bulk_new);
/* The nucleation sites as set (cell, critical undercooling); copies min(cap, n). */
amca_status amca_remelt_sites(amca_remelt* d, int32_t cap, int32_t* cell, double* dT_crit,
                              int32_t* n);
/* Euler angles of every orientation row pushed or drawn so far. */
amca_status amca_remelt_euler(amca_remelt* d, int32_t cap, double* euler3, int32_t* n);
/* CafeSolver::time_integration_weak_coupling (CafeSolver.cpp:178-219) until
 * termination_time, the end of the frames, or max_steps steps (< 0: no limit). */
amca_status amca_remelt_run(amca_remelt* d, int64_t max_steps, amca_remelt_counters* out);
/* Batched form: runs n independent domains (scan variants, seeds) at once,
 * each as amca_remelt_run would, queuing every domain's next steps on its own
 * stream before waiting for any, so the GPU runs the domains' small kernels
 * side by side.  No reference counterpart (the reference runs one track per
 * job).  outs (n entries) may be NULL. */
amca_status amca_remelt_run_many(amca_remelt* const* ds, int32_t n, int64_t max_steps,
                                 amca_remelt_counters* outs);
amca_status amca_remelt_get_counters(amca_remelt* d, amca_remelt_counters* out);
amca_status amca_remelt_enable_log(amca_remelt* d, int32_t capacity);
amca_status amca_remelt_read_log(amca_remelt* d, int32_t cap, amca_remelt_step_record* rows,
                                 int32_t* n_rows);
/* cellOrientationID_, cellGrainID_, cellTemperature_, hasBeenMelted_, cooling_
 * per cell; grain records (envelope, cell, orientation) by grain id. */
amca_status amca_remelt_download(amca_remelt* d, int32_t* orient, int32_t* grain, double* T,
                                 uint8_t* melted, uint8_t* cooling, int32_t grain_cap,
                                 double* genv, int32_t* gcell, int32_t* gorient,
                                 int32_t* n_grains);
/* thetaArray_: the nodal field the next step interpolates from. */
amca_status amca_remelt_theta(amca_remelt* d, double* theta);
/* Temperatures from the trained surrogate instead of the FE path: the network
 * is evaluated once on moving-frame lattices (cell y, z and node y, z; xi
 * spacing h) for the scan; each step a cell takes the table value unless it is
 * within `band` K of Tm or in an element one of whose nodes differs from the
 * FE nodal field by more than `resid` K, where it takes the FE path.  The mesh
 * must be the uniform lattice `sd` describes.  Results then differ from the
 * FE-driven run (no counterpart in the reference). */
amca_status amca_remelt_use_surrogate(amca_remelt* d, amca_mlp* m, const amca_surrogate_desc* sd,
                                      const amca_scan* scan, double band, double resid);
/* Surrogate error against the FE path at the current FE time: out6 = RMS over
 * cells with material, their count, RMS over the cells the surrogate serves,
 * their count, their max |error|, cells on the FE path. */
amca_status amca_remelt_surrogate_error(amca_remelt* d, double* out6);
/* Active box (on by default): each step's grid kernels run over the box of
 * cells that can change in that step -- elements at or above Tm - 1 K or
 * changing between void and material, liquid cells with material, last
 * step's captures, widened by two cells -- and a run ends by writing the last
 * step's T and cooling to the cells outside it.  Results are identical with
 * it off (every step over the whole grid, as the reference's loops run). */
amca_status amca_remelt_set_window(amca_remelt* d, int32_t on);
/* Orientation rows drawn ahead of the rows in use (default 4096).  A step
 * runs only when every site whose cell is liquid at its start could nucleate
 * with the rows pushed; otherwise the run draws more and repeats it.  Results
 * do not depend on it. */
amca_status amca_remelt_set_row_reserve(amca_remelt* d, int32_t rows);
/* Steps per HIP graph when a run is not timed (default 8; 0: every kernel is
 * launched on the domain's stream).  Results do not depend on it. */
amca_status amca_remelt_set_graph_steps(amca_remelt* d, int32_t steps);
/* The last step's box (i0, j0, k0, i1, j1, k1, ghost layer included). */
amca_status amca_remelt_last_box(amca_remelt* d, int32_t* box6);
/* Per-kernel device time from HIP events around every launch (off by default);
 * enabling resets the totals.  Kernel i is named amca_remelt_kernel_name(i). */
amca_status amca_remelt_enable_timing(amca_remelt* d, int32_t on);
int32_t amca_remelt_kernel_count(void);
const char* amca_remelt_kernel_name(int32_t i);
amca_status amca_remelt_kernel_times(amca_remelt* d, double* ms, int64_t* launches);

/* Input generators (host only, no GPU).  Nodal frames of a Rosenthal moving
 * point source, T0 + eta*P/(2 pi k R) exp(-v (R + xi)/(2 alpha)) with
 * xi = x - (x_start + v t), R floored at r_min, for t = t0 + f*dt; out is
 * [n_frames][n_nodes].  No counterpart in the reference, which reads such
 * frames from text files (LoadDataFromFile.cpp:21-191). */
amca_status amca_rosenthal_frames(int64_t n_nodes, const double* node_xyz, int32_t n_frames,
                                  double t0, double dt, double P, double eta, double v, double k,
                                  double alpha, double T0, double x_start, double y_c,
                                  double z_top, double r_min, double* out);
/* Voronoi substrate: label[NG] = nearest of n_seeds seed points (interior
 * cells; ghosts -1), euler3[3*n_seeds] uniform on SO(3), from a splitmix64
 * stream of `seed`.  No counterpart in the reference, which loads such a
 * microstructure from grainInfo text files (:165-307). */
amca_status amca_voronoi_substrate(int32_t nx, int32_t ny, int32_t nz, double h, double x0,
                                   double y0, double z0, int32_t n_seeds, uint64_t seed,
                                   int32_t* label, double* euler3);
/* Writes nodal frames as the text file a finite_element realm reads
 * (LoadDataFromFile.cpp:21-191 with lines_for_subtitle 5; format of the
 * README "PBF_x" example): `title_lines` title lines, then per frame a
 * five-line subtitle block whose fourth line is
 * "f+1 f+1 t t 0 fnx-1 0 fny-1 0 fnz-1", and one "x y z T" line per node
 * (x fastest); frames after the first open with an empty line.  Doubles carry
 * 17 significant digits, so the reader recovers them exactly.  node_xyz is
 * [nodes][3] as written (the deck's length unit and offset apply on reading),
 * times[n_frames] absolute, theta [n_frames][nodes].  No counterpart in the
 * reference, whose frames come from an external FE code; with
 * amca_rosenthal_frames it produces deck-ready input.  AMCA_ERR_IO if the
 * file cannot be written. */
amca_status amca_write_fe_frames(const char* path, int32_t title_lines, int32_t fnx, int32_t fny,
                                 int32_t fnz, const double* node_xyz, int32_t n_frames,
                                 const double* times, const double* theta);

/* out[i] = pow(x[i], y) evaluated by the device code of the remelting growth
 * velocity (ProblemPhysics.cpp:152 calls glibc's pow(dT, b)), which restates
 * glibc 2.35's pow and answers the exponents 1, 2 and 3 from exact products
 * where that gives glibc's result: for finite x[i] > 0 and y == 0 or |y| in
 * [2^-65, 2^63) the values equal glibc's bit for bit.  A check of the device
 * arithmetic; AMCA_ERR_INVALID for other arguments. */
amca_status amca_pow_selftest(int64_t n, const double* x, double y, double* out);
const char* amca_remelt_last_error(void);

/* ---- realm layer: YAML input deck -> device domain ------------------------
 * The reference's Simulation / CellularAutomataSolver layer
 * (Simulation.cpp:36-93, CellularAutomataSolver.cpp:40-202) with the same
 * realm keys; host RNG draws (sites, orientations) happen here. */
typedef struct amca_sim amca_sim;

typedef struct {
  int32_t nx, ny, nz;
  double h;
  double termination_time, time_step;
  int32_t n_sites, n_surface_sites, n_bulk_sites;
  int64_t n_site_draws;
  int32_t coupled;          /* 1: finite_element + cellular_automata_remelting (CafeSolver) */
  int32_t n_frames, fe_nodes, n_init_grains;
} amca_sim_info;

/* Simulation::load on a YAML document held in memory / in a file.  With the
 * `finite_element_method` solver, a `finite_element` realm (frames from its
 * `mesh` text file, load_data_from_file keys) and a
 * `cellular_automata_remelting` realm make a coupled run (CafeSolver); file
 * names are taken relative to the deck's directory. */
amca_status amca_sim_create(const char* yaml_text, amca_sim** out);
amca_status amca_sim_create_from_file(const char* path, amca_sim** out);
/* The host half of CellularAutomataManager_3D::initialize (:366-379): the
 * rand()-stream draws of nucleation sites and orientation rows.  Needs no GPU. */
amca_status amca_sim_prepare(amca_sim* s);
/* Simulation::initialize: prepare (if not done) and upload to a new domain. */
amca_status amca_sim_initialize(amca_sim* s);
/* Simulation::run -> time_integration; max_steps < 0 runs to termination. */
amca_status amca_sim_run(amca_sim* s, int64_t max_steps, amca_counters* out);
amca_domain* amca_sim_domain(amca_sim* s);
/* The coupled domain of a `finite_element_method` run (NULL for plain runs). */
amca_remelt* amca_sim_remelt(amca_sim* s);
/* The end of CellularAutomataSolver / CafeSolver::execute
 * (CellularAutomataSolver.cpp:197-200, CafeSolver.cpp:252-260): writes the
 * output_microstructure_information file (rank suffix "0.txt") when the deck
 * asks for it, then finalGrains.vtk (in the deck's folder, or the working
 * directory for a deck given as text) and finalResult.txt (in the folder of
 * output_data_base_name, created if missing).  The VTR time series is not
 * written. */
amca_status amca_sim_finalize(amca_sim* s);
/* The VTR time series of the realm's output block (VtrManager, VtrManager.cpp:41-561;
 * off until enabled, as a library call should not scatter folders): with it on,
 * amca_sim_run writes a frame after every step that output_frequency (or,
 * below 1, output_time_interval) names, and amca_sim_finalize the last one --
 * <dir>/grainsTimeStates/timeStep_NNNNN/<name>00000.vtr and <name>NNNNN.pvtr,
 * and <dir>/<name>.pvd, where <dir>/<name> is output_data_base_name (relative
 * to the deck's folder).  The amca3d driver turns it on. */
amca_status amca_sim_enable_vtr(amca_sim* s, int32_t on);
/* CellularAutomataManager_3D::writeToFile (CellularAutomataManager_3D.cpp:1381-1639):
 * finalGrains.vtk (legacy ASCII rectilinear grid: grain_id, number_cells,
 * melted_flag, EulerAngle) and finalResult.txt (one row per cell) from host
 * arrays.  orient[ng] and melted[ng] are ghost-inclusive (ng = (nx+2)(ny+2)(nz+2),
 * x fastest; melted may be null: every flag 1); euler3 holds 3 angles per
 * orientation id.  A null path skips that file.  No GPU is used. */
amca_status amca_write_final_grains(const char* vtk_path, const char* result_path, int32_t nx,
                                    int32_t ny, int32_t nz, double x0, double y0, double z0,
                                    double h, const int32_t* orient, const int32_t* melted,
                                    const double* euler3, int32_t n_orient);
amca_status amca_sim_get_info(amca_sim* s, amca_sim_info* info);
/* The sites and orientation rows drawn at initialisation (for checking). */
amca_status amca_sim_initial_sites(amca_sim* s, int32_t cap, int32_t* cell, double* dT_crit,
                                   double* euler3, int32_t* n);
amca_status amca_sim_destroy(amca_sim* s);
const char* amca_sim_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* AMCA3D_H */
