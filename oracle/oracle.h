/* ORACLE -- test infrastructure only (CPU restatement of the reference path).
 * Only tests/, the smoke check and the benchmark's CPU-baseline leg load this. */
#ifndef AMCA_ORACLE_H
#define AMCA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t nx, ny, nz;
  double h, x0, y0, z0;
  double ns, dTs_mean, dTs_sd;   /* surface site density / Gaussian undercooling */
  double nv, dTv_mean, dTv_sd;   /* bulk */
  int32_t surface_mode;          /* 0: the reference's six-face list, 1: z-min face only */
  uint32_t seed;                 /* srand() seed; 0 is the np=1 reference value */
  double T0, Tm, Tdot, a1, a2, a3;
  double trunc;                  /* envelope truncation coefficient (float-rounded) */
  double tsf, max_dt, final_time;
  int32_t max_steps;             /* < 0: run until final_time or full solidification */
} oracle_plain_params;

typedef struct {
  double time_before, dt;
  int32_t n_nucleated, n_active_before, n_captured, n_active_after, n_liquid, max_wave;
  int32_t n_wave1; /* captures made by grains active at the step's start */
} oracle_step_record;

typedef struct {
  /* caller-provided buffers (any may be NULL) */
  int32_t *ori, *gid;            /* [ng] */
  double *nuc_dt;                /* [ng] */
  double *grain_env;             /* [4*grain_cap] xc,yc,zc,L */
  int32_t *grain_cell, *grain_orient;
  uint8_t *grain_active;
  int32_t grain_cap;
  double *euler;                 /* [3*orient_cap] */
  int32_t orient_cap;
  int32_t *site_cell; double *site_dt; int32_t site_cap;
  oracle_step_record *steps; int32_t steps_cap;
  /* results */
  int32_t ng, n_steps, n_records, done, n_grains, n_orient, n_active, n_sites;
  int64_t n_site_draws;
  double time;
} oracle_plain_out;

int oracle_plain_run(const oracle_plain_params* prm, oracle_plain_out* out);
/* same loop with explicit sites and orientation angles (consumed in
 * nucleation order) instead of the rand() draws */
int oracle_plain_run_with(const oracle_plain_params* prm, int n_sites, const int32_t* site_cell,
                          const double* site_dt, int n_orient, const double* euler3,
                          oracle_plain_out* out);
void oracle_rotation(const double* euler3, double* l2g9);
int oracle_capture_test(const double* env4, const double* l2g9, const double* cell3);
void oracle_child_envelope(const double* gc3, double L, double h, double trunc,
                           const double* l2g9, double* out4);
void oracle_glibc_rand(uint32_t seed, int n, int32_t* out);
void oracle_std_normal(double mean, double sd, int n, double* out);
void oracle_minstd(int n, int64_t* out);
float oracle_sqrtf(float x);
/* capture test + child envelope over n independent inputs */
void oracle_envelope_batch(int n, const double* env4, const double* l2g9, const double* cell3,
                           double h, double trunc, int32_t* inside, double* child4);

/* ---- coupled CA-FE (remelting) path ---------------------------------- */
typedef struct {
  int32_t nx, ny, nz;
  double h, x0, y0, z0;
  /* structured FE mesh: node (i,j,k) has id (k*fny + j)*fnx + i */
  int32_t fnx, fny, fnz;
  const double* node_xyz;        /* [3*nodes], after offset and length scale */
  int32_t n_frames;
  const double* frame_time;      /* [n_frames], after time scale */
  const double* frame_theta;     /* [n_frames*nodes] */
  int32_t bin[3];
  double tol;
  int32_t max_iter;
  double void_temp;
  /* physics */
  double Tm, Ts, Tmin_nuc, a1, a2, a3, b1, b2, b3;
  double trunc, tsf, cutoff_factor, final_time;
  int32_t max_steps;
  int32_t initial_melted;        /* phase_state.has_been_melted */
  /* nucleation sites (criterion_to_activate_grain.using_nucleation_seed) */
  int32_t use_sites;
  double ns, dTs_mean, dTs_sd, nv, dTv_mean, dTv_sd;
  uint32_t seed;
  /* loaded microstructure (NULL init_ori: none) */
  const int32_t* init_ori;       /* [ng] */
  const int32_t* init_gid;       /* [ng] */
  const uint8_t* init_melted;    /* [ng] */
  int32_t n_init_orient;
  const double* init_euler;      /* [3*n_init_orient] */
  int32_t n_init_grains;
  const int32_t* init_grain_cell;
  const int32_t* init_grain_orient;
  /* 1: random_microstructure_information instead of a loaded state */
  int32_t random_micro;
  /* seeds of a loaded microstructure file (load_seeds): per-cell critical
     undercooling, > -1 marks a site; NULL: none */
  const double* init_site_dt;    /* [ng] */
} oracle_remelt_params;

typedef struct {
  double time_before, dt;
  int32_t n_nucleated, n_captured, n_fresh, n_flagged, n_liquid;
} oracle_remelt_step;

typedef struct {
  int32_t *ori, *gid;            /* [ng] */
  double* T;                     /* [ng] */
  uint8_t *melted, *cooling;     /* [ng] */
  int32_t* tag;                  /* [ng] element tags */
  double* par;                   /* [3*ng] parametric coordinates */
  double* theta;                 /* [nodes] interpolated nodal field of the last step */
  double* grain_env;             /* [4*grain_cap] */
  int32_t *grain_cell, *grain_orient;
  int32_t grain_cap;
  double* euler;
  int32_t orient_cap;
  oracle_remelt_step* steps;
  int32_t steps_cap;
  int32_t* site_cell;            /* [site_cap] nucleation sites in list order */
  double* site_dt;               /* their critical undercoolings */
  int32_t site_cap;
  /* results */
  int32_t n_steps, n_records, n_grains, n_orient, frames_exhausted, n_sites;
  double time, time_fe;
} oracle_remelt_out;

int oracle_remelt_run(const oracle_remelt_params* prm, oracle_remelt_out* out);

#ifdef __cplusplus
}
#endif
#endif
