/*
 * pbf_oracle.h - CPU restatement of the reference's PBF step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared
 * against; it is not part of the product.  Only tests/, the smoke() check in
 * __graft_entry__.py and bench.py's cpu_baseline leg may load it, and they use
 * it as the checker or as the timed CPU baseline, never as the measured or
 * shipped path.
 *
 * It restates the per-frame step of niklasandersson/PBD-Simulation-Framework-CUDA
 * (src/app/simulation/Fluid.cu:14-97 and the kernels in
 * src/app/simulation/kernel/ headers).  Arithmetic:
 *   - the predict stage (applyForces + confineToBox) and the cell ids are
 *     computed in FP32 with the exact operation order given below and no
 *     contraction, so that cell ids and the sorted order are bit-exact against
 *     the GPU path;
 *   - everything after the reorder (density solver, collisions, update,
 *     vorticity, XSPH) is FP64.
 * Each function cites the reference file:line it follows.  Where the
 * reference is racy or has a defect, the deterministic definition used here
 * is stated next to the function.
 */
#ifndef PBF_ORACLE_H
#define PBF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "pbd/pbd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Parameters derived from pbd_params for one step (kernel/Initialize.h:20-86). */
typedef struct orc_derived {
  int variant;                 /* PBD_KERNEL_* */
  uint32_t cap;                /* maxNeighboursPerParticle */
  uint32_t kw;                 /* kernelWidthNeighbours */
  double diameter;             /* 2 r (Initialize.h:43) */
  double h_spiky;              /* 2 r * Density.kernelWidthSpiky (Initialize.h:45) */
  double h_poly;               /* Viscosity.kernelWidthPoly (Initialize.h:53) */
  double dq;                   /* h_spiky * Density.qSCorr (Initialize.h:49) */
  double w_dq;                 /* W_poly(dq) */
  double rho0;
  double k_scorr;
  int n_scorr;
  double eps;                  /* Density.relaxation */
  double stiffness;
  double c_visc;
  double e_vort;
  double dt;                   /* delta_t / substeps */
  float dt_f;
  float gravity_f;
  float bmin_f[3], bmax_f[3];
  float vdamp_f, pdamp_f;
  float cell_f, inv_cell_f;    /* grid cell size and its FP32 reciprocal */
  float origin_f[3];
  uint32_t dims[3];
  uint32_t table_size;         /* morton(dims - 1) + 1 */
  int collision_type;
  uint32_t solver_iters, stab_iters, max_batches;
  int g_apply_forces, g_find_contacts, g_find_neighbours, g_lambda, g_delta,
      g_apply_delta, g_update, g_omega, g_vorticity, g_viscosity;
  double mu_s, mu_k;           /* Collision.frictionStatic / frictionKinetic */
} orc_derived;

/* In/out arrays of one step.  State arrays (float) are read and updated in
 * place; output arrays (double / uint32) are optional (NULL = not wanted). */
typedef struct orc_io {
  uint32_t n;
  /* state, FP32, float4 per particle, updated in place (sorted order after the step) */
  float* x4;          /* positions4 */
  float* p4;          /* predictedPositions4 */
  float* v4;          /* velocities4 */
  float* c4;          /* colors4 */
  float* f4;          /* d_externalForces */
  double* lambda;     /* d_lambdas, persistent across steps (n) */
  double* dp4;        /* d_deltaPositions, persistent across steps (4n) */
  /* FP64 outputs of the step (4n / n / 3n) */
  double* out_x4;
  double* out_p4;
  double* out_v4;
  double* out_f4;
  double* out_density;
  double* out_omega3;
  /* integer outputs */
  uint32_t* out_keys_unsorted; /* cell ids before the sort (n) */
  uint32_t* out_keys;          /* sorted cell ids (n) */
  uint32_t* out_ids;           /* sorted previous indices (n) */
  uint32_t* out_starts;        /* table_size */
  uint32_t* out_ends;          /* table_size */
  uint32_t* out_nbr_count;     /* n */
  uint32_t* out_contact_count; /* n */
  /* statistics of pair distances seen by support tests (min |r - cutoff|) */
  double min_cutoff_margin;
  /* phases (PBD_PHASE_*, n; NULL = all fluid), updated in place (sorted order after the step) */
  uint32_t* phase;
} orc_io;

/* ---- building blocks ---- */
int orc_derive(const pbd_params* p, orc_derived* d);
uint32_t orc_expand_bits(uint32_t v);
uint32_t orc_morton(uint32_t x, uint32_t y, uint32_t z);
uint32_t orc_cell_key(const orc_derived* d, const float p[3], uint32_t ijk[3]);
double orc_w(const orc_derived* d, double r);                     /* poly6 variant, width h_poly */
void orc_grad(const orc_derived* d, const double r[3], double g[3]); /* spiky gradient, width h_spiky */
double orc_lattice_rest_density(const orc_derived* d, double spacing);
void orc_predict(const orc_derived* d, uint32_t n, float* x4, float* v4, const float* f4, float* p4);
void orc_sort(uint32_t n, const uint32_t* keys_in, uint32_t* keys_out, uint32_t* ids_out);
void orc_cell_ranges(uint32_t n, const uint32_t* keys_sorted, uint32_t table_size,
                     uint32_t* starts, uint32_t* ends);
/* Neighbour list of sorted particle i in the reference order; returns the
 * count and the contact count (first part of the list). */
uint32_t orc_neighbours(const orc_derived* d, uint32_t n, const uint32_t* keys_sorted,
                        const uint32_t* starts, const uint32_t* ends, uint32_t i,
                        uint32_t* out, uint32_t* n_contacts);

/* ---- the step ---- */
int orc_step(const pbd_params* p, orc_io* io);
/* ---- FP64 multi-step forward and its reverse mode (pbf_adjoint.c) ----
 * Textbook variant, collisionType 0, at most one stabilization pass, all
 * stages on (else PBD_ERR_UNSUPPORTED); a neighbour list reaching the cap is
 * PBD_ERR_UNSUPPORTED as well.  State: n x 4 doubles (w ignored).  rho0 <= 0
 * and eps < 0 select the values derived from the parameters.  phase: n phases
 * (PBD_PHASE_*; NULL: all fluid), carried through the sorts like orig; solid
 * particles and friction as in orc_step. */
int orc64_forward(const pbd_params* p, double rho0, double eps, uint32_t n, uint32_t nsub, double* x4,
                  double* v4, double* f4, uint32_t* orig, uint32_t* phase);
/* dL/dx0, dL/dv0 (order of the initial state), dL/drho0, dL/deps of a loss with
 * adjoints gx4, gv4 of the final positions and velocities (final order);
 * phase0: the initial phases (NULL: all fluid). */
int orc64_gradient(const pbd_params* p, double rho0, double eps, uint32_t n, uint32_t nsub, const double* x4,
                   const double* v4, const double* f4, const uint32_t* phase0, const double* gx4,
                   const double* gv4, double* gx0, double* gv0, double* grho0, double* geps);

/* Number of OpenMP threads the oracle uses (1 when built without OpenMP). */
int orc_threads(void);
void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif

#endif /* PBF_ORACLE_H */
