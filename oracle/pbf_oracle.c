/*
 * pbf_oracle.c - CPU restatement of the reference's PBF step.
 *
 * TEST INFRASTRUCTURE ONLY (see pbf_oracle.h): the checker for the HIP path
 * and the timed CPU baseline; never the product path.
 *
 * Reference files restated (paths relative to PBD-Simulation-Framework-CUDA/src):
 *   app/simulation/Fluid.cu:14-97            stage order and Active.* gates
 *   app/simulation/kernel/Initialize.h:20-86 derived parameters
 *   app/simulation/kernel/Forces.h:19-101    predict + box clamp (FP32 here)
 *   app/simulation/kernel/Hash.h:10-34       Morton cell id (FP32 clamp/floor here)
 *   app/simulation/kernel/SortReorder.h      cell ids, stable sort, reorder, cell ranges
 *   app/simulation/kernel/Collision.h        contact / neighbour lists, collisions
 *   app/simulation/kernel/Density.h          kernels, lambda, delta p, omega, vorticity, XSPH
 *   app/simulation/kernel/Update.h:11-32     velocity / position update
 *
 * Build-side definitions where the reference is racy or defective:
 *   - collisions are a gather over the pre-collision state (Jacobi) instead
 *     of FP32 atomics (Collision.h:229-236);
 *   - XSPH reads the pre-XSPH velocities (Jacobi) instead of updating in
 *     place (Density.h:307-308);
 *   - cell ranges use the empty-cell convention {UINT_MAX, N} without the
 *     last-cell write to cellEndings[idx] (SortReorder.h:183-185);
 *   - the neighbour stencil is the integer cell stencil around the particle's
 *     own cell (the reference hashes p + offset * diameter, which is the same
 *     stencil when the cell size equals the diameter, up to FP32 rounding of
 *     the offset position and the low-edge clamp of Hash.h:21-23);
 *   - the s_corr divisor W(dq) = 0 gives s_corr = 0 instead of a division by 0.
 *
 * Textbook variant (Macklin & Mueller 2013 with Mueller et al. 2003 kernels),
 * made dimensionally consistent for world units (h = 0.1 m):
 *   - s_corr = -k h_spiky^2 (W/W(dq))^n: the paper states k for a unit kernel
 *     radius; the h^2 gives s_corr the units of lambda;
 *   - XSPH v_i += c (1/rho0) sum v_ij W_ij (Monaghan's XSPH with the particle
 *     volume m/rho0, mass 1) and the vorticity omega_i = (1/rho0) sum v_ij x grad_j W_ij.
 * Without these factors c = 0.01 and k = 0.1 at h = 0.1 amplify velocities
 * by ~100x per step and the dam break diverges within three steps.
 */
#include "pbf_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI_REF 3.14159265359 /* M_PI as #defined by the reference, kernel/Globals.h:122 */
#define ORC_PI 3.14159265358979323846

int orc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ---------------------------------------------------------------------------
 * Morton code.  Hash.h:10-16 spreads a 10-bit integer so that bit b lands at
 * bit 3b; Hash.h:21-29 interleaves x (bit 3b+2), y (3b+1), z (3b).
 * Restated here as a plain bit loop.
 * ------------------------------------------------------------------------- */
uint32_t orc_expand_bits(uint32_t v) {
  uint32_t out = 0;
  for (uint32_t b = 0; b < 10; ++b) out |= ((v >> b) & 1u) << (3u * b);
  return out;
}

uint32_t orc_morton(uint32_t x, uint32_t y, uint32_t z) {
  return (orc_expand_bits(x) << 2) + (orc_expand_bits(y) << 1) + orc_expand_bits(z);
}

static uint32_t orc_compact_bits(uint32_t v) {
  uint32_t out = 0;
  for (uint32_t b = 0; b < 10; ++b) out |= ((v >> (3u * b)) & 1u) << b;
  return out;
}

/* Cell id of an FP32 position (Hash.h:21-34 generalised to an origin, a cell
 * size and per-axis dims): t = (p - origin) * inv_cell in FP32 (two rounded
 * ops), i = (uint)clamp(floor(t), 0, dim - 1).  With origin 0 and cell 1 this
 * is the reference's trunc(clamp(p, 0, 1023)) for every p the reference can
 * hash without leaving its 128^3 table. */
uint32_t orc_cell_key(const orc_derived* d, const float p[3], uint32_t ijk[3]) {
  uint32_t c[3];
  for (int a = 0; a < 3; ++a) {
    volatile float diff = p[a] - d->origin_f[a];
    volatile float t = diff * d->inv_cell_f;
    float f = floorf(t);
    f = fmaxf(f, 0.0f);
    f = fminf(f, (float)(d->dims[a] - 1u));
    c[a] = (uint32_t)f;
  }
  if (ijk) { ijk[0] = c[0]; ijk[1] = c[1]; ijk[2] = c[2]; }
  return orc_morton(c[0], c[1], c[2]);
}

/* ---------------------------------------------------------------------------
 * Kernels.
 *   reference poly6Viscosity, Density.h:14-25:
 *     W = 315 ((h^2 - r^2)^3)^2 / (0.001 + 64 pi h^9),  0.001 < r < h - 0.001
 *   reference spiky, Density.h:31-42:
 *     grad = 15 (h - r)^2 / (pi h^6) * (p_i - p_j)/r,  r > 1e-4, h - r > 1e-4
 *   textbook (Mueller et al. 2003; Macklin & Mueller 2013):
 *     W = 315 / (64 pi h^9) (h^2 - r^2)^3,  0 <= r < h
 *     grad_i W = -45 / (pi h^6) (h - r)^2 (p_i - p_j)/r,  0 < r < h
 * ------------------------------------------------------------------------- */
double orc_w(const orc_derived* d, double r) {
  const double h = d->h_poly;
  if (d->variant == PBD_KERNEL_REFERENCE) {
    if (r > 0.001 && r < (h - 0.001)) {
      double t = h * h - r * r;
      double t3 = t * t * t;
      return (315.0 * t3 * t3) / (0.001 + 64.0 * ORC_PI_REF * pow(h, 9));
    }
    return 0.0;
  }
  if (r >= 0.0 && r < h) {
    double t = h * h - r * r;
    return 315.0 / (64.0 * ORC_PI * pow(h, 9)) * t * t * t;
  }
  return 0.0;
}

void orc_grad(const orc_derived* d, const double r[3], double g[3]) {
  const double h = d->h_spiky;
  double len = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  double s = 0.0;
  if (d->variant == PBD_KERNEL_REFERENCE) {
    if (len > 0.0001 && (h - len) > 0.0001) {
      double hr = h - len;
      s = 15.0 * (hr * hr) / (ORC_PI_REF * pow(h, 6)) / len;
    }
  } else {
    if (len > 0.0 && len < h) {
      double hr = h - len;
      s = -45.0 / (ORC_PI * pow(h, 6)) * hr * hr / len;
    }
  }
  g[0] = s * r[0];
  g[1] = s * r[1];
  g[2] = s * r[2];
}

double orc_lattice_rest_density(const orc_derived* d, double s) {
  if (!(s > 0.0)) return 0.0;
  int m = (int)ceil(d->h_poly / s) + 1;
  double sum = 0.0;
  for (int a = -m; a <= m; ++a)
    for (int b = -m; b <= m; ++b)
      for (int c = -m; c <= m; ++c) sum += orc_w(d, s * sqrt((double)(a * a + b * b + c * c)));
  return sum;
}

/* Derived parameters, following initializeFrame (Initialize.h:34-53): the
 * widths are FP32 products exactly as the reference forms them. */
int orc_derive(const pbd_params* p, orc_derived* d) {
  memset(d, 0, sizeof(*d));
  if (!p) return PBD_ERR_INVALID;
  if (p->kernel_variant != PBD_KERNEL_REFERENCE && p->kernel_variant != PBD_KERNEL_TEXTBOOK)
    return PBD_ERR_INVALID;
  if (p->max_neighbours_per_particle == 0 || p->substeps == 0) return PBD_ERR_INVALID;
  for (int a = 0; a < 3; ++a)
    if (p->grid_dims[a] < 1 || p->grid_dims[a] > 1024) return PBD_ERR_INVALID;
  d->variant = p->kernel_variant;
  d->cap = p->max_neighbours_per_particle;
  d->kw = p->kernel_width_neighbours;
  float diam = 2.0f * p->particle_radius;
  float hs = diam * p->kernel_width_spiky;
  float dq = hs * p->q_scorr;
  d->diameter = diam;
  d->h_spiky = hs;
  d->h_poly = p->kernel_width_poly;
  d->dq = dq;
  d->w_dq = orc_w(d, d->dq);
  d->k_scorr = p->k_scorr;
  d->n_scorr = p->n_scorr;
  d->eps = p->relaxation;
  d->stiffness = p->stiffness;
  d->c_visc = p->c_viscosity;
  d->e_vort = p->e_vorticity;
  d->dt_f = p->delta_t / (float)p->substeps;
  d->dt = d->dt_f;
  d->gravity_f = p->gravity;
  for (int a = 0; a < 3; ++a) {
    d->bmin_f[a] = p->bounds_min[a];
    d->bmax_f[a] = p->bounds_max[a];
    d->origin_f[a] = p->grid_origin[a];
    d->dims[a] = p->grid_dims[a];
  }
  d->vdamp_f = p->enclosure_velocity_damping;
  d->pdamp_f = p->enclosure_position_damping;
  d->cell_f = p->grid_cell_size > 0.0f ? p->grid_cell_size : diam;
  if (!(d->cell_f > 0.0f)) return PBD_ERR_INVALID;
  d->inv_cell_f = 1.0f / d->cell_f;
  d->table_size = orc_morton(d->dims[0] - 1u, d->dims[1] - 1u, d->dims[2] - 1u) + 1u;
  d->collision_type = p->collision_type;
  d->solver_iters = p->solver_iterations;
  d->stab_iters = p->stabilization_iterations;
  d->max_batches = p->max_collision_batches;
  d->g_apply_forces = p->active_apply_forces != 0;
  d->g_find_contacts = p->active_find_contacts != 0;
  d->g_find_neighbours = p->active_find_neighbours != 0;
  d->g_lambda = p->active_compute_lambda != 0;
  d->g_delta = p->active_compute_delta_positions != 0;
  d->g_apply_delta = p->active_apply_delta_positions != 0;
  d->g_update = p->active_update_positions != 0;
  d->g_omega = p->active_compute_omegas != 0;
  d->g_vorticity = p->active_compute_vorticity != 0;
  d->g_viscosity = p->active_compute_viscosity != 0;
  d->mu_s = p->friction_static;
  d->mu_k = p->friction_kinetic;
  if (p->rest_density_from_lattice)
    d->rho0 = orc_lattice_rest_density(d, p->box_spacing);
  else
    d->rho0 = p->rest_density;
  if (!(d->rho0 > 0.0)) return PBD_ERR_INVALID;
  return PBD_OK;
}

/* ---------------------------------------------------------------------------
 * applyForces + confineToBox (Forces.h:19-101), FP32, no contraction:
 *   v.y = v.y + (1*g)*dt;  v.c = v.c + f.c*dt;  p.c = x.c + v.c*dt  (c = x,y,z,w)
 *   per axis: p < min -> v = damp*v, p = min + 0.001;  p > max -> v = damp*v, p = max - 0.001
 *   any clamp -> x.c = x.c + pdamp*(p.c - x.c)   (c = x,y,z,w; p.w not yet reset)
 *   p.w = 1
 * ------------------------------------------------------------------------- */
void orc_predict(const orc_derived* d, uint32_t n, float* x4, float* v4, const float* f4, float* p4) {
  const float dt = d->dt_f;
  for (uint32_t i = 0; i < n; ++i) {
    volatile float v[4], x[4], p[4], t;
    for (int c = 0; c < 4; ++c) { v[c] = v4[4 * i + c]; x[c] = x4[4 * i + c]; }
    t = 1.0f * d->gravity_f;
    t = t * dt;
    v[1] = v[1] + t;
    for (int c = 0; c < 4; ++c) {
      t = f4[4 * i + c] * dt;
      v[c] = v[c] + t;
    }
    for (int c = 0; c < 4; ++c) {
      t = v[c] * dt;
      p[c] = x[c] + t;
    }
    int upd = 0;
    for (int a = 0; a < 3; ++a) {
      if (p[a] < d->bmin_f[a]) {
        v[a] = d->vdamp_f * v[a];
        p[a] = d->bmin_f[a] + 0.001f;
        upd = 1;
      } else if (p[a] > d->bmax_f[a]) {
        v[a] = d->vdamp_f * v[a];
        p[a] = d->bmax_f[a] - 0.001f;
        upd = 1;
      }
    }
    if (upd) {
      for (int c = 0; c < 4; ++c) {
        t = p[c] - x[c];
        t = d->pdamp_f * t;
        x[c] = x[c] + t;
      }
    }
    p[3] = 1.0f;
    for (int c = 0; c < 4; ++c) {
      v4[4 * i + c] = v[c];
      x4[4 * i + c] = x[c];
      p4[4 * i + c] = p[c];
    }
  }
}

/* Stable sort of (cell id, index) (SortReorder.h:39-47 via CUB SortPairs):
 * sorting the unique 64-bit pairs key<<32 | index gives the stable order. */
static int cmp_u64(const void* a, const void* b) {
  uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return (x > y) - (x < y);
}

void orc_sort(uint32_t n, const uint32_t* keys_in, uint32_t* keys_out, uint32_t* ids_out) {
  uint64_t* t = (uint64_t*)malloc((size_t)(n ? n : 1) * sizeof(uint64_t));
  for (uint32_t i = 0; i < n; ++i) t[i] = ((uint64_t)keys_in[i] << 32) | i;
  qsort(t, n, sizeof(uint64_t), cmp_u64);
  for (uint32_t i = 0; i < n; ++i) {
    keys_out[i] = (uint32_t)(t[i] >> 32);
    ids_out[i] = (uint32_t)(t[i] & 0xFFFFFFFFu);
  }
  free(t);
}

/* resetCellInfo + computeCellInfo (SortReorder.h:135-188): empty cell =
 * {UINT_MAX, N}; a cell's start is its first sorted index, its end the
 * first index of the next occupied cell. */
void orc_cell_ranges(uint32_t n, const uint32_t* keys, uint32_t table_size, uint32_t* starts,
                     uint32_t* ends) {
  for (uint32_t c = 0; c < table_size; ++c) {
    starts[c] = 0xFFFFFFFFu;
    ends[c] = n;
  }
  for (uint32_t s = 0; s < n; ++s) {
    if (s == 0 || keys[s] != keys[s - 1]) starts[keys[s]] = s;
    if (s > 0 && keys[s] != keys[s - 1]) ends[keys[s - 1]] = s;
  }
}

/* findContacts (Collision.h:11-76) then findNeighbours (Collision.h:87-160):
 * the 27 cells around the particle's cell in (x, y, z) nested order -1..1,
 * then the shells 2..kernelWidthNeighbours; every particle j != i of each
 * occupied cell in sorted order; capped at maxNeighboursPerParticle. */
uint32_t orc_neighbours(const orc_derived* d, uint32_t n, const uint32_t* keys, const uint32_t* starts,
                        const uint32_t* ends, uint32_t i, uint32_t* out, uint32_t* n_contacts) {
  (void)n;
  const uint32_t key = keys[i];
  const int ix = (int)orc_compact_bits(key >> 2), iy = (int)orc_compact_bits(key >> 1),
            iz = (int)orc_compact_bits(key);
  const uint32_t cap = d->cap;
  uint32_t cnt = 0;
  if (d->g_find_contacts) {
    for (int dx = -1; dx <= 1; ++dx)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dz = -1; dz <= 1; ++dz) {
          int cx = ix + dx, cy = iy + dy, cz = iz + dz;
          if (cx < 0 || cy < 0 || cz < 0 || cx >= (int)d->dims[0] || cy >= (int)d->dims[1] ||
              cz >= (int)d->dims[2])
            continue;
          uint32_t k = orc_morton((uint32_t)cx, (uint32_t)cy, (uint32_t)cz);
          uint32_t s = starts[k];
          if (s == 0xFFFFFFFFu) continue;
          uint32_t e = ends[k];
          for (uint32_t j = s; j < e; ++j) {
            if (j == i) continue;
            out[cnt++] = j;
            if (cnt == cap) goto contacts_done;
          }
        }
  }
contacts_done:
  *n_contacts = cnt;
  if (d->g_find_neighbours && cnt < cap) {
    const int kw = (int)d->kw;
    for (int sh = 2; sh <= kw; ++sh)
      for (int dx = -sh; dx <= sh; ++dx)
        for (int dy = -sh; dy <= sh; ++dy)
          for (int dz = -sh; dz <= sh; ++dz) {
            if (abs(dx) != sh && abs(dy) != sh && abs(dz) != sh) continue;
            int cx = ix + dx, cy = iy + dy, cz = iz + dz;
            if (cx < 0 || cy < 0 || cz < 0 || cx >= (int)d->dims[0] || cy >= (int)d->dims[1] ||
                cz >= (int)d->dims[2])
              continue;
            uint32_t k = orc_morton((uint32_t)cx, (uint32_t)cy, (uint32_t)cz);
            uint32_t s = starts[k];
            if (s == 0xFFFFFFFFu) continue;
            uint32_t e = ends[k];
            for (uint32_t j = s; j < e; ++j) {
              if (j == i) continue;
              out[cnt++] = j;
              if (cnt == cap) return cnt;
            }
          }
  }
  return cnt;
}

/* ---------------------------------------------------------------------------
 * The step (Fluid::compute, Fluid.cu:14-97).
 * ------------------------------------------------------------------------- */
typedef struct orc_work {
  const orc_derived* d;
  uint32_t n;
  const uint32_t* keys;
  const uint32_t* starts;
  const uint32_t* ends;
  const uint32_t* ph; /* phases in sorted order, NULL = all fluid */
} orc_work;

/* Unified phase (beyond the reference): a solid particle has no density
 * constraint, vorticity or XSPH, and the fluid sums run over fluid neighbours. */
static inline int solid_at(const orc_work* w, uint32_t i) { return w->ph && w->ph[i] == PBD_PHASE_SOLID; }

static inline void sub3(const double* a, const double* b, double* r) {
  r[0] = a[0] - b[0];
  r[1] = a[1] - b[1];
  r[2] = a[2] - b[2];
}
static inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline double len3(const double* a) { return sqrt(dot3(a, a)); }
static inline void cross3(const double* a, const double* b, double* r) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

static double margin_ref(const orc_derived* d, double r) {
  double m = fabs(r - 0.001);
  double t = fabs(r - (d->h_poly - 0.001));
  if (t < m) m = t;
  t = fabs(r - 0.0001);
  if (t < m) m = t;
  t = fabs(r - (d->h_spiky - 0.0001));
  if (t < m) m = t;
  return m;
}

/* computeLambda (Density.h:48-93); also the density of the pass. */
static double pass_lambda(const orc_work* w, const double* P, double* lambda, double* density) {
  const orc_derived* d = w->d;
  double minm = DBL_MAX;
#pragma omp parallel reduction(min : minm)
  {
    uint32_t* nb = (uint32_t*)malloc(d->cap * sizeof(uint32_t));
#pragma omp for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)w->n; ++ii) {
      uint32_t i = (uint32_t)ii, nc;
      uint32_t cnt = orc_neighbours(d, w->n, w->keys, w->starts, w->ends, i, nb, &nc);
      double A = (d->variant == PBD_KERNEL_TEXTBOOK) ? orc_w(d, 0.0) : 0.0;
      double B = 0.0, C[3] = {0, 0, 0};
      for (uint32_t k = 0; k < cnt; ++k) {
        uint32_t j = nb[k];
        if (solid_at(w, j)) continue;
        double r[3], g[3];
        sub3(&P[4 * i], &P[4 * j], r);
        double rl = len3(r);
        if (d->variant == PBD_KERNEL_REFERENCE) {
          double m = margin_ref(d, rl);
          if (m < minm) minm = m;
        }
        A += orc_w(d, rl);
        orc_grad(d, r, g);
        B += dot3(g, g);
        C[0] += g[0];
        C[1] += g[1];
        C[2] += g[2];
      }
      double lam;
      if (d->variant == PBD_KERNEL_REFERENCE) {
        double a = A / d->rho0 - 1.0, b = B / d->rho0, c = dot3(C, C) / d->rho0;
        lam = -a / (b + c + d->eps);
      } else {
        double cc = A / d->rho0 - 1.0;
        double s = (dot3(C, C) + B) / (d->rho0 * d->rho0);
        lam = -cc / (s + d->eps);
      }
      lambda[i] = solid_at(w, i) ? 0.0 : lam;
      if (density) density[i] = A;
    }
    free(nb);
  }
  return minm;
}

/* computeDeltaPositions (Density.h:104-145). */
static void pass_delta(const orc_work* w, const double* P, const double* lambda, double* dp) {
  const orc_derived* d = w->d;
#pragma omp parallel
  {
    uint32_t* nb = (uint32_t*)malloc(d->cap * sizeof(uint32_t));
#pragma omp for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)w->n; ++ii) {
      uint32_t i = (uint32_t)ii, nc;
      uint32_t cnt = orc_neighbours(d, w->n, w->keys, w->starts, w->ends, i, nb, &nc);
      double acc[3] = {0, 0, 0};
      const double li = lambda[i];
      for (uint32_t k = 0; k < cnt; ++k) {
        uint32_t j = nb[k];
        if (solid_at(w, i) || solid_at(w, j)) continue;
        double r[3], g[3];
        sub3(&P[4 * i], &P[4 * j], r);
        double W = orc_w(d, len3(r));
        double sc = (d->w_dq != 0.0) ? -d->k_scorr * pow(W / d->w_dq, (double)d->n_scorr) : 0.0;
        if (d->variant == PBD_KERNEL_TEXTBOOK) sc *= d->h_spiky * d->h_spiky;
        orc_grad(d, r, g);
        double s = li + lambda[j] + sc;
        acc[0] += s * g[0];
        acc[1] += s * g[1];
        acc[2] += s * g[2];
      }
      dp[4 * i + 0] = acc[0] / d->rho0;
      dp[4 * i + 1] = acc[1] / d->rho0;
      dp[4 * i + 2] = acc[2] / d->rho0;
      dp[4 * i + 3] = 0.0;
    }
    free(nb);
  }
}

/* solveCollisions (Collision.h:171-242) as a gather over the contact list
 * using the pre-collision predicted positions; pushes p*_i and x_i. */
static double pass_collide(const orc_work* w, double* P, double* X, double* tmp, double* fric) {
  const orc_derived* d = w->d;
  double minm = DBL_MAX;
#pragma omp parallel reduction(min : minm)
  {
    uint32_t* nb = (uint32_t*)malloc(d->cap * sizeof(uint32_t));
#pragma omp for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)w->n; ++ii) {
      uint32_t i = (uint32_t)ii, nc;
      orc_neighbours(d, w->n, w->keys, w->starts, w->ends, i, nb, &nc);
      double push[3] = {0, 0, 0}, fr[3] = {0, 0, 0};
      for (uint32_t k = 0; k < nc; ++k) {
        uint32_t j = nb[k];
        double dir[3];
        sub3(&P[4 * j], &P[4 * i], dir);
        double len = len3(dir);
        double ov = d->diameter - len;
        double m = fabs(len - 0.00001);
        if (m < minm) minm = m;
        if (ov > 0.0) {
          if (len <= 0.00001) continue;
          double s = -0.5 * ov * (1.0 - d->stiffness) / len;
          push[0] += s * dir[0];
          push[1] += s * dir[1];
          push[2] += s * dir[2];
          /* friction with a solid particle (Macklin et al. 2014, eq. 24-25): the
           * tangential part t of the relative displacement in the step is removed
           * when |t| < mu_s d (static), else shortened by mu_k d (kinetic); equal
           * masses take half each.  Applied to p* only. */
          if (d->mu_s > 0.0 && (solid_at(w, i) || solid_at(w, j))) {
            double nrm[3] = {-dir[0] / len, -dir[1] / len, -dir[2] / len}, u[3], t[3];
            for (int c = 0; c < 3; ++c)
              u[c] = (P[4 * i + c] - X[4 * i + c]) - (P[4 * j + c] - X[4 * j + c]);
            const double un = dot3(u, nrm);
            for (int c = 0; c < 3; ++c) t[c] = u[c] - un * nrm[c];
            const double tl = len3(t);
            double f = 1.0;
            if (!(tl < d->mu_s * ov)) f = fmin(d->mu_k * ov / tl, 1.0);
            for (int c = 0; c < 3; ++c) fr[c] -= 0.5 * f * t[c];
          }
        }
      }
      tmp[3 * i + 0] = push[0];
      tmp[3 * i + 1] = push[1];
      tmp[3 * i + 2] = push[2];
      fric[3 * i + 0] = fr[0];
      fric[3 * i + 1] = fr[1];
      fric[3 * i + 2] = fr[2];
    }
    free(nb);
  }
  for (uint32_t i = 0; i < w->n; ++i)
    for (int c = 0; c < 3; ++c) {
      P[4 * i + c] += tmp[3 * i + c] + fric[3 * i + c];
      X[4 * i + c] += tmp[3 * i + c];
    }
  return minm;
}

/* collisionType 1: contact batches (Collision.h:291-464, Fluid.cu:62-71).
 *
 * The contacts are the rows of the neighbour table: the list orc_neighbours
 * returns for each particle (27-cell contacts, then shells, capped), contact
 * c = i * cap + k joining i and its k-th entry.  One stabilization iteration
 * clears the solved flags (resetContactConstraintSuccess) and runs
 * maxCollisionBatches batches.  In a batch every unsolved contact claims its two
 * particles (setupCollisionConstraintBatches); a contact holding both is marked
 * solved and, if the particles overlap (len < diameter, len > 1e-5), pushes them
 * apart by half the overlap times (1 - stiffness) each, in p* and in x
 * (setupCollisionConstraintBatchesCheck).
 *
 * The reference serialises the claims in index order within a thread block and
 * lets blocks race for a particle.  Deterministic definition used here: a
 * contact wins a particle when its index is the smallest among the unsolved
 * contacts at that particle (the in-block order applied across blocks).  The
 * contacts winning both particles form a matching, so their pushes are
 * independent and applied in place. */
static void pass_collide_batches(const orc_work* w, double* P, double* X, const uint32_t* nbl,
                                 const uint32_t* nbc, uint8_t* solved, uint64_t* minc) {
  const orc_derived* d = w->d;
  const uint64_t cap = d->cap;
  const uint32_t n = w->n;
  memset(solved, 0, (size_t)n * cap);
  for (uint32_t b = 0; b < d->max_batches; ++b) {
    for (uint32_t i = 0; i < n; ++i) minc[i] = UINT64_MAX;
    for (uint32_t i = 0; i < n; ++i)
      for (uint32_t k = 0; k < nbc[i]; ++k) {
        const uint64_t c = (uint64_t)i * cap + k;
        if (solved[c]) continue;
        const uint32_t j = nbl[c];
        if (c < minc[i]) minc[i] = c;
        if (c < minc[j]) minc[j] = c;
      }
    for (uint32_t i = 0; i < n; ++i) {
      const uint64_t c = minc[i];
      if (c == UINT64_MAX || c / cap != i) continue;
      const uint32_t j = nbl[c];
      if (minc[j] != c) continue;
      solved[c] = 1;
      double dir[3];
      sub3(&P[4 * j], &P[4 * i], dir);
      const double len = len3(dir);
      const double half = (d->diameter - len) / 2.0;
      if (half > 0.0 && len > 0.00001) {
        const double s = half * (1.0 - d->stiffness) / len;
        for (int a = 0; a < 3; ++a) {
          P[4 * i + a] -= s * dir[a];
          X[4 * i + a] -= s * dir[a];
          P[4 * j + a] += s * dir[a];
          X[4 * j + a] += s * dir[a];
        }
        P[4 * i + 3] = 0.0; /* the pushed p* is written back with w = 0 (Collision.h:403,408,428-429) */
        P[4 * j + 3] = 0.0;
      }
    }
  }
}

/* computeOmega (Density.h:232-269). Textbook: omega_i = sum v_ij x grad_j W_ij. */
static void pass_omega(const orc_work* w, const double* P, const double* V, double* Om) {
  const orc_derived* d = w->d;
#pragma omp parallel
  {
    uint32_t* nb = (uint32_t*)malloc(d->cap * sizeof(uint32_t));
#pragma omp for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)w->n; ++ii) {
      uint32_t i = (uint32_t)ii, nc;
      uint32_t cnt = orc_neighbours(d, w->n, w->keys, w->starts, w->ends, i, nb, &nc);
      double om[3] = {0, 0, 0};
      for (uint32_t k = 0; k < cnt; ++k) {
        uint32_t j = nb[k];
        if (solid_at(w, i) || solid_at(w, j)) continue;
        double r[3], g[3], vij[3], cr[3];
        sub3(&P[4 * i], &P[4 * j], r);
        sub3(&V[4 * j], &V[4 * i], vij);
        orc_grad(d, r, g);
        if (d->variant == PBD_KERNEL_TEXTBOOK) {
          g[0] = -g[0];
          g[1] = -g[1];
          g[2] = -g[2];
        }
        cross3(vij, g, cr);
        om[0] += cr[0];
        om[1] += cr[1];
        om[2] += cr[2];
      }
      const double vol = (d->variant == PBD_KERNEL_TEXTBOOK) ? 1.0 / d->rho0 : 1.0;
      Om[3 * i + 0] = vol * om[0];
      Om[3 * i + 1] = vol * om[1];
      Om[3 * i + 2] = vol * om[2];
    }
    free(nb);
  }
}

/* computeVorticity (Density.h:179-221).  Reference: component-wise
 * sum |w_j - w_i| / (p_j - p_i + 1e-8), N = (g + 1e-8)/(|g| + 1e-8).
 * Textbook: eta = sum (|w_j| - |w_i|) grad_i W_ij, N = eta/|eta| (0 if eta = 0).
 * f = e_vort (N x w_i), f.w = 0. */
static void pass_vorticity(const orc_work* w, const double* P, const double* Om, double* F) {
  const orc_derived* d = w->d;
#pragma omp parallel
  {
    uint32_t* nb = (uint32_t*)malloc(d->cap * sizeof(uint32_t));
#pragma omp for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)w->n; ++ii) {
      uint32_t i = (uint32_t)ii, nc;
      uint32_t cnt = orc_neighbours(d, w->n, w->keys, w->starts, w->ends, i, nb, &nc);
      const double* oi = &Om[3 * i];
      double N[3];
      if (d->variant == PBD_KERNEL_REFERENCE) {
        const double EPS = 0.00000001;
        double gs[3] = {0, 0, 0};
        for (uint32_t k = 0; k < cnt; ++k) {
          uint32_t j = nb[k];
          if (solid_at(w, i) || solid_at(w, j)) continue;
          double dom[3];
          sub3(&Om[3 * j], oi, dom);
          double ol = len3(dom);
          gs[0] += ol / (P[4 * j + 0] - P[4 * i + 0] + EPS);
          gs[1] += ol / (P[4 * j + 1] - P[4 * i + 1] + EPS);
          gs[2] += ol / (P[4 * j + 2] - P[4 * i + 2] + EPS);
        }
        double inv = 1.0 / (len3(gs) + EPS);
        N[0] = inv * (gs[0] + EPS);
        N[1] = inv * (gs[1] + EPS);
        N[2] = inv * (gs[2] + EPS);
      } else {
        double eta[3] = {0, 0, 0};
        double wi = len3(oi);
        for (uint32_t k = 0; k < cnt; ++k) {
          uint32_t j = nb[k];
          if (solid_at(w, i) || solid_at(w, j)) continue;
          double r[3], g[3];
          sub3(&P[4 * i], &P[4 * j], r);
          orc_grad(d, r, g);
          double s = len3(&Om[3 * j]) - wi;
          eta[0] += s * g[0];
          eta[1] += s * g[1];
          eta[2] += s * g[2];
        }
        double el = len3(eta);
        if (el > 0.0) {
          N[0] = eta[0] / el;
          N[1] = eta[1] / el;
          N[2] = eta[2] / el;
        } else {
          N[0] = N[1] = N[2] = 0.0;
        }
      }
      double f[3];
      cross3(N, oi, f);
      F[4 * i + 0] = d->e_vort * f[0];
      F[4 * i + 1] = d->e_vort * f[1];
      F[4 * i + 2] = d->e_vort * f[2];
      F[4 * i + 3] = 0.0;
    }
    free(nb);
  }
}

/* computeViscosity (Density.h:280-310), Jacobi: v_i += c sum (v_j - v_i) W_ij
 * over all four components. */
static void pass_xsph(const orc_work* w, const double* P, const double* V, double* Vn) {
  const orc_derived* d = w->d;
#pragma omp parallel
  {
    uint32_t* nb = (uint32_t*)malloc(d->cap * sizeof(uint32_t));
#pragma omp for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)w->n; ++ii) {
      uint32_t i = (uint32_t)ii, nc;
      uint32_t cnt = orc_neighbours(d, w->n, w->keys, w->starts, w->ends, i, nb, &nc);
      double s[4] = {0, 0, 0, 0};
      for (uint32_t k = 0; k < cnt; ++k) {
        uint32_t j = nb[k];
        if (solid_at(w, i) || solid_at(w, j)) continue;
        double r[3];
        sub3(&P[4 * i], &P[4 * j], r);
        double W = orc_w(d, len3(r));
        for (int c = 0; c < 4; ++c) s[c] += (V[4 * j + c] - V[4 * i + c]) * W;
      }
      const double vol = (d->variant == PBD_KERNEL_TEXTBOOK) ? 1.0 / d->rho0 : 1.0;
      for (int c = 0; c < 4; ++c) Vn[4 * i + c] = V[4 * i + c] + d->c_visc * vol * s[c];
    }
    free(nb);
  }
}

static void count_pass(const orc_work* w, uint32_t* cnt_out, uint32_t* cc_out) {
  const orc_derived* d = w->d;
#pragma omp parallel
  {
    uint32_t* nb = (uint32_t*)malloc(d->cap * sizeof(uint32_t));
#pragma omp for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)w->n; ++ii) {
      uint32_t nc;
      uint32_t cnt = orc_neighbours(d, w->n, w->keys, w->starts, w->ends, (uint32_t)ii, nb, &nc);
      cnt_out[ii] = cnt;
      cc_out[ii] = nc;
    }
    free(nb);
  }
}

int orc_step(const pbd_params* prm, orc_io* io) {
  orc_derived D;
  int rc = orc_derive(prm, &D);
  if (rc != PBD_OK) return rc;
  const uint32_t n = io->n;
  io->min_cutoff_margin = DBL_MAX;
  if (n == 0) return PBD_OK; /* Fluid.cu:19 */
  if (D.collision_type != 0 && D.collision_type != 1) return PBD_ERR_INVALID;

  /* 1. applyForces (Fluid.cu:21-23) */
  if (D.g_apply_forces) orc_predict(&D, n, io->x4, io->v4, io->f4, io->p4);

  /* 2. initializeCellIds, 3. sortIds (Fluid.cu:25-27) */
  uint32_t* keys0 = (uint32_t*)malloc((size_t)n * 4);
  uint32_t* keys = (uint32_t*)malloc((size_t)n * 4);
  uint32_t* ids = (uint32_t*)malloc((size_t)n * 4);
  uint32_t* starts = (uint32_t*)malloc((size_t)D.table_size * 4);
  uint32_t* ends = (uint32_t*)malloc((size_t)D.table_size * 4);
  double* X = (double*)malloc((size_t)n * 32);
  double* P = (double*)malloc((size_t)n * 32);
  double* V = (double*)malloc((size_t)n * 32);
  double* Vn = (double*)malloc((size_t)n * 32);
  double* F = (double*)malloc((size_t)n * 32);
  double* Om = (double*)calloc((size_t)n * 3, sizeof(double));
  double* tmp = (double*)malloc((size_t)n * 24);
  double* fric = (double*)calloc((size_t)n * 3, sizeof(double));
  uint32_t* phs = io->phase ? (uint32_t*)malloc((size_t)n * 4) : NULL;
  float* c4s = (float*)malloc((size_t)n * 16);
  uint32_t* cnt = (uint32_t*)malloc((size_t)n * 4);
  uint32_t* ccnt = (uint32_t*)malloc((size_t)n * 4);
  double* dens = (double*)calloc(n, sizeof(double));
  uint32_t *nbl = NULL, *nbc = NULL; /* collisionType 1: contact table, built on first use */
  uint8_t* solved = NULL;
  uint64_t* minc = NULL;
  if (!keys0 || !keys || !ids || !starts || !ends || !X || !P || !V || !Vn || !F || !Om || !tmp ||
      !c4s || !cnt || !ccnt || !dens || !fric || (io->phase && !phs)) {
    rc = PBD_ERR_NOMEM;
    goto out;
  }
  for (uint32_t i = 0; i < n; ++i) keys0[i] = orc_cell_key(&D, &io->p4[4 * i], NULL);
  if (io->out_keys_unsorted) memcpy(io->out_keys_unsorted, keys0, (size_t)n * 4);
  orc_sort(n, keys0, keys, ids);

  /* 4. reorderStorage (SortReorder.h:53-129): positions, predicted, velocities, colours */
  for (uint32_t s = 0; s < n; ++s) {
    uint32_t j = ids[s];
    for (int c = 0; c < 4; ++c) {
      X[4 * s + c] = io->x4[4 * j + c];
      P[4 * s + c] = io->p4[4 * j + c];
      V[4 * s + c] = io->v4[4 * j + c];
      c4s[4 * s + c] = io->c4[4 * j + c];
    }
    if (phs) phs[s] = io->phase[j];
  }
  /* externalForces, lambdas and deltaPositions are indexed by slot and not reordered */
  for (uint32_t s = 0; s < n; ++s)
    for (int c = 0; c < 4; ++c) F[4 * s + c] = io->f4[4 * s + c];

  /* 5. resetCellInfo + computeCellInfo (Fluid.cu:31-33) */
  orc_cell_ranges(n, keys, D.table_size, starts, ends);
  orc_work W = {&D, n, keys, starts, ends, phs};

  /* 6. findContacts / findNeighbours (Fluid.cu:41-47): counts and colours */
  count_pass(&W, cnt, ccnt);
  if (D.g_find_neighbours) {
    for (uint32_t s = 0; s < n; ++s) {
      volatile float q = (float)cnt[s] / (float)D.cap;
      c4s[4 * s + 0] = q;
      c4s[4 * s + 1] = q;
      c4s[4 * s + 2] = q;
      c4s[4 * s + 3] = 1.0f;
    }
  }

  /* 7. solver loop (Fluid.cu:49-80) */
  for (uint32_t it = 0; it < D.solver_iters; ++it) {
    if (D.g_lambda) {
      double m = pass_lambda(&W, P, io->lambda, it == 0 ? dens : NULL);
      if (m < io->min_cutoff_margin) io->min_cutoff_margin = m;
    }
    if (D.g_delta) pass_delta(&W, P, io->lambda, io->dp4);
    for (uint32_t sj = 0; sj < D.stab_iters; ++sj) {
      if (D.collision_type == 0) {
        double m = pass_collide(&W, P, X, tmp, fric);
        if (m < io->min_cutoff_margin) io->min_cutoff_margin = m;
      } else {
        if (!nbl) { /* the contact table of the step (cells are fixed within it) */
          nbl = (uint32_t*)malloc((size_t)n * D.cap * 4);
          nbc = (uint32_t*)malloc((size_t)n * 4);
          solved = (uint8_t*)malloc((size_t)n * D.cap);
          minc = (uint64_t*)malloc((size_t)n * 8);
          if (!nbl || !nbc || !solved || !minc) {
            rc = PBD_ERR_NOMEM;
            goto out;
          }
          for (uint32_t s = 0; s < n; ++s) {
            uint32_t nc;
            nbc[s] = orc_neighbours(&D, n, keys, starts, ends, s, &nbl[(size_t)s * D.cap], &nc);
          }
        }
        pass_collide_batches(&W, P, X, nbl, nbc, solved, minc);
      }
    }
    if (D.g_apply_delta) {
      if (D.variant == PBD_KERNEL_REFERENCE) {
        for (uint32_t s = 0; s < n; ++s)
          for (int c = 0; c < 4; ++c) P[4 * s + c] -= io->dp4[4 * s + c];
      } else {
        /* textbook: p += dp, then project onto the enclosure */
        for (uint32_t s = 0; s < n; ++s) {
          for (int c = 0; c < 4; ++c) P[4 * s + c] += io->dp4[4 * s + c];
          for (int a = 0; a < 3; ++a) {
            double lo = (double)(D.bmin_f[a] + 0.001f), hi = (double)(D.bmax_f[a] - 0.001f);
            if (P[4 * s + a] < lo) P[4 * s + a] = lo;
            else if (P[4 * s + a] > hi) P[4 * s + a] = hi;
          }
        }
      }
    }
  }

  /* 8. updatePositions (Fluid.cu:82-84) */
  if (D.g_update) {
    for (uint32_t s = 0; s < n; ++s)
      for (int c = 0; c < 4; ++c) {
        V[4 * s + c] = (P[4 * s + c] - X[4 * s + c]) / D.dt;
        X[4 * s + c] = P[4 * s + c];
      }
  }
  /* 9. computeOmega, 10. computeVorticity, 11. computeViscosity (Fluid.cu:86-96) */
  if (D.g_omega) pass_omega(&W, P, V, Om);
  if (D.g_vorticity) pass_vorticity(&W, P, Om, F);
  if (D.g_viscosity) {
    pass_xsph(&W, P, V, Vn);
    memcpy(V, Vn, (size_t)n * 32);
  }

  /* outputs and FP32 state */
  for (uint32_t s = 0; s < n; ++s) {
    for (int c = 0; c < 4; ++c) {
      io->x4[4 * s + c] = (float)X[4 * s + c];
      io->p4[4 * s + c] = (float)P[4 * s + c];
      io->v4[4 * s + c] = (float)V[4 * s + c];
      io->f4[4 * s + c] = (float)F[4 * s + c];
      io->c4[4 * s + c] = c4s[4 * s + c];
    }
  }
  if (phs) memcpy(io->phase, phs, (size_t)n * 4);
  if (io->out_x4) memcpy(io->out_x4, X, (size_t)n * 32);
  if (io->out_p4) memcpy(io->out_p4, P, (size_t)n * 32);
  if (io->out_v4) memcpy(io->out_v4, V, (size_t)n * 32);
  if (io->out_f4) memcpy(io->out_f4, F, (size_t)n * 32);
  if (io->out_density) memcpy(io->out_density, dens, (size_t)n * 8);
  if (io->out_omega3) memcpy(io->out_omega3, Om, (size_t)n * 24);
  if (io->out_keys) memcpy(io->out_keys, keys, (size_t)n * 4);
  if (io->out_ids) memcpy(io->out_ids, ids, (size_t)n * 4);
  if (io->out_starts) memcpy(io->out_starts, starts, (size_t)D.table_size * 4);
  if (io->out_ends) memcpy(io->out_ends, ends, (size_t)D.table_size * 4);
  if (io->out_nbr_count) memcpy(io->out_nbr_count, cnt, (size_t)n * 4);
  if (io->out_contact_count) memcpy(io->out_contact_count, ccnt, (size_t)n * 4);
  rc = PBD_OK;
out:
  free(keys0);
  free(keys);
  free(ids);
  free(starts);
  free(ends);
  free(X);
  free(P);
  free(V);
  free(Vn);
  free(F);
  free(Om);
  free(tmp);
  free(fric);
  free(phs);
  free(c4s);
  free(cnt);
  free(ccnt);
  free(dens);
  free(nbl);
  free(nbc);
  free(solved);
  free(minc);
  return rc;
}
