/*
 * pbf_adjoint.c - FP64 multi-step forward and reverse mode of the PBF step.
 *
 * TEST INFRASTRUCTURE ONLY (see pbf_oracle.h): the checker of the GPU
 * adjoint.  orc64_forward is the FP64 map that the central finite differences
 * evaluate; orc64_gradient is its hand-derived reverse mode, itself pinned by
 * those finite differences in the tests.
 *
 * The step is the textbook variant of Fluid::compute (src/app/simulation/
 * Fluid.cu:14-97) with collisionType 0 and at most one stabilization pass,
 * carried in FP64 from step to step (orc_step reads and writes FP32 state):
 *   predict   v += dt (g y + f); p = x + dt v; box clamp (Forces.h:19-101)
 *   sort      cell ids of (float) p, stable sort, reorder (SortReorder.h)
 *   iterate   lambda_i = -C_i / (|sum_j G_ij|^2 + sum_j |G_ij|^2)/rho0^2 + eps)
 *             dp_i = 1/rho0 sum_j (lambda_i + lambda_j + s_corr_ij) G_ij
 *             push_i = sum_{r_ij < d} 0.5 (1 - stiffness) (d - r)/r r_ij
 *             p <- clamp(p + push + dp), x <- x + push      (Density.h, Collision.h)
 *   update    u = (p - x)/dt, x = p                          (Update.h:11-32)
 *   omega     w_i = 1/rho0 sum_j G_ij x (u_j - u_i)          (Density.h:232-269)
 *   vorticity eta_i = sum_j (|w_j| - |w_i|) G_ij, f_i = e (eta/|eta|) x w_i
 *   XSPH      v_i = u_i + c/rho0 sum_j (u_j - u_i) W_ij      (Density.h:280-310)
 * with W the poly6 kernel (width h_poly), G_ij = grad_i W_spiky(x_i - x_j)
 * (width h_spiky) and s_corr = -k h_spiky^2 (W/W(dq))^n, as in pbf_oracle.c.
 *
 * Reverse mode.  Every neighbour sum is over a symmetric relation (no
 * neighbour cap reached), so each adjoint is a gather over the same
 * neighbours: for a pair term f(x_i - x_j) the two rows that contain it give
 * the i-side and the j-side adjoint together, e.g.
 *   d/dx_i sum_k a_k . sum_j G_kj = sum_j H_ij (a_i - a_j),  H = dG/dx_i (symmetric, even).
 * The passes per step, last stage first:
 *   A  XSPH (self-adjoint gather), eta and the vorticity-force adjoint
 *   B  |w| adjoint through eta, position terms of eta
 *   C  omega adjoint, then the update
 *   per iteration, last first:
 *   R1 lambda adjoint from dp, position terms of dp and push, then the
 *      per-particle adjoints of rho_i, a_i = sum_j G_ij and b_i = sum_j |G_ij|^2
 *   R2 position terms of rho, a and b
 *   predict, scattered back through the sort permutation.
 * Nonsmooth points (box clamps, contact onset, kernel supports, eta = 0) take
 * the derivative of the branch the forward took.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "pbf_oracle.h"

#define ORC_PI 3.14159265358979323846

typedef struct k64 {
  double hp, hp2, hs, wc, gc, w0, k, inv_wdq, diam, kc;
  int ns;
} k64;

static void k64_init(const orc_derived* d, k64* k) {
  k->hp = d->h_poly;
  k->hp2 = d->h_poly * d->h_poly;
  k->hs = d->h_spiky;
  k->wc = 315.0 / (64.0 * ORC_PI * pow(d->h_poly, 9));
  k->gc = -45.0 / (ORC_PI * pow(d->h_spiky, 6));
  k->w0 = k->wc * k->hp2 * k->hp2 * k->hp2;
  k->k = d->k_scorr * d->h_spiky * d->h_spiky;
  k->inv_wdq = d->w_dq != 0.0 ? 1.0 / d->w_dq : 0.0;
  k->ns = d->n_scorr;
  k->diam = d->diameter;
  k->kc = 0.5 * (1.0 - d->stiffness);
}

/* Pair quantities of r = x_i - x_j (true units). */
typedef struct pair64 {
  double r2, r;
  double W;       /* poly6 */
  double Wg[3];   /* grad_i W */
  double G[3];    /* grad_i W_spiky */
  double psi;     /* G = gc psi r_vec */
  double hq;      /* gc (h - r)(h + r)/r^3: H = gc psi I - hq r r^T */
} pair64;

static void pair_eval(const k64* k, const double* r, pair64* q) {
  q->r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  q->r = sqrt(q->r2);
  q->W = 0.0;
  q->Wg[0] = q->Wg[1] = q->Wg[2] = 0.0;
  if (q->r2 < k->hp2) {
    const double t = k->hp2 - q->r2;
    q->W = k->wc * t * t * t;
    const double s = -6.0 * k->wc * t * t;
    q->Wg[0] = s * r[0];
    q->Wg[1] = s * r[1];
    q->Wg[2] = s * r[2];
  }
  q->psi = 0.0;
  q->hq = 0.0;
  q->G[0] = q->G[1] = q->G[2] = 0.0;
  if (q->r > 0.0 && q->r < k->hs) {
    const double hr = k->hs - q->r;
    q->psi = hr * hr / q->r;
    q->hq = k->gc * hr * (k->hs + q->r) / (q->r2 * q->r);
    const double s = k->gc * q->psi;
    q->G[0] = s * r[0];
    q->G[1] = s * r[1];
    q->G[2] = s * r[2];
  }
}

/* out += H(r) v */
static void hmul_add(const k64* k, const pair64* q, const double* r, const double* v, double scale, double* out) {
  if (q->psi == 0.0) return;
  const double rv = r[0] * v[0] + r[1] * v[1] + r[2] * v[2];
  const double a = k->gc * q->psi, b = q->hq * rv;
  for (int c = 0; c < 3; ++c) out[c] += scale * (a * v[c] - b * r[c]);
}

static double scorr64(const k64* k, double W) {
  if (k->inv_wdq == 0.0) return 0.0;
  return -k->k * pow(W * k->inv_wdq, (double)k->ns);
}
static double dscorr64(const k64* k, double W) { /* d s_corr / dW */
  if (k->inv_wdq == 0.0 || k->ns == 0) return 0.0;
  return -k->k * (double)k->ns * pow(W * k->inv_wdq, (double)(k->ns - 1)) * k->inv_wdq;
}

static inline void sub3(const double* a, const double* b, double* r) {
  r[0] = a[0] - b[0];
  r[1] = a[1] - b[1];
  r[2] = a[2] - b[2];
}
static inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void cross3(const double* a, const double* b, double* r) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

/* One step's record for the reverse pass (sorted order of the step unless noted). */
typedef struct tape64 {
  uint32_t* perm;   /* slot -> index before the sort */
  uint32_t* nbo;    /* n + 1 offsets into nbl */
  uint32_t* nbl;    /* neighbour slots, reference order */
  uint32_t* ncon;   /* contacts: the first ncon[i] entries of the list */
  double* xin;      /* 3n state at the step's start, order before the sort */
  double* vin;
  double* fin;
  double* Q;        /* (iters + 1) x 3n positions at each iteration's start; last = final */
  double* lam;      /* iters x n */
  double* U;        /* 3n velocities after the update */
  double* OM;       /* 3n vorticity */
  uint8_t* sol;     /* n: 1 for a solid particle (NULL: all fluid) */
  double* Xq;       /* iters x 3n: x at each iteration's start (friction) */
} tape64;

static void tape_free(tape64* t) {
  free(t->perm);
  free(t->nbo);
  free(t->nbl);
  free(t->ncon);
  free(t->xin);
  free(t->vin);
  free(t->fin);
  free(t->Q);
  free(t->lam);
  free(t->U);
  free(t->OM);
  free(t->sol);
  free(t->Xq);
  memset(t, 0, sizeof(*t));
}

typedef struct ctx64 {
  orc_derived d;
  k64 k;
  double rho0, eps, lo[3], hi[3];
  double mu_s, mu_k; /* friction on contacts with a solid particle (0: none) */
  uint32_t n, iters, stab;
} ctx64;

static int ctx_init(const pbd_params* p, double rho0, double eps, uint32_t n, ctx64* c) {
  memset(c, 0, sizeof(*c));
  int rc = orc_derive(p, &c->d);
  if (rc != PBD_OK) return rc;
  if (c->d.variant != PBD_KERNEL_TEXTBOOK || c->d.collision_type != 0 || c->d.stab_iters > 1)
    return PBD_ERR_UNSUPPORTED;
  if (!c->d.g_apply_forces || !c->d.g_find_contacts || !c->d.g_lambda || !c->d.g_delta || !c->d.g_apply_delta ||
      !c->d.g_update || !c->d.g_omega || !c->d.g_vorticity || !c->d.g_viscosity)
    return PBD_ERR_UNSUPPORTED;
  k64_init(&c->d, &c->k);
  c->rho0 = rho0 > 0.0 ? rho0 : c->d.rho0;
  c->eps = eps >= 0.0 ? eps : c->d.eps;
  for (int a = 0; a < 3; ++a) {
    c->lo[a] = (double)(c->d.bmin_f[a] + 0.001f);
    c->hi[a] = (double)(c->d.bmax_f[a] - 0.001f);
  }
  c->mu_s = c->d.mu_s;
  c->mu_k = c->d.mu_k;
  c->n = n;
  c->iters = c->d.solver_iters;
  c->stab = c->d.stab_iters;
  return PBD_OK;
}

/* predict in FP64; returns the clamp flags (bits 0-2 per axis, bit 3 any). */
static unsigned predict64(const ctx64* c, const double* x, const double* v, const double* f, double* xo,
                          double* vo, double* po) {
  const orc_derived* d = &c->d;
  const double dt = d->dt;
  double vv[3] = {v[0] + dt * f[0], v[1] + dt * ((double)d->gravity_f + f[1]), v[2] + dt * f[2]};
  double p[3];
  unsigned fl = 0;
  for (int a = 0; a < 3; ++a) {
    p[a] = x[a] + dt * vv[a];
    if (p[a] < (double)d->bmin_f[a]) {
      vv[a] *= (double)d->vdamp_f;
      p[a] = (double)(d->bmin_f[a] + 0.001f);
      fl |= 1u << a;
    } else if (p[a] > (double)d->bmax_f[a]) {
      vv[a] *= (double)d->vdamp_f;
      p[a] = (double)(d->bmax_f[a] - 0.001f);
      fl |= 1u << a;
    }
  }
  for (int a = 0; a < 3; ++a) {
    xo[a] = fl ? x[a] + (double)d->pdamp_f * (p[a] - x[a]) : x[a];
    vo[a] = vv[a];
    po[a] = p[a];
  }
  return fl ? (fl | 8u) : 0u;
}

/* Neighbour lists (CSR) of the sorted step. */
static int build_lists(const ctx64* c, const uint32_t* keys, const uint32_t* st, const uint32_t* en, tape64* t) {
  const uint32_t n = c->n, cap = c->d.cap;
  t->nbo = (uint32_t*)malloc(((size_t)n + 1) * 4);
  t->ncon = (uint32_t*)malloc((size_t)n * 4);
  uint32_t* tmp = (uint32_t*)malloc((size_t)cap * 4);
  if (!t->nbo || !t->ncon || !tmp) {
    free(tmp);
    return PBD_ERR_NOMEM;
  }
  size_t total = 0;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t nc;
    uint32_t cnt = orc_neighbours(&c->d, n, keys, st, en, i, tmp, &nc);
    if (cnt >= cap) {
      free(tmp);
      return PBD_ERR_UNSUPPORTED; /* a capped list is not symmetric */
    }
    t->nbo[i] = (uint32_t)total;
    t->ncon[i] = nc;
    total += cnt;
  }
  t->nbo[n] = (uint32_t)total;
  t->nbl = (uint32_t*)malloc((total ? total : 1) * 4);
  if (!t->nbl) {
    free(tmp);
    return PBD_ERR_NOMEM;
  }
#pragma omp parallel
  {
    uint32_t* buf = (uint32_t*)malloc((size_t)cap * 4);
#pragma omp for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
      uint32_t nc;
      uint32_t cnt = orc_neighbours(&c->d, n, keys, st, en, (uint32_t)ii, buf, &nc);
      memcpy(&t->nbl[t->nbo[ii]], buf, (size_t)cnt * 4);
    }
    free(buf);
  }
  free(tmp);
  return PBD_OK;
}

/* Friction of one contact with a solid particle (Macklin et al. 2014, eq.
 * 24-25), as pass_collide in pbf_oracle.c: r = p_i - p_j (length len),
 * u = (p_i - x_i) - (p_j - x_j); adds -0.5 f t to out, t the tangential part
 * of u, f = 1 when |t| < mu_s d (d = diameter - len) else min(mu_k d / |t|, 1). */
static void fric64(const ctx64* c, const double* r, double len, const double* u, double* out) {
  const double ov = c->k.diam - len;
  double nr[3], t[3];
  for (int m = 0; m < 3; ++m) nr[m] = r[m] / len;
  const double un = dot3(u, nr);
  for (int m = 0; m < 3; ++m) t[m] = u[m] - un * nr[m];
  const double tl = sqrt(dot3(t, t));
  double f = 1.0;
  if (!(tl < c->mu_s * ov)) f = fmin(c->mu_k * ov / tl, 1.0);
  for (int m = 0; m < 3; ++m) out[m] -= 0.5 * f * t[m];
}

/* Adjoint of fric64's term F(r, u) in a pair term e . F: gr = J_r^T e and
 * gu = J_u^T e, the derivative of the branch the forward took.  With n = r/len,
 * un = u.n, t = u - un n, T = |t|, e_p = e - (e.n) n:
 *   f = 1:  F = -t/2,  gu = -e_p/2,  gr = ((n.e) t + un e_p) / (2 len);
 *   f = mu_k d/T:  F = -mu_k d t/(2T); with w = e - (e.t/T) t/T, w_p = w - (w.n) n,
 *           gu = -f w_p/2,  gr = mu_k (e.t/T) n/2 + f ((n.w) t + un w_p) / (2 len). */
static void fric64_adj(const ctx64* c, const double* r, double len, const double* u, const double* e, double* gr,
                       double* gu) {
  const double ov = c->k.diam - len;
  double nr[3], t[3], ep[3];
  for (int m = 0; m < 3; ++m) nr[m] = r[m] / len;
  const double un = dot3(u, nr), ne = dot3(nr, e);
  for (int m = 0; m < 3; ++m) {
    t[m] = u[m] - un * nr[m];
    ep[m] = e[m] - ne * nr[m];
  }
  const double tl = sqrt(dot3(t, t));
  if (tl < c->mu_s * ov || c->mu_k * ov >= tl) {
    for (int m = 0; m < 3; ++m) {
      gu[m] = -0.5 * ep[m];
      gr[m] = 0.5 / len * (ne * t[m] + un * ep[m]);
    }
    return;
  }
  const double f = c->mu_k * ov / tl;
  double th[3], w[3];
  for (int m = 0; m < 3; ++m) th[m] = t[m] / tl;
  const double te = dot3(th, e);
  for (int m = 0; m < 3; ++m) w[m] = e[m] - te * th[m];
  const double nw = dot3(nr, w);
  for (int m = 0; m < 3; ++m) {
    const double wp = w[m] - nw * nr[m];
    gu[m] = -0.5 * f * wp;
    gr[m] = 0.5 * c->mu_k * te * nr[m] + 0.5 * f / len * (nw * t[m] + un * wp);
  }
}

/* One FP64 step.  X, V, F: 3n state in/out (order before the sort on entry,
 * sorted on exit); orig and phase (NULL: all fluid) carried through the sort.
 * t: record (NULL = none). */
static int step64(const ctx64* c, double* X, double* V, double* F, uint32_t* orig, uint32_t* phase, tape64* t) {
  const uint32_t n = c->n, iters = c->iters;
  const k64* k = &c->k;
  const double dt = c->d.dt, rho0 = c->rho0, eps = c->eps;
  int rc = PBD_OK;
  tape64 own;
  memset(&own, 0, sizeof(own));
  tape64* T = t ? t : &own;
  double* Xp = (double*)malloc((size_t)n * 24);
  double* Pp = (double*)malloc((size_t)n * 24);
  double* Xs = (double*)malloc((size_t)n * 24);
  double* Ps = (double*)malloc((size_t)n * 24);
  double* lam = (double*)malloc((size_t)n * 8);
  double* dp = (double*)malloc((size_t)n * 24);
  double* pu = (double*)malloc((size_t)n * 24);
  double* om = (double*)malloc((size_t)n * 24);
  uint32_t* keys0 = (uint32_t*)malloc((size_t)n * 4);
  uint32_t* keys = (uint32_t*)malloc((size_t)n * 4);
  uint32_t* perm = (uint32_t*)malloc((size_t)n * 4);
  uint32_t* st = (uint32_t*)malloc((size_t)c->d.table_size * 4);
  uint32_t* en = (uint32_t*)malloc((size_t)c->d.table_size * 4);
  uint32_t* ob = (uint32_t*)malloc((size_t)n * 4);
  double* fr3 = (double*)calloc((size_t)n * 3, 8);
  uint8_t* sol = phase ? (uint8_t*)malloc((size_t)n) : NULL;
  if (!Xp || !Pp || !Xs || !Ps || !lam || !dp || !pu || !om || !keys0 || !keys || !perm || !st || !en || !ob ||
      !fr3 || (phase && !sol)) {
    rc = PBD_ERR_NOMEM;
    goto out;
  }
  memset(keys0, 0, (size_t)n * 4);
  if (t) {
    t->xin = (double*)malloc((size_t)n * 24);
    t->vin = (double*)malloc((size_t)n * 24);
    t->fin = (double*)malloc((size_t)n * 24);
    t->Q = (double*)malloc((size_t)(iters + 1) * n * 24);
    t->lam = (double*)malloc((size_t)(iters ? iters : 1) * n * 8);
    t->U = (double*)malloc((size_t)n * 24);
    t->OM = (double*)malloc((size_t)n * 24);
    t->Xq = (double*)malloc((size_t)(iters ? iters : 1) * n * 24);
    if (!t->xin || !t->vin || !t->fin || !t->Q || !t->lam || !t->U || !t->OM || !t->Xq) {
      rc = PBD_ERR_NOMEM;
      goto out;
    }
    memcpy(t->xin, X, (size_t)n * 24);
    memcpy(t->vin, V, (size_t)n * 24);
    memcpy(t->fin, F, (size_t)n * 24);
  }
  /* predict, cell ids */
  for (uint32_t i = 0; i < n; ++i) {
    double vo[3];
    predict64(c, &X[3 * i], &V[3 * i], &F[3 * i], &Xp[3 * i], vo, &Pp[3 * i]);
    float pf[3] = {(float)Pp[3 * i], (float)Pp[3 * i + 1], (float)Pp[3 * i + 2]};
    keys0[i] = orc_cell_key(&c->d, pf, NULL);
  }
  orc_sort(n, keys0, keys, perm);
  for (uint32_t s = 0; s < n; ++s) {
    const uint32_t j = perm[s];
    memcpy(&Xs[3 * s], &Xp[3 * j], 24);
    memcpy(&Ps[3 * s], &Pp[3 * j], 24);
    ob[s] = orig[j];
  }
  memcpy(orig, ob, (size_t)n * 4);
  if (phase) {
    for (uint32_t s = 0; s < n; ++s) ob[s] = phase[perm[s]];
    memcpy(phase, ob, (size_t)n * 4);
    for (uint32_t s = 0; s < n; ++s) sol[s] = phase[s] == PBD_PHASE_SOLID;
  }
  orc_cell_ranges(n, keys, c->d.table_size, st, en);
  rc = build_lists(c, keys, st, en, T);
  if (rc != PBD_OK) goto out;
  const uint32_t* nbo = T->nbo;
  const uint32_t* nbl = T->nbl;
  const uint32_t* ncon = T->ncon;

  for (uint32_t it = 0; it < iters; ++it) {
    if (t) {
      memcpy(&t->Q[(size_t)it * 3 * n], Ps, (size_t)n * 24);
      memcpy(&t->Xq[(size_t)it * 3 * n], Xs, (size_t)n * 24);
    }
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
      const uint32_t i = (uint32_t)ii;
      double A = k->w0, a[3] = {0, 0, 0}, B = 0.0;
      for (uint32_t e = nbo[i]; e < nbo[i + 1]; ++e) {
        const uint32_t j = nbl[e];
        if (sol && sol[j]) continue; /* fluid neighbours only */
        double r[3];
        pair64 q;
        sub3(&Ps[3 * i], &Ps[3 * j], r);
        pair_eval(k, r, &q);
        A += q.W;
        for (int m = 0; m < 3; ++m) a[m] += q.G[m];
        B += dot3(q.G, q.G);
      }
      lam[i] = (sol && sol[i]) ? 0.0 : -(A / rho0 - 1.0) / ((dot3(a, a) + B) / (rho0 * rho0) + eps);
    }
    if (t) memcpy(&t->lam[(size_t)it * n], lam, (size_t)n * 8);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
      const uint32_t i = (uint32_t)ii;
      const int si = sol && sol[i];
      double acc[3] = {0, 0, 0}, push[3] = {0, 0, 0}, fr[3] = {0, 0, 0};
      for (uint32_t e = nbo[i]; e < nbo[i + 1]; ++e) {
        const uint32_t j = nbl[e];
        const int sj = sol && sol[j];
        double r[3];
        pair64 q;
        sub3(&Ps[3 * i], &Ps[3 * j], r);
        pair_eval(k, r, &q);
        if (!si && !sj) { /* delta p between fluid particles only */
          const double s = lam[i] + lam[j] + scorr64(k, q.W);
          for (int m = 0; m < 3; ++m) acc[m] += s * q.G[m];
        }
        if (c->stab && e - nbo[i] < ncon[i] && q.r < k->diam && q.r > 0.00001) {
          const double f = k->kc * (k->diam - q.r) / q.r;
          for (int m = 0; m < 3; ++m) push[m] += f * r[m];
          if (c->mu_s > 0.0 && (si || sj)) {
            double u[3];
            for (int m = 0; m < 3; ++m) u[m] = (Ps[3 * i + m] - Xs[3 * i + m]) - (Ps[3 * j + m] - Xs[3 * j + m]);
            fric64(c, r, q.r, u, fr);
          }
        }
      }
      for (int m = 0; m < 3; ++m) {
        dp[3 * i + m] = acc[m] / rho0;
        pu[3 * i + m] = push[m];
        fr3[3 * i + m] = fr[m];
      }
    }
    for (uint32_t i = 0; i < n; ++i)
      for (int m = 0; m < 3; ++m) {
        double p = Ps[3 * i + m] + pu[3 * i + m] + fr3[3 * i + m] + dp[3 * i + m];
        if (p < c->lo[m]) p = c->lo[m];
        else if (p > c->hi[m]) p = c->hi[m];
        Ps[3 * i + m] = p;
        Xs[3 * i + m] += pu[3 * i + m];
      }
  }
  if (t) memcpy(&t->Q[(size_t)iters * 3 * n], Ps, (size_t)n * 24);
  /* update: u = (p - x)/dt, x = p */
  for (uint32_t i = 0; i < n; ++i)
    for (int m = 0; m < 3; ++m) {
      V[3 * i + m] = (Ps[3 * i + m] - Xs[3 * i + m]) / dt;
      X[3 * i + m] = Ps[3 * i + m];
    }
  if (t) memcpy(t->U, V, (size_t)n * 24);
  /* omega */
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
    const uint32_t i = (uint32_t)ii;
    double w[3] = {0, 0, 0};
    for (uint32_t e = nbo[i]; e < nbo[i + 1]; ++e) {
      const uint32_t j = nbl[e];
      if (sol && (sol[i] || sol[j])) continue; /* fluid pairs only */
      double r[3], u[3], cr[3];
      pair64 q;
      sub3(&X[3 * i], &X[3 * j], r);
      pair_eval(k, r, &q);
      sub3(&V[3 * j], &V[3 * i], u);
      cross3(q.G, u, cr);
      for (int m = 0; m < 3; ++m) w[m] += cr[m];
    }
    for (int m = 0; m < 3; ++m) om[3 * i + m] = w[m] / rho0;
  }
  if (t) memcpy(t->OM, om, (size_t)n * 24);
  /* vorticity force and XSPH (Jacobi) */
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
    const uint32_t i = (uint32_t)ii;
    const double mi = sqrt(dot3(&om[3 * i], &om[3 * i]));
    double eta[3] = {0, 0, 0}, xs[3] = {0, 0, 0};
    for (uint32_t e = nbo[i]; e < nbo[i + 1]; ++e) {
      const uint32_t j = nbl[e];
      if (sol && (sol[i] || sol[j])) continue; /* fluid pairs only */
      double r[3];
      pair64 q;
      sub3(&X[3 * i], &X[3 * j], r);
      pair_eval(k, r, &q);
      const double s = sqrt(dot3(&om[3 * j], &om[3 * j])) - mi;
      for (int m = 0; m < 3; ++m) {
        eta[m] += s * q.G[m];
        xs[m] += (V[3 * j + m] - V[3 * i + m]) * q.W;
      }
    }
    const double el = sqrt(dot3(eta, eta));
    double N[3] = {0, 0, 0}, f[3];
    if (el > 0.0)
      for (int m = 0; m < 3; ++m) N[m] = eta[m] / el;
    cross3(N, &om[3 * i], f);
    for (int m = 0; m < 3; ++m) {
      F[3 * i + m] = c->d.e_vort * f[m];
      Ps[3 * i + m] = V[3 * i + m] + c->d.c_visc / rho0 * xs[m]; /* new velocity, staged */
    }
  }
  memcpy(V, Ps, (size_t)n * 24);
  if (t) {
    t->perm = perm;
    perm = NULL;
    t->sol = sol;
    sol = NULL;
  }
out:
  if (!t) tape_free(&own);
  free(Xp);
  free(Pp);
  free(Xs);
  free(Ps);
  free(lam);
  free(dp);
  free(pu);
  free(om);
  free(keys0);
  free(keys);
  free(perm);
  free(st);
  free(en);
  free(ob);
  free(fr3);
  free(sol);
  return rc;
}

/* Reverse of one step.  gX, gV, gF: 3n adjoints of the step's outputs (sorted
 * order) in, adjoints of its inputs (order before the sort) out. */
static int reverse64(const ctx64* c, const tape64* t, double* gX, double* gV, double* gF, double* grho0,
                     double* geps) {
  const uint32_t n = c->n, iters = c->iters;
  const k64* k = &c->k;
  const double dt = c->d.dt, rho0 = c->rho0, eps = c->eps, ev = c->d.e_vort, cv = c->d.c_visc / rho0;
  const uint32_t *nbo = t->nbo, *nbl = t->nbl, *ncon = t->ncon;
  const double* X = &t->Q[(size_t)iters * 3 * n];
  const double *U = t->U, *OM = t->OM;
  const uint8_t* sol = t->sol; /* phases: fluid sums over fluid pairs only; friction on solid contacts */
  double* Ub = (double*)calloc((size_t)n * 3, 8);
  double* Pb = (double*)calloc((size_t)n * 3, 8);
  double* Pn = (double*)calloc((size_t)n * 3, 8);
  double* Xb = (double*)calloc((size_t)n * 3, 8);
  double* Eb = (double*)calloc((size_t)n * 3, 8);
  double* Ob = (double*)calloc((size_t)n * 3, 8);
  double* DPb = (double*)calloc((size_t)n * 3, 8);
  double* PUb = (double*)calloc((size_t)n * 3, 8);
  double* Ab = (double*)calloc((size_t)n * 3, 8);
  double* Bb = (double*)calloc((size_t)n, 8);
  double* Rb = (double*)calloc((size_t)n, 8);
  double* gr = (double*)calloc((size_t)n, 8); /* per-particle d L / d rho0 terms */
  double* ge = (double*)calloc((size_t)n, 8); /* per-particle d L / d eps terms */
  double* XF = (double*)calloc((size_t)n * 3, 8); /* friction's x adjoint of one iteration */
  int rc = PBD_OK;
  if (!Ub || !Pb || !Pn || !Xb || !Eb || !Ob || !DPb || !PUb || !Ab || !Bb || !Rb || !gr || !ge || !XF) {
    rc = PBD_ERR_NOMEM;
    goto out;
  }
  /* A: XSPH, eta, vorticity force */
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
    const uint32_t i = (uint32_t)ii;
    const double* gvi = &gV[3 * i];
    const double mi = sqrt(dot3(&OM[3 * i], &OM[3 * i]));
    double ub[3] = {gvi[0], gvi[1], gvi[2]}, xb[3] = {0, 0, 0}, eta[3] = {0, 0, 0}, xs[3] = {0, 0, 0};
    for (uint32_t e = nbo[i]; e < nbo[i + 1]; ++e) {
      const uint32_t j = nbl[e];
      if (sol && (sol[i] || sol[j])) continue;
      double r[3], du[3], dg[3];
      pair64 q;
      sub3(&X[3 * i], &X[3 * j], r);
      pair_eval(k, r, &q);
      sub3(&U[3 * j], &U[3 * i], du);
      sub3(&gV[3 * j], gvi, dg);
      for (int m = 0; m < 3; ++m) {
        ub[m] += cv * q.W * dg[m];
        xs[m] += du[m] * q.W;
      }
      const double s = -cv * dot3(du, dg); /* (u_j - u_i).(gv_i - gv_j) */
      for (int m = 0; m < 3; ++m) xb[m] += s * q.Wg[m];
      const double sm = sqrt(dot3(&OM[3 * j], &OM[3 * j])) - mi;
      for (int m = 0; m < 3; ++m) eta[m] += sm * q.G[m];
    }
    gr[i] += -cv * dot3(gvi, xs) / rho0;
    const double el = sqrt(dot3(eta, eta));
    double N[3] = {0, 0, 0}, ob[3], nb[3], eb[3] = {0, 0, 0};
    if (el > 0.0)
      for (int m = 0; m < 3; ++m) N[m] = eta[m] / el;
    cross3(&gF[3 * i], N, ob);
    cross3(&OM[3 * i], &gF[3 * i], nb);
    for (int m = 0; m < 3; ++m) {
      ob[m] *= ev;
      nb[m] *= ev;
    }
    if (el > 0.0) {
      const double nn = dot3(N, nb);
      for (int m = 0; m < 3; ++m) eb[m] = (nb[m] - N[m] * nn) / el;
    }
    for (int m = 0; m < 3; ++m) {
      Ub[3 * i + m] = ub[m];
      Xb[3 * i + m] = xb[m];
      Eb[3 * i + m] = eb[m];
      Ob[3 * i + m] = ob[m];
    }
  }
  /* B: |omega| through eta */
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
    const uint32_t i = (uint32_t)ii;
    const double mi = sqrt(dot3(&OM[3 * i], &OM[3 * i]));
    double mb = 0.0, xb[3] = {0, 0, 0};
    for (uint32_t e = nbo[i]; e < nbo[i + 1]; ++e) {
      const uint32_t j = nbl[e];
      if (sol && (sol[i] || sol[j])) continue;
      double r[3], es[3];
      pair64 q;
      sub3(&X[3 * i], &X[3 * j], r);
      pair_eval(k, r, &q);
      for (int m = 0; m < 3; ++m) es[m] = Eb[3 * i + m] + Eb[3 * j + m];
      mb -= dot3(es, q.G);
      const double sm = sqrt(dot3(&OM[3 * j], &OM[3 * j])) - mi;
      hmul_add(k, &q, r, es, sm, xb);
    }
    for (int m = 0; m < 3; ++m) {
      Xb[3 * i + m] += xb[m];
      if (mi > 0.0) Pn[3 * i + m] = Ob[3 * i + m] + mb * OM[3 * i + m] / mi; /* complete omega adjoint */
      else Pn[3 * i + m] = Ob[3 * i + m];
    }
  }
  memcpy(Ob, Pn, (size_t)n * 24);
  /* C: omega, then the update */
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
    const uint32_t i = (uint32_t)ii;
    double ub[3] = {0, 0, 0}, xb[3] = {0, 0, 0};
    for (uint32_t e = nbo[i]; e < nbo[i + 1]; ++e) {
      const uint32_t j = nbl[e];
      if (sol && (sol[i] || sol[j])) continue;
      double r[3], os[3], cr[3], du[3];
      pair64 q;
      sub3(&X[3 * i], &X[3 * j], r);
      pair_eval(k, r, &q);
      for (int m = 0; m < 3; ++m) os[m] = Ob[3 * i + m] + Ob[3 * j + m];
      cross3(os, q.G, cr);
      for (int m = 0; m < 3; ++m) ub[m] -= cr[m] / rho0;
      sub3(&U[3 * j], &U[3 * i], du);
      cross3(du, os, cr);
      hmul_add(k, &q, r, cr, 1.0 / rho0, xb);
    }
    gr[i] += -dot3(&Ob[3 * i], &OM[3 * i]) / rho0;
    for (int m = 0; m < 3; ++m) {
      const double u = Ub[3 * i + m] + ub[m];
      Pb[3 * i + m] = gX[3 * i + m] + Xb[3 * i + m] + xb[m] + u / dt;
      Xb[3 * i + m] = -u / dt; /* x before the update, constant through the iterations */
    }
  }
  /* the solver iterations, last first */
  for (uint32_t it1 = iters; it1-- > 0;) {
    const double* Q = &t->Q[(size_t)it1 * 3 * n];
    const double* Qn = &t->Q[(size_t)(it1 + 1) * 3 * n];
    const double* lam = &t->lam[(size_t)it1 * n];
    const double* Xq = &t->Xq[(size_t)it1 * 3 * n];
    for (uint32_t i = 0; i < n; ++i)
      for (int m = 0; m < 3; ++m) {
        const double q = Qn[3 * i + m];
        const double b = (q == c->lo[m] || q == c->hi[m]) ? 0.0 : Pb[3 * i + m];
        DPb[3 * i + m] = b;
        PUb[3 * i + m] = c->stab ? b + Xb[3 * i + m] : 0.0;
      }
    /* R1 */
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
      const uint32_t i = (uint32_t)ii;
      double A = k->w0, a[3] = {0, 0, 0}, B = 0.0, lb = 0.0, pn[3], dpi[3] = {0, 0, 0}, xf[3] = {0, 0, 0};
      for (int m = 0; m < 3; ++m) pn[m] = DPb[3 * i + m];
      const int si = sol && sol[i];
      for (uint32_t e = nbo[i]; e < nbo[i + 1]; ++e) {
        const uint32_t j = nbl[e];
        const int sj = sol && sol[j];
        double r[3], dd[3];
        pair64 q;
        sub3(&Q[3 * i], &Q[3 * j], r);
        pair_eval(k, r, &q);
        const int contact = c->stab && e - nbo[i] < ncon[i] && q.r < k->diam && q.r > 0.00001;
        if (contact && c->mu_s > 0.0 && (si || sj)) {
          double u[3], ef[3], gr3[3], gu3[3];
          for (int m = 0; m < 3; ++m) {
            u[m] = (Q[3 * i + m] - Xq[3 * i + m]) - (Q[3 * j + m] - Xq[3 * j + m]);
            ef[m] = DPb[3 * i + m] - DPb[3 * j + m];
          }
          fric64_adj(c, r, q.r, u, ef, gr3, gu3);
          for (int m = 0; m < 3; ++m) {
            pn[m] += gr3[m] + gu3[m];
            xf[m] -= gu3[m];
          }
        }
        if (contact) {
          double pd[3];
          sub3(&PUb[3 * i], &PUb[3 * j], pd);
          const double phi = (k->diam - q.r) / q.r;
          const double rr = dot3(r, pd) * k->diam / (q.r2 * q.r);
          for (int m = 0; m < 3; ++m) pn[m] += k->kc * (phi * pd[m] - rr * r[m]);
        }
        if (si || sj) continue; /* density terms between fluid particles only */
        A += q.W;
        for (int m = 0; m < 3; ++m) a[m] += q.G[m];
        B += dot3(q.G, q.G);
        sub3(&DPb[3 * i], &DPb[3 * j], dd);
        const double gd = dot3(q.G, dd);
        lb += gd / rho0;
        const double S = lam[i] + lam[j] + scorr64(k, q.W);
        for (int m = 0; m < 3; ++m) dpi[m] += S * q.G[m] / rho0;
        hmul_add(k, &q, r, dd, S / rho0, pn);
        const double ds = dscorr64(k, q.W) * gd / rho0;
        for (int m = 0; m < 3; ++m) pn[m] += ds * q.Wg[m];
      }
      gr[i] += -dot3(&DPb[3 * i], dpi) / rho0;
      const double s2 = (dot3(a, a) + B) / (rho0 * rho0);
      const double D = s2 + eps, Cc = A / rho0 - 1.0;
      const double cb = -lb / D;
      const double db = lb * Cc / (D * D);
      ge[i] += db;
      gr[i] += cb * (-A / (rho0 * rho0)) + db * (-2.0 * s2 / rho0);
      Rb[i] = cb / rho0;
      for (int m = 0; m < 3; ++m) Ab[3 * i + m] = 2.0 * db * a[m] / (rho0 * rho0);
      Bb[i] = db / (rho0 * rho0);
      for (int m = 0; m < 3; ++m) {
        Pn[3 * i + m] = pn[m];
        XF[3 * i + m] = xf[m];
      }
    }
    /* friction reads x of this iteration: its adjoint joins x's before it */
    if (sol && c->mu_s > 0.0)
      for (size_t m = 0; m < (size_t)n * 3; ++m) Xb[m] += XF[m];
    /* R2 */
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
      const uint32_t i = (uint32_t)ii;
      double pn[3] = {0, 0, 0};
      for (uint32_t e = nbo[i]; e < nbo[i + 1]; ++e) {
        const uint32_t j = nbl[e];
        if (sol && (sol[i] || sol[j])) continue;
        double r[3], da[3];
        pair64 q;
        sub3(&Q[3 * i], &Q[3 * j], r);
        pair_eval(k, r, &q);
        const double rs = Rb[i] + Rb[j];
        for (int m = 0; m < 3; ++m) pn[m] += rs * q.Wg[m];
        sub3(&Ab[3 * i], &Ab[3 * j], da);
        hmul_add(k, &q, r, da, 1.0, pn);
        hmul_add(k, &q, r, q.G, 2.0 * (Bb[i] + Bb[j]), pn);
      }
      for (int m = 0; m < 3; ++m) Pb[3 * i + m] = Pn[3 * i + m] + pn[m];
    }
  }
  /* predict, back through the sort */
  for (uint32_t s = 0; s < n; ++s) {
    const uint32_t j = t->perm[s];
    double xo[3], vo[3], po[3];
    const unsigned fl = predict64(c, &t->xin[3 * j], &t->vin[3 * j], &t->fin[3 * j], xo, vo, po);
    const double pd = (fl & 8u) ? (double)c->d.pdamp_f : 0.0;
    for (int m = 0; m < 3; ++m) {
      const double xn = Xb[3 * s + m];
      const double pt = (fl & (1u << m)) ? 0.0 : Pb[3 * s + m] + pd * xn;
      gX[3 * j + m] = (1.0 - pd) * xn + pt;
      gV[3 * j + m] = dt * pt;
      gF[3 * j + m] = dt * dt * pt;
    }
  }
  {
    double sr = 0.0, se = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
      sr += gr[i];
      se += ge[i];
    }
    *grho0 += sr;
    *geps += se;
  }
out:
  free(Ub);
  free(Pb);
  free(Pn);
  free(Xb);
  free(Eb);
  free(Ob);
  free(DPb);
  free(PUb);
  free(Ab);
  free(Bb);
  free(Rb);
  free(gr);
  free(ge);
  free(XF);
  return rc;
}

static void to3(uint32_t n, const double* a4, double* a3) {
  for (uint32_t i = 0; i < n; ++i)
    for (int m = 0; m < 3; ++m) a3[3 * i + m] = a4[4 * i + m];
}
static void to4(uint32_t n, const double* a3, double* a4) {
  for (uint32_t i = 0; i < n; ++i) {
    for (int m = 0; m < 3; ++m) a4[4 * i + m] = a3[3 * i + m];
    a4[4 * i + 3] = 0.0;
  }
}

int orc64_forward(const pbd_params* p, double rho0, double eps, uint32_t n, uint32_t nsub, double* x4, double* v4,
                  double* f4, uint32_t* orig, uint32_t* phase) {
  ctx64 c;
  int rc = ctx_init(p, rho0, eps, n, &c);
  if (rc != PBD_OK || n == 0) return rc;
  double* X = (double*)malloc((size_t)n * 24);
  double* V = (double*)malloc((size_t)n * 24);
  double* F = (double*)malloc((size_t)n * 24);
  if (!X || !V || !F) {
    rc = PBD_ERR_NOMEM;
    goto out;
  }
  to3(n, x4, X);
  to3(n, v4, V);
  to3(n, f4, F);
  for (uint32_t s = 0; s < nsub && rc == PBD_OK; ++s) rc = step64(&c, X, V, F, orig, phase, NULL);
  if (rc == PBD_OK) {
    to4(n, X, x4);
    to4(n, V, v4);
    to4(n, F, f4);
  }
out:
  free(X);
  free(V);
  free(F);
  return rc;
}

int orc64_gradient(const pbd_params* p, double rho0, double eps, uint32_t n, uint32_t nsub, const double* x4,
                   const double* v4, const double* f4, const uint32_t* phase0, const double* gx4, const double* gv4,
                   double* gx0, double* gv0, double* grho0, double* geps) {
  ctx64 c;
  int rc = ctx_init(p, rho0, eps, n, &c);
  if (rc != PBD_OK) return rc;
  *grho0 = 0.0;
  *geps = 0.0;
  if (n == 0) return PBD_OK;
  tape64* tp = (tape64*)calloc(nsub ? nsub : 1, sizeof(tape64));
  double* X = (double*)malloc((size_t)n * 24);
  double* V = (double*)malloc((size_t)n * 24);
  double* F = (double*)malloc((size_t)n * 24);
  double* gX = (double*)malloc((size_t)n * 24);
  double* gV = (double*)malloc((size_t)n * 24);
  double* gF = (double*)calloc((size_t)n * 3, 8);
  uint32_t* orig = (uint32_t*)malloc((size_t)n * 4);
  uint32_t* phase = phase0 ? (uint32_t*)malloc((size_t)n * 4) : NULL;
  if (!tp || !X || !V || !F || !gX || !gV || !gF || !orig || (phase0 && !phase)) {
    rc = PBD_ERR_NOMEM;
    goto out;
  }
  to3(n, x4, X);
  to3(n, v4, V);
  to3(n, f4, F);
  for (uint32_t i = 0; i < n; ++i) orig[i] = i;
  if (phase) memcpy(phase, phase0, (size_t)n * 4);
  for (uint32_t s = 0; s < nsub && rc == PBD_OK; ++s) rc = step64(&c, X, V, F, orig, phase, &tp[s]);
  if (rc != PBD_OK) goto out;
  to3(n, gx4, gX);
  to3(n, gv4, gV);
  for (uint32_t s = nsub; s-- > 0 && rc == PBD_OK;) rc = reverse64(&c, &tp[s], gX, gV, gF, grho0, geps);
  if (rc == PBD_OK) {
    to4(n, gX, gx0);
    to4(n, gV, gv0);
  }
out:
  if (tp)
    for (uint32_t s = 0; s < nsub; ++s) tape_free(&tp[s]);
  free(tp);
  free(X);
  free(V);
  free(F);
  free(gX);
  free(gV);
  free(gF);
  free(orig);
  free(phase);
  return rc;
}
