// particle_ops.hpp - per-pair terms and per-particle epilogues of the
// neighbour passes, shared by the per-particle (global memory) kernels and
// the LDS cell-block tile kernels.
//
// Reference kernels (src/app/simulation/kernel/): computeLambda Density.h:48-93,
// computeDeltaPositions Density.h:104-145, solveCollisions Collision.h:171-242,
// applyDeltaPositions Density.h:156-168, updatePositions Update.h:11-32,
// computeOmega Density.h:232-269, computeVorticity Density.h:179-221,
// computeViscosity Density.h:280-310.
#pragma once

#include "device.hpp"
#include "kernels.hpp"

namespace pbd {

// ---------------------------------------------------------------- lambda
struct LambdaAcc {
  float A = 0.0f, B = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f;
};

template <int VAR>
__device__ __forceinline__ void lambda_pair(const DevParams& P, float rx, float ry, float rz, LambdaAcc& a) {
  const float r2 = rx * rx + ry * ry + rz * rz;
  a.A += w_unnorm<VAR>(P, r2);
  float hr4;
  const float g = grad_factor<VAR>(P, r2, hr4);
  a.gx += g * rx;
  a.gy += g * ry;
  a.gz += g * rz;
  a.B += hr4;
}

// lambda_i = -C_i / (sum_k |grad_k C_i|^2 + eps); W is the density of the pass.
// eps: the relaxation of the particle's scene (scene_eps)
template <int VAR>
__device__ __forceinline__ float lambda_finish(const DevParams& P, const LambdaAcc& a, float& W, float eps) {
  W = P.wc * a.A + P.w0;
  const float G2 = P.gc * P.gc * (a.gx * a.gx + a.gy * a.gy + a.gz * a.gz);
  const float B2 = P.gc * P.gc * a.B;
  if (VAR == PBD_KERNEL_TEXTBOOK) return -(W * P.inv_rho0 - 1.0f) / ((G2 + B2) * P.inv_rho0_sq + eps);
  return -(W * P.inv_rho0 - 1.0f) / (B2 * P.inv_rho0 + G2 * P.inv_rho0 + eps);
}

// ---------------------------------------------------------------- delta p + collision
struct DeltaAcc {
  float ax = 0.0f, ay = 0.0f, az = 0.0f;  // sum (l_i + l_j + s_corr) g r
  float px = 0.0f, py = 0.0f, pz = 0.0f;  // collision push
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;  // friction (applied to p* only)
};

// Position-based friction of one contact with a solid particle (Macklin et al.
// 2014, eq. 24-25; the reference has no friction).  r = p*_i - p*_j (len, 1/len),
// (dix, diy, diz) and dj = p* - x of i and j: their displacement in the step.
// The tangential part t of the relative displacement is removed when
// |t| < mu_s d (static) and shortened by mu_k d otherwise (kinetic), d = diameter
// - len the penetration; equal masses take half each.
__device__ __forceinline__ void friction_pair(const DevParams& P, float rx, float ry, float rz, float len, float il,
                                              float dix, float diy, float diz, float djx, float djy, float djz,
                                              DeltaAcc& a) {
  const float nx = rx * il, ny = ry * il, nz = rz * il;
  const float ux = dix - djx, uy = diy - djy, uz = diz - djz;
  const float un = ux * nx + uy * ny + uz * nz;
  const float tx = ux - un * nx, ty = uy - un * ny, tz = uz - un * nz;
  const float tl = sqrtf(tx * tx + ty * ty + tz * tz);
  const float d = P.diameter - len;
  float s = 1.0f;
  if (!(tl < P.mu_s * d)) s = fminf(P.mu_k * d / tl, 1.0f);  // tl > 0 here: mu_s d > 0
  a.fx -= 0.5f * s * tx;
  a.fy -= 0.5f * s * ty;
  a.fz -= 0.5f * s * tz;
}

// s_corr exponent: NS > 0 compile-time, NS == 0 the runtime nSCorr.
template <int NS>
__device__ __forceinline__ float scorr_pow(const DevParams& P, float x) {
  if (NS > 0) return ipow_c<(NS > 0 ? NS : 1)>(x);
  return ipow(x, P.n_scorr);
}

template <int VAR, bool DELTA, bool COLLIDE, int NS = 0>
__device__ __forceinline__ void delta_pair(const DevParams& P, float li, float lj, float rx, float ry, float rz,
                                           bool contact, DeltaAcc& a) {
  const float r2 = rx * rx + ry * ry + rz * rz;
  if (VAR == PBD_KERNEL_TEXTBOOK) {
    // one 1/r for the spiky gradient and the contact push
    const float r2c = fmaxf(r2, 1.0e-30f);
    const float ir = rsq(r2c);
    const float len = r2c * ir;
    if (DELTA) {
      const float t = fmaxf(P.hp2 - r2, 0.0f);
      const float sc = -P.k_scorr * scorr_pow<NS>(P, P.wc * (t * t * t) * P.inv_w_dq);
      const float hr = fmaxf(P.hs - len, 0.0f);
      const float c = (li + lj + sc) * (hr * hr * ir);
      a.ax += c * rx;
      a.ay += c * ry;
      a.az += c * rz;
    }
    if (COLLIDE && contact) {
      // push = coll_scale * ov * (p_j - p_i)/len = -coll_scale * ov / len * r, for ov > 0, len > 1e-5
      const float ov = fmaxf(P.diameter - len, 0.0f);
      const float t = len > 0.00001f ? -P.coll_scale * ov * ir : 0.0f;
      a.px += t * rx;
      a.py += t * ry;
      a.pz += t * rz;
    }
    return;
  }
  if (DELTA) {
    const float w = P.wc * w_unnorm<VAR>(P, r2);
    const float sc = -P.k_scorr * ipow(w * P.inv_w_dq, P.n_scorr);
    float hr4;
    const float g = grad_factor<VAR>(P, r2, hr4);
    const float t = (li + lj + sc) * g;
    a.ax += t * rx;
    a.ay += t * ry;
    a.az += t * rz;
  }
  if (COLLIDE && contact && r2 > 0.0f) {
    const float il = rsq(r2);
    const float len = r2 * il;
    const float ov = P.diameter - len;
    if (ov > 0.0f && len > 0.00001f) {
      // push = coll_scale * ov * (p_j - p_i)/len = -coll_scale * ov / len * r
      const float t = -P.coll_scale * ov * il;
      a.px += t * rx;
      a.py += t * ry;
      a.pz += t * rz;
    }
  }
}

// Apply the pass to particle s (cell `key`, corner c): Qo/QXo <- (q, qx) + push -+ dp
// (+ projection onto the scene's enclosure), update.  Returns the new position (cell frame).
template <int VAR, bool DELTA>
__device__ __forceinline__ float4 delta_finish(const DevParams& P, uint32_t flags, uint32_t s, float4 qi,
                                             const float4* __restrict__ QX, const DeltaAcc& a, float cx, float cy,
                                             float cz, uint32_t key, float4* __restrict__ Qo,
                                             float4* __restrict__ QXo, float4* __restrict__ DP,
                                             float4* __restrict__ VS) {
  float dx, dy, dz;
  if (DELTA) {
    const float sc = P.gc * P.inv_rho0;
    dx = a.ax * sc;
    dy = a.ay * sc;
    dz = a.az * sc;
    if (flags & DF_WRITE_DP) DP[s] = make_float4(dx, dy, dz, 0.0f);
  } else {
    const float4 d = (flags & DF_APPLY) ? DP[s] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    dx = d.x;
    dy = d.y;
    dz = d.z;
  }
  const float4 qxi = QX[s];
  float nx = qi.x + a.px + a.fx, ny = qi.y + a.py + a.fy, nz = qi.z + a.pz + a.fz;
  float4 qx = make_float4(qxi.x + a.px, qxi.y + a.py, qxi.z + a.pz, qxi.w);
  if (flags & DF_APPLY) {
    if (VAR == PBD_KERNEL_REFERENCE) {
      nx -= dx;
      ny -= dy;
      nz -= dz;
    } else {
      nx += dx;
      ny += dy;
      nz += dz;
      const ProjBox b = scene_box(P, key);
      nx = fminf(fmaxf(nx, b.lo[0] - cx), b.hi[0] - cx);
      ny = fminf(fmaxf(ny, b.lo[1] - cy), b.hi[1] - cy);
      nz = fminf(fmaxf(nz, b.lo[2] - cz), b.hi[2] - cz);
    }
  }
  if (flags & DF_UPDATE) {
    VS[s] = make_float4((nx - qx.x) / P.dt, (ny - qx.y) / P.dt, (nz - qx.z) / P.dt, (qi.w - qx.w) / P.dt);
    qx = make_float4(nx, ny, nz, qi.w);
  }
  const float4 q = make_float4(nx, ny, nz, qi.w);
  Qo[s] = q;
  QXo[s] = qx;
  return q;
}

// ---------------------------------------------------------------- omega
struct OmegaAcc {
  float ox = 0.0f, oy = 0.0f, oz = 0.0f;
};

template <int VAR>
__device__ __forceinline__ void omega_pair(const DevParams& P, float rx, float ry, float rz, float vx, float vy,
                                           float vz, OmegaAcc& a) {
  const float r2 = rx * rx + ry * ry + rz * rz;
  float hr4;
  const float g = grad_factor<VAR>(P, r2, hr4);
  const float gx = g * rx, gy = g * ry, gz = g * rz;
  a.ox += vy * gz - vz * gy;
  a.oy += vz * gx - vx * gz;
  a.oz += vx * gy - vy * gx;
}

// reference omega_i = sum v_ij x grad(p_i - p_j); textbook (1/rho0) sum v_ij x grad_j W_ij
template <int VAR>
__device__ __forceinline__ float4 omega_finish(const DevParams& P, const OmegaAcc& a) {
  const float sc = (VAR == PBD_KERNEL_TEXTBOOK) ? -P.gc * P.vol : P.gc;
  const float ox = a.ox * sc, oy = a.oy * sc, oz = a.oz * sc;
  return make_float4(ox, oy, oz, sqrtf(ox * ox + oy * oy + oz * oz));  // w = |omega|
}

// ---------------------------------------------------------------- vorticity + XSPH + output
struct FinalAcc {
  float sx = 0.0f, sy = 0.0f, sz = 0.0f, sw = 0.0f;  // XSPH sum
  float ex = 0.0f, ey = 0.0f, ez = 0.0f;             // vorticity gradient
};

// Textbook vorticity: eta = sum (|w_j| - |w_i|) g r (gc applied in the finish).
template <int VAR>
__device__ __forceinline__ void final_pair_textbook(const DevParams& P, bool vort, bool xsph, float rx, float ry,
                                                    float rz, float4 vi, float4 vj, float oil, float ojl,
                                                    FinalAcc& a) {
  const float r2 = rx * rx + ry * ry + rz * rz;
  if (xsph) {
    const float w = w_unnorm<VAR>(P, r2);
    a.sx += (vj.x - vi.x) * w;
    a.sy += (vj.y - vi.y) * w;
    a.sz += (vj.z - vi.z) * w;
    a.sw += (vj.w - vi.w) * w;
  }
  if (vort) {
    float hr4;
    const float g = grad_factor<VAR>(P, r2, hr4);
    const float d = (ojl - oil) * g;
    a.ex += d * rx;
    a.ey += d * ry;
    a.ez += d * rz;
  }
}

// Reference vorticity (Density.h:207-213): sum |w_j - w_i| / (p_j - p_i + 1e-8) per component.
__device__ __forceinline__ void final_pair_reference(const DevParams& P, bool vort, bool xsph, float rx, float ry,
                                                     float rz, float4 vi, float4 vj, float4 oi, float4 oj,
                                                     FinalAcc& a) {
  const float r2 = rx * rx + ry * ry + rz * rz;
  if (xsph) {
    const float w = w_unnorm<PBD_KERNEL_REFERENCE>(P, r2);
    a.sx += (vj.x - vi.x) * w;
    a.sy += (vj.y - vi.y) * w;
    a.sz += (vj.z - vi.z) * w;
    a.sw += (vj.w - vi.w) * w;
  }
  if (vort) {
    const float dx = oj.x - oi.x, dy = oj.y - oi.y, dz = oj.z - oi.z;
    const float ol = sqrtf(dx * dx + dy * dy + dz * dz);
    a.ex += ol / (-rx + 0.00000001f);
    a.ey += ol / (-ry + 0.00000001f);
    a.ez += ol / (-rz + 0.00000001f);
  }
}

template <int VAR>
__device__ __forceinline__ void final_finish(const DevParams& P, uint32_t flags, uint32_t s, float4 qi, float4 vi,
                                             float4 oi, const FinalAcc& a, float cx, float cy, float cz,
                                             const float4* __restrict__ QX, const float4* __restrict__ CS,
                                             const uint32_t* __restrict__ ncnt, float4* __restrict__ Xo,
                                             float4* __restrict__ Po, float4* __restrict__ Vo,
                                             float4* __restrict__ Co, float4* __restrict__ Fo) {
  if (flags & FF_VORT) {
    float nx, ny, nz;
    if (VAR == PBD_KERNEL_REFERENCE) {
      const float EPS = 0.00000001f;
      const float inv = 1.0f / (sqrtf(a.ex * a.ex + a.ey * a.ey + a.ez * a.ez) + EPS);
      nx = inv * (a.ex + EPS);
      ny = inv * (a.ey + EPS);
      nz = inv * (a.ez + EPS);
    } else {
      // eta carries the factor gc; its direction is sign(gc) * eta
      const float sg = P.gc < 0.0f ? -1.0f : 1.0f;
      const float el = sqrtf(a.ex * a.ex + a.ey * a.ey + a.ez * a.ez);
      const float inv = el > 0.0f ? sg / el : 0.0f;
      nx = a.ex * inv;
      ny = a.ey * inv;
      nz = a.ez * inv;
    }
    Fo[s] = make_float4(P.e_vort * (ny * oi.z - nz * oi.y), P.e_vort * (nz * oi.x - nx * oi.z),
                        P.e_vort * (nx * oi.y - ny * oi.x), 0.0f);
  }
  float4 v = vi;
  if (flags & FF_XSPH) {
    const float c = P.c_visc * P.wc * P.vol;
    v = make_float4(vi.x + c * a.sx, vi.y + c * a.sy, vi.z + c * a.sz, vi.w + c * a.sw);
  }
  const float4 xl = (flags & FF_UPDATED) ? qi : QX[s];
  Xo[s] = make_float4(xl.x + cx, xl.y + cy, xl.z + cz, xl.w);
  Po[s] = make_float4(qi.x + cx, qi.y + cy, qi.z + cz, qi.w);
  if (!(flags & FF_XSPH_DONE)) Vo[s] = v;
  if (P.gates & G_FIND_NEIGHBOURS) {
    const float q = (float)ncnt[s] / (float)P.cap;
    Co[s] = make_float4(q, q, q, 1.0f);
  } else {
    Co[s] = CS[s];
  }
}

// ---------------------------------------------------------------- per-particle (global) bodies
// P.phs: phases in sorted order (nullptr: all fluid).  A solid particle has no
// density constraint (lambda 0) and a fluid one counts only fluid neighbours.
template <int VAR, bool FIRST>
__device__ __forceinline__ void lambda_particle(const DevParams& P, uint32_t s, const float4* __restrict__ Q,
                                                const uint32_t* __restrict__ keys, const uint32_t* __restrict__ cs,
                                                const uint32_t* __restrict__ ce, float* __restrict__ lam,
                                                float* __restrict__ dens, uint32_t* __restrict__ ncnt,
                                                uint32_t* __restrict__ ccnt) {
  const uint32_t* __restrict__ ph = P.phs;
  const float4 qi = Q[s];
  const CellFrame f = make_frame(P, keys[s]);
  LambdaAcc a;
  uint32_t nc = 0;
  const uint32_t cnt = visit_neighbours(P, s, f, qi.x, qi.y, qi.z, (P.gates & G_FIND_CONTACTS) != 0,
                                        (P.gates & G_FIND_NEIGHBOURS) != 0, cs, ce, nc,
                                        [&](uint32_t j, float bx, float by, float bz, bool) {
                                          if (ph && ph[j]) return;
                                          const float4 qj = Q[j];
                                          lambda_pair<VAR>(P, bx - qj.x, by - qj.y, bz - qj.z, a);
                                        });
  float W;
  lam[s] = lambda_finish<VAR>(P, a, W, scene_eps(P, keys[s]));
  if (ph && ph[s]) lam[s] = 0.0f;
  if (FIRST) {
    dens[s] = W;
    ncnt[s] = cnt;
    ccnt[s] = nc;
  }
}

// P.phs: phases (nullptr: all fluid).  Delta p acts between fluid particles only;
// the contact push acts on every pair and friction on contacts with a solid.
template <int VAR, bool DELTA, bool COLLIDE>
__device__ __forceinline__ float4 delta_particle(const DevParams& P, uint32_t flags, uint32_t s,
                                               const float4* __restrict__ Q, const float4* __restrict__ QX,
                                               const float* __restrict__ lam, const uint32_t* __restrict__ keys,
                                               const uint32_t* __restrict__ cs, const uint32_t* __restrict__ ce,
                                               float4* __restrict__ Qo, float4* __restrict__ QXo,
                                               float4* __restrict__ DP, float4* __restrict__ VS) {
  const uint32_t* __restrict__ ph = P.phs;
  const float4 qi = Q[s];
  const CellFrame f = make_frame(P, keys[s]);
  DeltaAcc a;
  if (DELTA || COLLIDE) {
    const float li = DELTA ? lam[s] : 0.0f;
    const bool si = ph && ph[s];
    const bool fric = COLLIDE && ph && P.mu_s > 0.0f;
    float dix = 0.0f, diy = 0.0f, diz = 0.0f;
    if (fric) {
      const float4 x = QX[s];
      dix = qi.x - x.x;
      diy = qi.y - x.y;
      diz = qi.z - x.z;
    }
    uint32_t nc = 0;
    visit_neighbours(P, s, f, qi.x, qi.y, qi.z, (P.gates & G_FIND_CONTACTS) != 0,
                     DELTA && (P.gates & G_FIND_NEIGHBOURS) != 0, cs, ce, nc,
                     [&](uint32_t j, float bx, float by, float bz, bool contact) {
                       const float4 qj = Q[j];
                       const float rx = bx - qj.x, ry = by - qj.y, rz = bz - qj.z;
                       const bool sj = ph && ph[j];
                       if (si || sj) {
                         delta_pair<VAR, false, COLLIDE>(P, 0.0f, 0.0f, rx, ry, rz, contact, a);
                         if (fric && contact) {
                           const float r2 = rx * rx + ry * ry + rz * rz;
                           if (r2 > 0.0f) {
                             const float il = rsq(r2), len = r2 * il;
                             if (len < P.diameter && len > 0.00001f) {
                               const float4 xj = QX[j];
                               friction_pair(P, rx, ry, rz, len, il, dix, diy, diz, qj.x - xj.x, qj.y - xj.y,
                                             qj.z - xj.z, a);
                             }
                           }
                         }
                       } else {
                         delta_pair<VAR, DELTA, COLLIDE>(P, li, DELTA ? lam[j] : 0.0f, rx, ry, rz, contact, a);
                       }
                     });
  }
  return delta_finish<VAR, DELTA>(P, flags, s, qi, QX, a, f.cx, f.cy, f.cz, keys[s], Qo, QXo, DP, VS);
}

// P.phs: phases (nullptr: all fluid); omega of a solid particle is 0, a fluid one
// sums over fluid neighbours.
template <int VAR>
__device__ __forceinline__ void omega_particle(const DevParams& P, uint32_t s, const float4* __restrict__ Q,
                                               const float4* __restrict__ VS, const uint32_t* __restrict__ keys,
                                               const uint32_t* __restrict__ cs, const uint32_t* __restrict__ ce,
                                               float4* __restrict__ OM) {
  const uint32_t* __restrict__ ph = P.phs;
  const float4 qi = Q[s], vi = VS[s];
  const CellFrame f = make_frame(P, keys[s]);
  OmegaAcc a;
  uint32_t nc = 0;
  const bool si = ph && ph[s];
  visit_neighbours(P, s, f, qi.x, qi.y, qi.z, (P.gates & G_FIND_CONTACTS) != 0, (P.gates & G_FIND_NEIGHBOURS) != 0, cs,
                   ce, nc, [&](uint32_t j, float bx, float by, float bz, bool) {
                     if (si || (ph && ph[j])) return;
                     const float4 qj = Q[j], vj = VS[j];
                     omega_pair<VAR>(P, bx - qj.x, by - qj.y, bz - qj.z, vj.x - vi.x, vj.y - vi.y, vj.z - vi.z, a);
                   });
  OM[s] = omega_finish<VAR>(P, a);
}

// XSPH alone for particle s, with final_particle's neighbour order, arithmetic
// and phase masks (P.phs: a solid particle gets none, a fluid one sums over
// fluid neighbours).
template <int VAR>
__device__ __forceinline__ float4 xsph_particle(const DevParams& P, uint32_t s, const float4* __restrict__ Q,
                                                const float4* __restrict__ VS, const uint32_t* __restrict__ keys,
                                                const uint32_t* __restrict__ cs, const uint32_t* __restrict__ ce) {
  const uint32_t* __restrict__ ph = P.phs;
  const float4 qi = Q[s], vi = VS[s];
  if (ph && ph[s]) return vi;
  const CellFrame f = make_frame(P, keys[s]);
  FinalAcc a;
  uint32_t nc = 0;
  visit_neighbours(P, s, f, qi.x, qi.y, qi.z, (P.gates & G_FIND_CONTACTS) != 0, (P.gates & G_FIND_NEIGHBOURS) != 0, cs,
                   ce, nc, [&](uint32_t j, float bx, float by, float bz, bool) {
                     if (ph && ph[j]) return;
                     const float4 qj = Q[j];
                     final_pair_textbook<VAR>(P, false, true, bx - qj.x, by - qj.y, bz - qj.z, vi, VS[j], 0.0f, 0.0f,
                                              a);
                   });
  const float c = P.c_visc * P.wc * P.vol;
  return make_float4(vi.x + c * a.sx, vi.y + c * a.sy, vi.z + c * a.sz, vi.w + c * a.sw);
}

template <int VAR>
__device__ __forceinline__ void final_particle(const DevParams& P, uint32_t flags, uint32_t s,
                                               const float4* __restrict__ Q, const float4* __restrict__ QX,
                                               const float4* __restrict__ VS, const float4* __restrict__ OM,
                                               const uint32_t* __restrict__ keys, const uint32_t* __restrict__ cs,
                                               const uint32_t* __restrict__ ce, const float4* __restrict__ CS,
                                               const uint32_t* __restrict__ ncnt, float4* __restrict__ Xo,
                                               float4* __restrict__ Po, float4* __restrict__ Vo,
                                               float4* __restrict__ Co, float4* __restrict__ Fo) {
  const uint32_t* __restrict__ ph = P.phs;
  const float4 qi = Q[s], vi = VS[s];
  const CellFrame f = make_frame(P, keys[s]);
  const bool vort = (flags & FF_VORT) != 0, xsph = (flags & FF_XSPH) != 0;
  const float4 oi = vort ? OM[s] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const bool si = ph && ph[s];  // solid: no vorticity force, no XSPH
  FinalAcc a;
  if (vort || xsph) {
    uint32_t nc = 0;
    visit_neighbours(P, s, f, qi.x, qi.y, qi.z, (P.gates & G_FIND_CONTACTS) != 0,
                     (P.gates & G_FIND_NEIGHBOURS) != 0, cs, ce, nc,
                     [&](uint32_t j, float bx, float by, float bz, bool) {
                       if (si || (ph && ph[j])) return;
                       const float4 qj = Q[j];
                       const float4 vj = xsph ? VS[j] : vi;
                       const float rx = bx - qj.x, ry = by - qj.y, rz = bz - qj.z;
                       if (VAR == PBD_KERNEL_REFERENCE) {
                         const float4 oj = vort ? OM[j] : oi;
                         final_pair_reference(P, vort, xsph, rx, ry, rz, vi, vj, oi, oj, a);
                       } else {
                         const float ojl = vort ? OM[j].w : 0.0f;
                         final_pair_textbook<VAR>(P, vort, xsph, rx, ry, rz, vi, vj, oi.w, ojl, a);
                       }
                     });
  }
  final_finish<VAR>(P, flags, s, qi, vi, oi, a, f.cx, f.cy, f.cz, QX, CS, ncnt, Xo, Po, Vo, Co, Fo);
}

}  // namespace pbd
