/*
 * smoke_oracle.c -- CPU restatement of the reference smoke step
 * (LCEBurton/python-game-engine, src/cfd/backends/numba/...).
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker for the HIP path and the CPU
 * baseline timed by bench.py.  Nothing in python-game-engine_amd/ links it.
 *
 * Built with -O2 -fno-fast-math -ffp-contract=off so every expression is
 * evaluated in the order the reference writes it, with no FMA contraction.
 * Loops are parallel over rows exactly where the reference uses prange, so
 * results do not depend on the thread count (every kernel is a pure map, or a
 * colour pass that reads only the other colour).
 */
#include "smoke_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX(j, i, w) ((size_t)(j) * (size_t)(w) + (size_t)(i))

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ---------------------------------------------------------------------- */
/* advection.py:6-19 clamp / lerp; :21-49 sample_scalar_field             */
/* ---------------------------------------------------------------------- */
static inline long clamp_index(double f, long hi) {
  /* int(math.floor(x)) then clamp(x0, 0, n-1).  Clamping before the
   * conversion keeps out-of-range doubles defined; NaN maps to 0. */
  if (!(f > 0.0)) return 0;
  if (f > (double)hi) return hi;
  return (long)f;
}

double orc_sample(const double* f, int fh, int fw, double x, double y) {
  long x0 = clamp_index(floor(x), fw - 1);
  long x1 = clamp_index(ceil(x), fw - 1);
  long y0 = clamp_index(floor(y), fh - 1);
  long y1 = clamp_index(ceil(y), fh - 1);
  double tx = x - (double)x0;
  double ty = y - (double)y0;
  double c00 = f[IDX(y0, x0, fw)];
  double c10 = f[IDX(y0, x1, fw)];
  double c01 = f[IDX(y1, x0, fw)];
  double c11 = f[IDX(y1, x1, fw)];
  double c0 = c00 + (c10 - c00) * tx;
  double c1 = c01 + (c11 - c01) * tx;
  return c0 + (c1 - c0) * ty;
}

/* advection.py:97-110: staggered offsets */
static inline double sample_u(const double* u, int uh, int uw, double x, double y) {
  return orc_sample(u, uh, uw, x, y - 0.5);
}
static inline double sample_v(const double* v, int vh, int vw, double x, double y) {
  return orc_sample(v, vh, vw, x - 0.5, y);
}
static inline double sample_c(const double* c, int h, int w, double x, double y) {
  return orc_sample(c, h, w, x - 0.5, y - 0.5);
}

/* advection.py:112-128 */
void orc_advect_scalar(const double* src, double* dst, int H, int W,
                       const double* u, int uh, int uw,
                       const double* v, int vh, int vw,
                       const uint8_t* is_fluid, double dt, double cell_size) {
  const double inv = 1.0 / cell_size;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < H; ++j) {
    for (int i = 0; i < W; ++i) {
      if (!is_fluid[IDX(j, i, W)]) { dst[IDX(j, i, W)] = 0.0; continue; }
      double px = i + 0.5, py = j + 0.5;
      double uu = sample_u(u, uh, uw, px, py);
      double vv = sample_v(v, vh, vw, px, py);
      double x = px - (uu * dt * inv);
      double y = py - (vv * dt * inv);
      dst[IDX(j, i, W)] = sample_c(src, H, W, x, y);
    }
  }
}

/* advection.py:130-154 */
void orc_advect_u(const double* u_src, double* u_dst, int uh, int uw,
                  const double* v, int vh, int vw,
                  const uint8_t* fmu, double dt, double cell_size) {
  const double inv = 1.0 / cell_size;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < uh; ++j) {
    for (int i = 0; i < uw; ++i) {
      if (!fmu[IDX(j, i, uw)]) { u_dst[IDX(j, i, uw)] = 0.0; continue; }
      double px = (double)i, py = j + 0.5;
      double uu = u_src[IDX(j, i, uw)];
      double vv = sample_v(v, vh, vw, px, py);
      double x = px - (uu * dt * inv);
      double y = py - (vv * dt * inv);
      u_dst[IDX(j, i, uw)] = sample_u(u_src, uh, uw, x, y);
    }
  }
}

/* advection.py:156-180 */
void orc_advect_v(const double* v_src, double* v_dst, int vh, int vw,
                  const double* u, int uh, int uw,
                  const uint8_t* fmv, double dt, double cell_size) {
  const double inv = 1.0 / cell_size;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < vh; ++j) {
    for (int i = 0; i < vw; ++i) {
      if (!fmv[IDX(j, i, vw)]) { v_dst[IDX(j, i, vw)] = 0.0; continue; }
      double px = i + 0.5, py = (double)j;
      double uu = sample_u(u, uh, uw, px, py);
      double vv = v_src[IDX(j, i, vw)];
      double x = px - (uu * dt * inv);
      double y = py - (vv * dt * inv);
      v_dst[IDX(j, i, vw)] = sample_v(v_src, vh, vw, x, y);
    }
  }
}

/* ---------------------------------------------------------------------- */
/* forces.py                                                              */
/* ---------------------------------------------------------------------- */
/* forces.py:19-46: float32 buffer of shape (H+1, W), row H stays 0 */
static inline float buoyancy_at(const double* rho, int H, int W, int j, int i,
                                double dc, double g) {
  if (j >= H) return 0.0f;
  double d;
  if (j == 0 || j == H - 1) d = rho[IDX(j, i, W)];
  else d = 0.5 * (rho[IDX(j - 1, i, W)] + rho[IDX(j, i, W)]);
  return (float)(-dc * d * g);
}
/* forces.py:48-76: float32 buffer of shape (Ht+1, Wt) */
static inline float temperature_at(const double* T, int Ht, int Wt, int j,
                                   int i, double Ta, double bc) {
  double t;
  if (T == NULL) t = 0.0;
  else if (j == 0) t = T[IDX(0, i, Wt)];
  else if (j == Ht) t = T[IDX(Ht - 1, i, Wt)];
  else t = 0.5 * (T[IDX(j - 1, i, Wt)] + T[IDX(j, i, Wt)]);
  return (float)(-bc * (t - Ta));
}

/* forces.py:78-105 */
void orc_apply_forces(double* v, int vh, int vw,
                      const double* density, int H, int W,
                      const double* temperature, int Ht, int Wt,
                      double dt, double gravity, double buoyancy_coefficient,
                      double density_coefficient, double ambient_temperature) {
  if (temperature == NULL) { Ht = H; Wt = W; }
#pragma omp parallel for schedule(static)
  for (int j = 0; j < vh; ++j) {
    for (int i = 0; i < vw; ++i) {
      float bf = (j <= H && i < W)
                     ? buoyancy_at(density, H, W, j, i, density_coefficient, gravity)
                     : 0.0f;
      float tf = (j <= Ht && i < Wt)
                     ? temperature_at(temperature, Ht, Wt, j, i,
                                      ambient_temperature, buoyancy_coefficient)
                     : 0.0f;
      float s = bf + tf; /* float32 + float32 in float32 (forces.py:105) */
      v[IDX(j, i, vw)] += dt * (double)s;
    }
  }
}

/* forces.py:108-123 */
void orc_cool_temperature(double* T, int H, int W, double ambient,
                          double cooling_rate, double dt) {
  const double decay = exp(-cooling_rate * dt);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < H; ++j)
    for (int i = 0; i < W; ++i)
      T[IDX(j, i, W)] = ambient + (T[IDX(j, i, W)] - ambient) * decay;
}

/* ---------------------------------------------------------------------- */
/* divergence.py:4-25                                                     */
/* ---------------------------------------------------------------------- */
void orc_divergence(const double* u, const double* v, double* div,
                    const uint8_t* fmu, const uint8_t* fmv,
                    int H, int W, double cell_size) {
  const double inv = 1.0 / cell_size;
  const int W1 = W + 1;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < H; ++j) {
    for (int i = 0; i < W; ++i) {
      double du = ((u[IDX(j, i + 1, W1)] * (double)fmu[IDX(j, i + 1, W1)]) -
                   (u[IDX(j, i, W1)] * (double)fmu[IDX(j, i, W1)])) * inv;
      double dv = ((v[IDX(j + 1, i, W)] * (double)fmv[IDX(j + 1, i, W)]) -
                   (v[IDX(j, i, W)] * (double)fmv[IDX(j, i, W)])) * inv;
      div[IDX(j, i, W)] = du + dv;
    }
  }
}

/* ---------------------------------------------------------------------- */
/* pressure_gradient.py:4-37                                              */
/* An open face at index -1 wraps like Python negative indexing (i=0 reads */
/* p[j, W-1]).  An open face past the end (i=W or j=H) reads outside the   */
/* array in the reference (undefined); here it reads the clamped cell.     */
/* Neither occurs with face masks built by the reference (outer faces are  */
/* always closed).                                                         */
/* ---------------------------------------------------------------------- */
void orc_pressure_gradient(double* u, double* v, const double* p,
                           const uint8_t* fmu, const uint8_t* fmv,
                           int H, int W, double rho, double cell_size,
                           double dt) {
  const double scale = dt / (rho * cell_size);
  const int W1 = W + 1;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < H; ++j) {
    for (int i = 0; i < W1; ++i) {
      if (fmu[IDX(j, i, W1)]) {
        int ir = i < W ? i : W - 1;
        int il = i - 1 < 0 ? W - 1 : i - 1;
        u[IDX(j, i, W1)] -= scale * (p[IDX(j, ir, W)] - p[IDX(j, il, W)]);
      } else {
        u[IDX(j, i, W1)] = 0.0;
      }
    }
  }
#pragma omp parallel for schedule(static)
  for (int j = 0; j < H + 1; ++j) {
    for (int i = 0; i < W; ++i) {
      if (fmv[IDX(j, i, W)]) {
        int jt = j < H ? j : H - 1;
        int jb = j - 1 < 0 ? H - 1 : j - 1;
        v[IDX(j, i, W)] -= scale * (p[IDX(jt, i, W)] - p[IDX(jb, i, W)]);
      } else {
        v[IDX(j, i, W)] = 0.0;
      }
    }
  }
}

/* boundaries.py:7-39 */
void orc_enforce_boundaries(double* u, double* v, const uint8_t* solid,
                            int H, int W) {
  const int W1 = W + 1;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < H; ++j) {
    for (int i = 0; i < W1; ++i) {
      int left = i > 0 ? solid[IDX(j, i - 1, W)] == 1 : 0;
      int right = i < W ? solid[IDX(j, i, W)] == 1 : 0;
      if (left || right) u[IDX(j, i, W1)] = 0.0;
    }
  }
#pragma omp parallel for schedule(static)
  for (int j = 0; j < H + 1; ++j) {
    for (int i = 0; i < W; ++i) {
      int top = j > 0 ? solid[IDX(j - 1, i, W)] == 1 : 0;
      int bottom = j < H ? solid[IDX(j, i, W)] == 1 : 0;
      if (top || bottom) v[IDX(j, i, W)] = 0.0;
    }
  }
}

/* ---------------------------------------------------------------------- */
/* Poisson stencil shared by residual.py:35-40, jacobi.py:18-24,           */
/* red_black_sor_gauss_seidel.py:19-25, gauss_seidel.py:20-26              */
/* ---------------------------------------------------------------------- */
static inline double neighbour_sum(const double* p, const uint8_t* fmu,
                                   const uint8_t* fmv, int H, int W, int j,
                                   int i) {
  const int W1 = W + 1;
  double c = p[IDX(j, i, W)];
  double left = fmu[IDX(j, i, W1)] ? p[IDX(j, (i - 1 > 0 ? i - 1 : 0), W)] : c;
  double right = fmu[IDX(j, i + 1, W1)] ? p[IDX(j, (i + 1 < W - 1 ? i + 1 : W - 1), W)] : c;
  double down = fmv[IDX(j, i, W)] ? p[IDX((j - 1 > 0 ? j - 1 : 0), i, W)] : c;
  double up = fmv[IDX(j + 1, i, W)] ? p[IDX((j + 1 < H - 1 ? j + 1 : H - 1), i, W)] : c;
  return left + right + down + up;
}

/* residual.py:6-42 */
void orc_residual(const double* p, const double* div, const uint8_t* is_fluid,
                  const uint8_t* fmu, const uint8_t* fmv, int H, int W,
                  double alpha, double reciprocal_beta, double* r) {
  const double beta = 1.0 / reciprocal_beta;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < H; ++j) {
    for (int i = 0; i < W; ++i) {
      if (!is_fluid[IDX(j, i, W)]) { r[IDX(j, i, W)] = 0.0; continue; }
      double s = neighbour_sum(p, fmu, fmv, H, W, j, i);
      r[IDX(j, i, W)] = (s + alpha * div[IDX(j, i, W)]) - beta * p[IDX(j, i, W)];
    }
  }
}

/* solvers/common.py:6-18 */
double orc_alpha(double cell_size, double density, double dt) {
  return (-(cell_size * cell_size) * density) / dt;
}
/* solvers/common.py:30-39 */
double orc_optimal_omega(int size) {
  double w = 2.0 / (1.0 + sin(M_PI / (double)size));
  return w < 1.95 ? w : 1.95;
}

/* masks.py:52-74 == jacobi.py:46 */
void orc_is_fluid(const uint8_t* fmu, const uint8_t* fmv, int H, int W,
                  uint8_t* is_fluid) {
  const int W1 = W + 1;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < H; ++j)
    for (int i = 0; i < W; ++i)
      is_fluid[IDX(j, i, W)] =
          (fmu[IDX(j, i, W1)] || fmu[IDX(j, i + 1, W1)] ||
           fmv[IDX(j, i, W)] || fmv[IDX(j + 1, i, W)]) ? 1 : 0;
}

/* jacobi.py:8-24 */
void orc_jacobi_sweep(const double* p, double* p_out, const double* div,
                      const uint8_t* is_fluid, const uint8_t* fmu,
                      const uint8_t* fmv, int H, int W, double alpha,
                      double reciprocal_beta) {
#pragma omp parallel for schedule(static)
  for (int j = 0; j < H; ++j) {
    for (int i = 0; i < W; ++i) {
      if (!is_fluid[IDX(j, i, W)]) { p_out[IDX(j, i, W)] = p[IDX(j, i, W)]; continue; }
      double s = neighbour_sum(p, fmu, fmv, H, W, j, i);
      p_out[IDX(j, i, W)] = (s + alpha * div[IDX(j, i, W)]) * reciprocal_beta;
    }
  }
}

/* jacobi.py:27-52 */
void orc_jacobi_solve(double* p, const double* div, const uint8_t* fmu,
                      const uint8_t* fmv, int H, int W, double rho,
                      double cell_size, double dt, int iterations) {
  const size_t n = (size_t)H * W;
  const double alpha = orc_alpha(cell_size, rho, dt);
  const double rb = 0.25;
  double* tmp = (double*)malloc(n * sizeof(double));
  uint8_t* fl = (uint8_t*)malloc(n);
  orc_is_fluid(fmu, fmv, H, W, fl);
  double* src = p;
  double* dst = tmp;
  for (int it = 0; it < iterations; ++it) {
    orc_jacobi_sweep(src, dst, div, fl, fmu, fmv, H, W, alpha, rb);
    double* t = src; src = dst; dst = t;
  }
  if (src != p) memcpy(p, src, n * sizeof(double));
  free(tmp);
  free(fl);
}

/* red_black_sor_gauss_seidel.py:8-27 */
void orc_rbsor_sweeps(double* p, const double* div, const uint8_t* is_fluid,
                      const uint8_t* fmu, const uint8_t* fmv, int H, int W,
                      double alpha, double reciprocal_beta, double omega,
                      int iterations) {
  for (int it = 0; it < iterations; ++it) {
    for (int color = 0; color < 2; ++color) {
#pragma omp parallel for schedule(static)
      for (int j = 0; j < H; ++j) {
        int start = (j % 2 == 0) ? color : 1 - color;
        for (int i = start; i < W; i += 2) {
          if (!is_fluid[IDX(j, i, W)]) continue;
          double s = neighbour_sum(p, fmu, fmv, H, W, j, i);
          double np_ = (s + alpha * div[IDX(j, i, W)]) * reciprocal_beta;
          p[IDX(j, i, W)] = (1.0 - omega) * p[IDX(j, i, W)] + omega * np_;
        }
      }
    }
  }
}

/* red_black_sor_gauss_seidel.py:30-42 */
void orc_rbsor_solve(double* p, const double* div, const uint8_t* fmu,
                     const uint8_t* fmv, int H, int W, double rho,
                     double cell_size, double dt, int iterations,
                     double omega) {
  const double alpha = orc_alpha(cell_size, rho, dt);
  uint8_t* fl = (uint8_t*)malloc((size_t)H * W);
  orc_is_fluid(fmu, fmv, H, W, fl);
  if (omega <= 0.0 || omega > 2.0) omega = orc_optimal_omega(H > W ? H : W);
  orc_rbsor_sweeps(p, div, fl, fmu, fmv, H, W, alpha, 0.25, omega, iterations);
  free(fl);
}

/* gauss_seidel.py:8-48 */
void orc_gs_solve(double* p, const double* div, const uint8_t* fmu,
                  const uint8_t* fmv, int H, int W, double rho,
                  double cell_size, double dt, int iterations) {
  const double alpha = orc_alpha(cell_size, rho, dt);
  uint8_t* fl = (uint8_t*)malloc((size_t)H * W);
  orc_is_fluid(fmu, fmv, H, W, fl);
  for (int it = 0; it < iterations; ++it)
    for (int j = 0; j < H; ++j)
      for (int i = 0; i < W; ++i) {
        if (!fl[IDX(j, i, W)]) continue;
        double s = neighbour_sum(p, fmu, fmv, H, W, j, i);
        p[IDX(j, i, W)] = (s + alpha * div[IDX(j, i, W)]) * 0.25;
      }
  free(fl);
}

double orc_norm2_fluid(const double* r, const uint8_t* is_fluid, int H, int W) {
  double acc = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : acc)
  for (int j = 0; j < H; ++j)
    for (int i = 0; i < W; ++i)
      if (is_fluid[IDX(j, i, W)]) acc += r[IDX(j, i, W)] * r[IDX(j, i, W)];
  return acc;
}

/* ||alpha*div||^2 over fluid cells, the reference point of the relative
 * residual (the residual of p = 0 equals alpha*div). */
static double rhs_norm2(const double* div, const uint8_t* fl, int H, int W,
                        double alpha) {
  double acc = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : acc)
  for (int j = 0; j < H; ++j)
    for (int i = 0; i < W; ++i)
      if (fl[IDX(j, i, W)]) {
        double b = alpha * div[IDX(j, i, W)];
        acc += b * b;
      }
  return acc;
}

int orc_rbsor_solve_tol(double* p, const double* div, const uint8_t* fmu,
                        const uint8_t* fmv, int H, int W, double rho,
                        double cell_size, double dt, double omega, double tol,
                        int check_every, int max_sweeps, double* rel_out) {
  const size_t n = (size_t)H * W;
  const double alpha = orc_alpha(cell_size, rho, dt);
  uint8_t* fl = (uint8_t*)malloc(n);
  double* r = (double*)malloc(n * sizeof(double));
  orc_is_fluid(fmu, fmv, H, W, fl);
  if (omega <= 0.0 || omega > 2.0) omega = orc_optimal_omega(H > W ? H : W);
  if (check_every < 1) check_every = 1;
  const double b2 = rhs_norm2(div, fl, H, W, alpha);
  int done = 0;
  double rel = 0.0;
  orc_residual(p, div, fl, fmu, fmv, H, W, alpha, 0.25, r);
  rel = b2 > 0.0 ? sqrt(orc_norm2_fluid(r, fl, H, W) / b2) : 0.0;
  while (rel > tol && done < max_sweeps) {
    int k = check_every;
    if (done + k > max_sweeps) k = max_sweeps - done;
    orc_rbsor_sweeps(p, div, fl, fmu, fmv, H, W, alpha, 0.25, omega, k);
    done += k;
    orc_residual(p, div, fl, fmu, fmv, H, W, alpha, 0.25, r);
    rel = b2 > 0.0 ? sqrt(orc_norm2_fluid(r, fl, H, W) / b2) : 0.0;
  }
  if (rel_out) *rel_out = rel;
  free(fl);
  free(r);
  return done;
}

/* ---------------------------------------------------------------------- */
/* multigrid                                                              */
/* ---------------------------------------------------------------------- */
/* masks.py:3-27 */
void orc_coarsen_solid(const uint8_t* sf, int H, int W, uint8_t* sc) {
  const int Hc = H / 2, Wc = W / 2;
#pragma omp parallel for schedule(static)
  for (int J = 0; J < Hc; ++J)
    for (int I = 0; I < Wc; ++I) {
      int j0 = 2 * J, i0 = 2 * I;
      sc[IDX(J, I, Wc)] = (sf[IDX(j0, i0, W)] || sf[IDX(j0, i0 + 1, W)] ||
                           sf[IDX(j0 + 1, i0, W)] || sf[IDX(j0 + 1, i0 + 1, W)]) ? 1 : 0;
    }
}

/* masks.py:30-49 */
void orc_build_face_masks(const uint8_t* s, int H, int W, uint8_t* fmu,
                          uint8_t* fmv) {
  const int W1 = W + 1;
  memset(fmu, 0, (size_t)H * W1);
  memset(fmv, 0, (size_t)(H + 1) * W);
  for (int j = 0; j < H; ++j)
    for (int i = 1; i < W; ++i)
      fmu[IDX(j, i, W1)] = !(s[IDX(j, i - 1, W)] || s[IDX(j, i, W)]);
  for (int j = 1; j < H; ++j)
    for (int i = 0; i < W; ++i)
      fmv[IDX(j, i, W)] = !(s[IDX(j - 1, i, W)] || s[IDX(j, i, W)]);
}

/* transfer.py:5-36 */
void orc_restrict(const double* fine, const uint8_t* fl, int H, int W,
                  double* coarse) {
  const int Hc = H / 2, Wc = W / 2;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < Hc; ++j)
    for (int i = 0; i < Wc; ++i) {
      int j0 = 2 * j, i0 = 2 * i;
      double total = 0.0;
      int count = 0;
      for (int dj = 0; dj < 2; ++dj)
        for (int di = 0; di < 2; ++di) {
          int jj = j0 + dj, ii = i0 + di;
          if (fl[IDX(jj, ii, W)]) { total += fine[IDX(jj, ii, W)]; count += 1; }
        }
      coarse[IDX(j, i, Wc)] = count > 0 ? total / (double)count : 0.0;
    }
}

/* transfer.py:38-72 */
void orc_prolong(const double* coarse, int Hc, int Wc, const uint8_t* fl,
                 int H, int W, double* fine) {
#pragma omp parallel for schedule(static)
  for (int j = 0; j < H; ++j)
    for (int i = 0; i < W; ++i) {
      if (!fl[IDX(j, i, W)]) { fine[IDX(j, i, W)] = 0.0; continue; }
      double xc = (i + 0.5) * 0.5;
      double yc = (j + 0.5) * 0.5;
      fine[IDX(j, i, W)] = sample_c(coarse, Hc, Wc, xc, yc);
    }
}

typedef struct {
  int H, W;
  uint8_t *fmu, *fmv, *fl, *solid;
  double *p, *r, *rhs, *corr;
  int owns_masks;
} orc_level;

/* solver.py:94-139 with the smoother contract of solver.py:107,137 */
static void mg_smooth(orc_level* L, double rho, double h, double dt, int smoother,
                      int n_iter, double omega_override) {
  switch (smoother) {
    case ORC_SMOOTH_JACOBI:
      orc_jacobi_solve(L->p, L->rhs, L->fmu, L->fmv, L->H, L->W, rho, h, dt, n_iter);
      break;
    case ORC_SMOOTH_GS:
      orc_gs_solve(L->p, L->rhs, L->fmu, L->fmv, L->H, L->W, rho, h, dt, n_iter);
      break;
    default:
      orc_rbsor_solve(L->p, L->rhs, L->fmu, L->fmv, L->H, L->W, rho, h, dt, n_iter,
                      omega_override > 0.0 ? omega_override : 0.0);
      break;
  }
}

/* The omega of the coarsest level's red-black smoothing: coarsest_omega > 0
 * as given, < 0 the optimal omega of that level's size (common.py:30-39),
 * 0 the other levels' (a new option; 0 is the reference's behaviour). */
static double coarsest_om(const orc_level* L, double om, double coarsest_omega) {
  if (coarsest_omega > 0.0) return coarsest_omega;
  if (coarsest_omega < 0.0) return orc_optimal_omega(L->H > L->W ? L->H : L->W);
  return om;
}

static void mg_vcycle(orc_level* lv, int nl, int l, double rho, double h,
                      double dt, int smoother, int si, int ci, double om,
                      int rediscretize, double com) {
  orc_level* L = &lv[l];
  const int coarsest = (l == nl - 1);
  const int n_iter = coarsest ? ci : si;
  mg_smooth(L, rho, h, dt, smoother, n_iter, coarsest ? coarsest_om(L, om, com) : om);
  if (coarsest) return;
  const double alpha = orc_alpha(h, rho, dt);
  orc_residual(L->p, L->rhs, L->fl, L->fmu, L->fmv, L->H, L->W, alpha, 0.25, L->r);
  orc_level* C = &lv[l + 1];
  orc_restrict(L->r, L->fl, L->H, L->W, C->rhs);
  /* solver.py:124-125 divides by alpha(2h): the coarse system is then
   * (4 - N_c) e = R r, a quarter of the rediscretised correction.  The
   * rediscretised option divides by alpha(h) instead (= 4 R r / alpha(2h)
   * exactly, the factor 4 being a power of two). */
  const double alpha_c = rediscretize ? alpha : orc_alpha(h * 2.0, rho, dt);
  const size_t nc = (size_t)C->H * C->W;
  for (size_t k = 0; k < nc; ++k) C->rhs[k] /= alpha_c; /* solver.py:125 */
  memset(C->p, 0, nc * sizeof(double));
  mg_vcycle(lv, nl, l + 1, rho, h * 2.0, dt, smoother, si, ci, om, rediscretize, com);
  orc_prolong(C->p, C->H, C->W, L->fl, L->H, L->W, L->corr);
  const size_t n = (size_t)L->H * L->W;
  for (size_t k = 0; k < n; ++k) L->p[k] += L->corr[k]; /* solver.py:134 */
  mg_smooth(L, rho, h, dt, smoother, n_iter, om);
}

int orc_mg_solve(double* p, const double* div, const uint8_t* fmu,
                 const uint8_t* fmv, int H, int W, double rho,
                 double cell_size, double dt, int v_cycles, int smoother,
                 int num_levels, int smoother_iterations,
                 int coarsest_iterations, double smoother_omega,
                 int rediscretize, double tol, double* rel_out) {
  return orc_mg_solve_cached(p, div, fmu, fmv, NULL, H, W, rho, cell_size, dt,
                             v_cycles, smoother, num_levels, smoother_iterations,
                             coarsest_iterations, smoother_omega, 0.0,
                             rediscretize, tol, rel_out);
}

/* As orc_mg_solve, with the level cache of solver.py:87-92 left in place
 * after the face-mask arrays changed: level_fluid (H x W, may be NULL =
 * derived from fmu/fmv) is level 0's is_fluid as the cache built it.  The
 * level-0 face masks are fmu/fmv (the cache holds the arrays, solver.py:45-46)
 * and each smoother call derives its own is_fluid from them
 * (red_black_sor_gauss_seidel.py:38, jacobi.py:46, gauss_seidel.py:46);
 * the residual, restriction and prolongation of level 0 read level_fluid
 * (solver.py:98,116,122,133), and the coarse levels are built from
 * solid = ~level_fluid (solver.py:40). */
int orc_mg_solve_cached(double* p, const double* div, const uint8_t* fmu,
                        const uint8_t* fmv, const uint8_t* level_fluid, int H,
                        int W, double rho, double cell_size, double dt,
                        int v_cycles, int smoother, int num_levels,
                        int smoother_iterations, int coarsest_iterations,
                        double smoother_omega, double coarsest_omega,
                        int rediscretize, double tol, double* rel_out) {
  if (num_levels < 1) num_levels = 1;
  orc_level* lv = (orc_level*)calloc((size_t)num_levels, sizeof(orc_level));
  /* solver.py:34-73 _build_levels */
  int nl = 0;
  {
    orc_level* L = &lv[0];
    L->H = H; L->W = W;
    L->fmu = (uint8_t*)fmu; L->fmv = (uint8_t*)fmv; L->owns_masks = 0;
    L->fl = (uint8_t*)malloc((size_t)H * W);
    L->solid = (uint8_t*)malloc((size_t)H * W);
    if (level_fluid)
      memcpy(L->fl, level_fluid, (size_t)H * W);
    else
      orc_is_fluid(fmu, fmv, H, W, L->fl);
    for (size_t k = 0; k < (size_t)H * W; ++k) L->solid[k] = !L->fl[k];
    nl = 1;
  }
  for (int l = 1; l < num_levels; ++l) {
    orc_level* F = &lv[l - 1];
    if (F->H % 2 != 0 || F->W % 2 != 0) break;
    orc_level* C = &lv[l];
    C->H = F->H / 2; C->W = F->W / 2;
    C->solid = (uint8_t*)malloc((size_t)C->H * C->W);
    orc_coarsen_solid(F->solid, F->H, F->W, C->solid);
    C->fmu = (uint8_t*)malloc((size_t)C->H * (C->W + 1));
    C->fmv = (uint8_t*)malloc((size_t)(C->H + 1) * C->W);
    C->owns_masks = 1;
    orc_build_face_masks(C->solid, C->H, C->W, C->fmu, C->fmv);
    C->fl = (uint8_t*)malloc((size_t)C->H * C->W);
    orc_is_fluid(C->fmu, C->fmv, C->H, C->W, C->fl);
    nl = l + 1;
  }
  /* solver.py:75-85 _build_buffers */
  for (int l = 0; l < nl; ++l) {
    size_t n = (size_t)lv[l].H * lv[l].W;
    lv[l].p = (double*)calloc(n, sizeof(double));
    lv[l].r = (double*)calloc(n, sizeof(double));
    lv[l].rhs = (double*)calloc(n, sizeof(double));
    lv[l].corr = (double*)calloc(n, sizeof(double));
  }
  const size_t n0 = (size_t)H * W;
  memcpy(lv[0].p, p, n0 * sizeof(double));
  memcpy(lv[0].rhs, div, n0 * sizeof(double));
  int cycles = 0;
  double rel = -1.0;
  if (tol > 0.0) {
    const double alpha = orc_alpha(cell_size, rho, dt);
    const double b2 = rhs_norm2(div, lv[0].fl, H, W, alpha);
    orc_residual(lv[0].p, lv[0].rhs, lv[0].fl, fmu, fmv, H, W, alpha, 0.25, lv[0].r);
    rel = b2 > 0.0 ? sqrt(orc_norm2_fluid(lv[0].r, lv[0].fl, H, W) / b2) : 0.0;
    while (rel > tol && cycles < v_cycles) {
      mg_vcycle(lv, nl, 0, rho, cell_size, dt, smoother, smoother_iterations,
                coarsest_iterations, smoother_omega, rediscretize, coarsest_omega);
      cycles++;
      orc_residual(lv[0].p, lv[0].rhs, lv[0].fl, fmu, fmv, H, W, alpha, 0.25, lv[0].r);
      rel = b2 > 0.0 ? sqrt(orc_norm2_fluid(lv[0].r, lv[0].fl, H, W) / b2) : 0.0;
    }
  } else {
    for (int c = 0; c < v_cycles; ++c) {
      mg_vcycle(lv, nl, 0, rho, cell_size, dt, smoother, smoother_iterations,
                coarsest_iterations, smoother_omega, rediscretize, coarsest_omega);
      cycles++;
    }
  }
  memcpy(p, lv[0].p, n0 * sizeof(double));
  if (rel_out) *rel_out = rel;
  for (int l = 0; l < nl; ++l) {
    free(lv[l].p); free(lv[l].r); free(lv[l].rhs); free(lv[l].corr);
    free(lv[l].fl); free(lv[l].solid);
    if (lv[l].owns_masks) { free(lv[l].fmu); free(lv[l].fmv); }
  }
  free(lv);
  return cycles;
}

/* ---------------------------------------------------------------------- */
/* emitters.py                                                            */
/* ---------------------------------------------------------------------- */
/* emitters.py:339-411 (python approach classes == numba :6-35) */
double orc_falloff(double distance, double radius, int approach) {
  switch (approach) {
    case 0: return distance < radius ? 1.0 : 0.0;
    case 1:
      if (distance >= radius) return 0.0;
      return 1.0 - (distance / radius);
    case 2: {
      if (distance >= radius) return 0.0;
      double q = distance / radius;
      return 1.0 - q * q;
    }
    case 3: {
      if (distance >= radius) return 0.0;
      double sigma = radius / 3.0;
      double q = distance / sigma;
      return exp(-0.5 * (q * q));
    }
    default: return 0.0;
  }
}

/* emitters.py:157-178 */
void orc_emit_scalar(double* field, int H, int W, double value, double dt,
                     double pos_x, double pos_y, double radius, int approach) {
#pragma omp parallel for schedule(static)
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      double dx = x - pos_x, dy = y - pos_y;
      double distance = sqrt(dx * dx + dy * dy);
      double f = orc_falloff(distance, radius, approach);
      field[IDX(y, x, W)] += value * f * dt;
    }
}
