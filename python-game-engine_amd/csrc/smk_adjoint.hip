// smk_adjoint.hip -- differentiable transport of the smoke step.
//
// Forward pieces whose controls live on the device (box strength read from a
// device array, buoyancy in the field type), and the adjoint (transposed
// Jacobian) of every transport operator of SmokeSimulation2D.substep
// (src/cfd/common/simulation.py:206-249):
//
//   box set      rho[box] = s[n]                      -> s_bar[n] += sum_box rho_bar; rho_bar[box] = 0
//   buoyancy     v += dt * (-dc*g) * face_avg(rho)     -> rho_bar += dt * (-dc*g) * face_avg^T(v_bar)
//   advect q     q'(x) = bilerp(q, x - dt*vel(x)/h)    -> q_bar  = bilerp^T(q'_bar)            (field)
//                                                        vel_bar += q'_bar * grad bilerp * (-dt/h) (backtrace)
//
// The projection needs no kernel of its own: with w = M B u_bar (closed-face
// mask M, solid-adjacent faces B) the adjoint is w - c G S D w, i.e. the same
// divergence / Poisson solve / gradient as the forward applied to w (the
// masked Poisson operator is symmetric).
//
// The transpose of the bilinear gather is computed two ways:
//   SMK_ADJ_GATHER  each source cell visits the targets whose backtrace can
//                   reach it (a window of radius max|displacement|+1, read
//                   from device memory) and sums their weights in a fixed
//                   order: deterministic, bit-reproducible;
//   SMK_ADJ_SCATTER each target adds its four weights with atomics.
// Both give the same values up to summation order.
//
// Kinks: where a backtrace lands exactly on a grid line the bilinear sample
// is not differentiable across it (zero velocity, as at the start of a
// rollout).  There the slope is the mean of the two one-sided slopes, which
// is what a central finite difference measures.
#include <algorithm>

#include "smk_common.hpp"

namespace smk {
namespace {

constexpr int AX = 64, AY = 4;
inline dim3 agrid(int rows, int cols) { return dim3(cdiv(cols, AX), cdiv(rows, AY)); }

// Footprint of sample_field at (x, y) on an fh x fw grid: the clamped corner
// indices and the interpolation parameters (0 where both corners coincide).
struct Foot {
  int x0, x1, y0, y1;
  double tx, ty;
};

__device__ __forceinline__ Foot foot_of(int fh, int fw, double x, double y) {
  Foot f;
  f.x0 = clamp_index(floor(x), fw - 1);
  f.x1 = clamp_index(ceil(x), fw - 1);
  f.y0 = clamp_index(floor(y), fh - 1);
  f.y1 = clamp_index(ceil(y), fh - 1);
  f.tx = (f.x0 == f.x1) ? 0.0 : x - (double)f.x0;
  f.ty = (f.y0 == f.y1) ? 0.0 : y - (double)f.y0;
  return f;
}

// d sample / d f[r][c]
__device__ __forceinline__ double foot_weight(const Foot& f, int r, int c) {
  double wx = 0.0, wy = 0.0;
  if (c == f.x0) wx += 1.0 - f.tx;
  if (c == f.x1 && f.x1 != f.x0) wx += f.tx;
  if (r == f.y0) wy += 1.0 - f.ty;
  if (r == f.y1 && f.y1 != f.y0) wy += f.ty;
  return wx * wy;
}

// A target whose (clamped) sample position is a full cell or more away from
// (r, c) on either axis cannot have (r, c) in its footprint.
__device__ __forceinline__ bool may_touch(int fh, int fw, double x, double y, int r, int c) {
  const double xc = fmin(fmax(x, 0.0), (double)(fw - 1));
  const double yc = fmin(fmax(y, 0.0), (double)(fh - 1));
  return fabs(xc - (double)c) < 1.0 && fabs(yc - (double)r) < 1.0;
}

// d sample / dx and d sample / dy at (x, y).  Off grid lines this is the
// derivative of the bilinear form; on an interior grid line it is the mean of
// the one-sided derivatives; in the clamped region outside the grid it is 0.
template <typename T>
__device__ __forceinline__ void sample_slopes(const T* __restrict__ f, int fh, int fw, double x,
                                              double y, double& dx, double& dy) {
  const double fx = floor(x), cx = ceil(x), fy = floor(y), cy = ceil(y);
  const int x0 = clamp_index(fx, fw - 1), x1 = clamp_index(cx, fw - 1);
  const int y0 = clamp_index(fy, fh - 1), y1 = clamp_index(cy, fh - 1);
  const double tx = (x0 == x1) ? 0.0 : x - (double)x0;
  const double ty = (y0 == y1) ? 0.0 : y - (double)y0;
  auto at = [&](int r, int c) { return (double)f[(size_t)r * fw + c]; };
  const bool xk = (fx == cx) && fx >= 0.0 && fx <= (double)(fw - 1);
  const bool yk = (fy == cy) && fy >= 0.0 && fy <= (double)(fh - 1);
  double sx0, sx1, sy0, sy1;
  if (xk) {
    const int xm = x0 > 0 ? x0 - 1 : 0, xp = x0 < fw - 1 ? x0 + 1 : fw - 1;
    sx0 = 0.5 * (at(y0, xp) - at(y0, xm));
    sx1 = 0.5 * (at(y1, xp) - at(y1, xm));
  } else {
    sx0 = at(y0, x1) - at(y0, x0);
    sx1 = at(y1, x1) - at(y1, x0);
  }
  if (yk) {
    const int ym = y0 > 0 ? y0 - 1 : 0, yp = y0 < fh - 1 ? y0 + 1 : fh - 1;
    sy0 = 0.5 * (at(yp, x0) - at(ym, x0));
    sy1 = 0.5 * (at(yp, x1) - at(ym, x1));
  } else {
    sy0 = at(y1, x0) - at(y0, x0);
    sy1 = at(y1, x1) - at(y0, x1);
  }
  dx = sx0 * (1.0 - ty) + sx1 * ty;
  dy = sy0 * (1.0 - tx) + sy1 * tx;
}

// whole cells a backtrace moved (window radius contribution)
__device__ __forceinline__ int disp_cells(double dx, double dy) {
  const double m = fmin(fmax(fabs(dx), fabs(dy)), 1.0e6);
  return (int)ceil(m);
}

__device__ __forceinline__ void wave_max_to(int r, int* dst) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) r = max(r, __shfl_xor(r, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(dst, r);
}

template <typename T>
__device__ __forceinline__ void scatter_foot(T* __restrict__ out, int fh, int fw, double x, double y,
                                             double g) {
  const Foot f = foot_of(fh, fw, x, y);
  const double ax = 1.0 - f.tx, ay = 1.0 - f.ty;
  atomicAdd(&out[(size_t)f.y0 * fw + f.x0], (T)(g * (ax * ay)));
  if (f.x1 != f.x0) atomicAdd(&out[(size_t)f.y0 * fw + f.x1], (T)(g * (f.tx * ay)));
  if (f.y1 != f.y0) {
    atomicAdd(&out[(size_t)f.y1 * fw + f.x0], (T)(g * (ax * f.ty)));
    if (f.x1 != f.x0) atomicAdd(&out[(size_t)f.y1 * fw + f.x1], (T)(g * (f.tx * f.ty)));
  }
}

// Sum over the targets of a window of y[t] * d sample_t / d f[r][c], where
// target t samples an fh x fw grid at (px[t], py[t]); targets and sources
// share the grid's indexing.
template <typename T>
__device__ __forceinline__ double window_gather(const T* __restrict__ yt,
                                                const double* __restrict__ px,
                                                const double* __restrict__ py, int fh, int fw,
                                                int r, int c, int R) {
  double acc = 0.0;
  const int j0 = max(r - R, 0), j1 = min(r + R, fh - 1);
  const int i0 = max(c - R, 0), i1 = min(c + R, fw - 1);
  for (int tj = j0; tj <= j1; ++tj) {
    for (int ti = i0; ti <= i1; ++ti) {
      const size_t t = (size_t)tj * fw + ti;
      const double y = (double)yt[t];
      if (y == 0.0) continue;
      const double x = px[t], yy = py[t];
      if (!may_touch(fh, fw, x, yy, r, c)) continue;
      acc += y * foot_weight(foot_of(fh, fw, x, yy), r, c);
    }
  }
  return acc;
}

template <typename T>
__global__ void k_fill(T* __restrict__ p, size_t n, T val) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    p[k] = val;
}

// ------------------------------------------------------------ forward pieces
template <typename T>
__global__ void k_box_set_indexed(T* __restrict__ f, int W, int x0, int y0, int x1, int y1,
                                  const T* __restrict__ values, int index) {
  const int i = x0 + blockIdx.x * AX + threadIdx.x, j = y0 + blockIdx.y * AY + threadIdx.y;
  if (i >= x1 || j >= y1) return;
  f[(size_t)j * W + i] = values[index];
}

// forces.py:19-46 face averaging with the force kept in the field type.
template <typename T>
__global__ void k_buoyancy(T* __restrict__ v, const T* __restrict__ rho, int H, int W, double dt,
                           double g, double dc) {
  const int i = blockIdx.x * AX + threadIdx.x, j = blockIdx.y * AY + threadIdx.y;
  if (j >= H || i >= W) return;  // row H: force 0
  double d;
  if (j == 0 || j == H - 1) {
    d = (double)rho[(size_t)j * W + i];
  } else {
    const T s = rho[(size_t)(j - 1) * W + i] + rho[(size_t)j * W + i];
    d = 0.5 * (double)s;
  }
  const size_t o = (size_t)j * W + i;
  v[o] = (T)((double)v[o] + dt * (-dc * d * g));
}

// ------------------------------------------------------------ adjoints
template <typename T>
__global__ void k_buoyancy_adj(const T* __restrict__ vbar, T* __restrict__ rbar, int H, int W,
                               double scale) {
  const int i = blockIdx.x * AX + threadIdx.x, j = blockIdx.y * AY + threadIdx.y;
  if (j >= H || i >= W) return;
  // cell j enters face j (alone on faces 0 and H-1, halved otherwise) and,
  // halved, face j+1 when that face averages (1 <= j+1 < H-1).
  const double ws = (j == 0 || j == H - 1) ? 1.0 : 0.5;
  double acc = ws * (double)vbar[(size_t)j * W + i];
  if (j + 1 < H - 1) acc += 0.5 * (double)vbar[(size_t)(j + 1) * W + i];
  const size_t o = (size_t)j * W + i;
  rbar[o] = (T)((double)rbar[o] + scale * acc);
}

template <int NT, typename T>
__global__ void __launch_bounds__(NT) k_box_set_adj(T* __restrict__ fbar, int W, int x0, int y0,
                                                    int x1, int y1, T* __restrict__ vbar,
                                                    int index) {
  __shared__ double red[NT / 64];
  const int bw = x1 - x0, n = bw * (y1 - y0);
  double acc = 0.0;
  for (int k = threadIdx.x; k < n; k += NT) {
    const size_t o = (size_t)(y0 + k / bw) * W + (x0 + k % bw);
    acc += (double)fbar[o];
  }
  acc = block_sum<NT>(acc, red);
  if (threadIdx.x == 0) vbar[index] = (T)((double)vbar[index] + acc);
  for (int k = threadIdx.x; k < n; k += NT) fbar[(size_t)(y0 + k / bw) * W + (x0 + k % bw)] = (T)0;
}

// advection.py:112-128 transposed, pass 1 (per target cell): masked adjoint,
// backtrace position (cell-index coordinates) and the backtrace sensitivity.
template <typename T>
__global__ void k_adj_scalar_sens(const T* __restrict__ ybar, const uint8_t* __restrict__ fluid,
                                  const T* __restrict__ src, const T* __restrict__ u,
                                  const T* __restrict__ v, int H, int W, double dt, double inv,
                                  T* __restrict__ ym, double* __restrict__ pxo,
                                  double* __restrict__ pyo, T* __restrict__ gx,
                                  T* __restrict__ gy, int* __restrict__ rmax, int scatter,
                                  T* __restrict__ sbar) {
  const int i = blockIdx.x * AX + threadIdx.x, j = blockIdx.y * AY + threadIdx.y;
  int r = 0;
  if (j < H && i < W) {
    const size_t o = (size_t)j * W + i;
    const double y = fluid[o] ? (double)ybar[o] : 0.0;
    const double px = i + 0.5, py = j + 0.5;
    const double uu = sample_known(u, H, W + 1, px, py - 0.5, i, i + 1, j, j);
    const double vv = sample_known(v, H + 1, W, px - 0.5, py, i, i, j, j + 1);
    const double x = px - (uu * dt * inv);
    const double yy = py - (vv * dt * inv);
    const double sx = x - 0.5, sy = yy - 0.5;
    double dx = 0.0, dy = 0.0;
    if (y != 0.0) sample_slopes(src, H, W, sx, sy, dx, dy);
    const double k = dt * inv;
    ym[o] = (T)y;
    pxo[o] = sx;
    pyo[o] = sy;
    gx[o] = (T)(-(y * dx) * k);
    gy[o] = (T)(-(y * dy) * k);
    r = disp_cells(sx - i, sy - j);
    if (scatter && y != 0.0) scatter_foot(sbar, H, W, sx, sy, y);
  }
  wave_max_to(r, rmax);
}

// pass 2: source gather (WINDOW) and the backtrace sensitivity pushed onto
// the velocity samples (cell velocity = mean of its two faces).
template <typename T, bool WINDOW>
__global__ void k_adj_scalar_gather(const T* __restrict__ ym, const double* __restrict__ pxo,
                                    const double* __restrict__ pyo, const T* __restrict__ gx,
                                    const T* __restrict__ gy, int H, int W,
                                    const int* __restrict__ rmax, T* __restrict__ sbar,
                                    T* __restrict__ ubar, T* __restrict__ vbar) {
  const int i = blockIdx.x * AX + threadIdx.x, j = blockIdx.y * AY + threadIdx.y;
  if (WINDOW && j < H && i < W) {
    const int R = *rmax + 1;
    sbar[(size_t)j * W + i] = (T)window_gather(ym, pxo, pyo, H, W, j, i, R);
  }
  if (ubar != nullptr && j < H && i <= W) {
    const size_t c = (size_t)j * W + i;
    double a = 0.0;
    if (i >= 1) a += 0.5 * (double)gx[c - 1];
    if (i < W) a += 0.5 * (double)gx[c];
    const size_t o = (size_t)j * (W + 1) + i;
    ubar[o] = (T)((double)ubar[o] + a);
  }
  if (vbar != nullptr && j <= H && i < W) {
    double a = 0.0;
    if (j >= 1) a += 0.5 * (double)gy[(size_t)(j - 1) * W + i];
    if (j < H) a += 0.5 * (double)gy[(size_t)j * W + i];
    const size_t o = (size_t)j * W + i;
    vbar[o] = (T)((double)vbar[o] + a);
  }
}

// advection.py:130-180 transposed, pass 1 (per target face).
template <typename T>
__global__ void k_adj_vel_sens(const T* __restrict__ udbar, const T* __restrict__ vdbar,
                               const uint8_t* __restrict__ fmu, const uint8_t* __restrict__ fmv,
                               const T* __restrict__ u, const T* __restrict__ v, int H, int W,
                               double dt, double inv, T* __restrict__ yu,
                               double* __restrict__ pxu, double* __restrict__ pyu,
                               T* __restrict__ sxu, T* __restrict__ syu, T* __restrict__ yv,
                               double* __restrict__ pxv, double* __restrict__ pyv,
                               T* __restrict__ sxv, T* __restrict__ syv, int* __restrict__ rmax,
                               int scatter, T* __restrict__ ubar, T* __restrict__ vbar) {
  const int i = blockIdx.x * AX + threadIdx.x, j = blockIdx.y * AY + threadIdx.y;
  const double k = dt * inv;
  int r = 0;
  if (j < H && i <= W) {  // u face: px = i, py = j + 1/2
    const size_t o = (size_t)j * (W + 1) + i;
    const double y = fmu[o] ? (double)udbar[o] : 0.0;
    const double px = (double)i, py = j + 0.5;
    const double uu = (double)u[o];
    const double vv = sample_known(v, H + 1, W, px - 0.5, py, i - 1, i, j, j + 1);
    const double x = px - (uu * dt * inv);
    const double yy = py - (vv * dt * inv);
    const double sx = x, sy = yy - 0.5;
    double dx = 0.0, dy = 0.0;
    if (y != 0.0) sample_slopes(u, H, W + 1, sx, sy, dx, dy);
    yu[o] = (T)y;
    pxu[o] = sx;
    pyu[o] = sy;
    sxu[o] = (T)(-(y * dx) * k);
    syu[o] = (T)(-(y * dy) * k);
    r = disp_cells(sx - i, sy - j);
    if (scatter && y != 0.0) scatter_foot(ubar, H, W + 1, sx, sy, y);
  }
  if (j <= H && i < W) {  // v face: px = i + 1/2, py = j
    const size_t o = (size_t)j * W + i;
    const double y = fmv[o] ? (double)vdbar[o] : 0.0;
    const double px = i + 0.5, py = (double)j;
    const double uu = sample_known(u, H, W + 1, px, py - 0.5, i, i + 1, j - 1, j);
    const double vv = (double)v[o];
    const double x = px - (uu * dt * inv);
    const double yy = py - (vv * dt * inv);
    const double sx = x - 0.5, sy = yy;
    double dx = 0.0, dy = 0.0;
    if (y != 0.0) sample_slopes(v, H + 1, W, sx, sy, dx, dy);
    yv[o] = (T)y;
    pxv[o] = sx;
    pyv[o] = sy;
    sxv[o] = (T)(-(y * dx) * k);
    syv[o] = (T)(-(y * dy) * k);
    r = max(r, disp_cells(sx - i, sy - j));
    if (scatter && y != 0.0) scatter_foot(vbar, H + 1, W, sx, sy, y);
  }
  wave_max_to(r, rmax);
}

// pass 2: per source face, the field transpose (WINDOW; otherwise the atomic
// scatter already left it in ubar / vbar), the direct backtrace term of the
// face's own velocity, and the transposed cross-component velocity sample
// (v sampled at u faces, u sampled at v faces).
template <typename T, bool WINDOW>
__global__ void k_adj_vel_gather(const T* __restrict__ yu, const double* __restrict__ pxu,
                                 const double* __restrict__ pyu, const T* __restrict__ sxu,
                                 const T* __restrict__ syu, const T* __restrict__ yv,
                                 const double* __restrict__ pxv, const double* __restrict__ pyv,
                                 const T* __restrict__ sxv, const T* __restrict__ syv, int H,
                                 int W, const int* __restrict__ rmax, T* __restrict__ ubar,
                                 T* __restrict__ vbar) {
  const int i = blockIdx.x * AX + threadIdx.x, j = blockIdx.y * AY + threadIdx.y;
  const int R = WINDOW ? *rmax + 1 : 0;
  if (j < H && i <= W) {
    const size_t o = (size_t)j * (W + 1) + i;
    double acc = WINDOW ? window_gather(yu, pxu, pyu, H, W + 1, j, i, R) : (double)ubar[o];
    acc += (double)sxu[o];
    // v faces (tj, ti) sample u at (ti + 1/2, tj - 1/2): rows tj-1, tj (clamped), columns ti, ti+1
#pragma unroll
    for (int dj = 0; dj < 2; ++dj) {
      const int tj = j + dj;
      if (tj > H) continue;
#pragma unroll
      for (int di = -1; di <= 0; ++di) {
        const int ti = i + di;
        if (ti < 0 || ti >= W) continue;
        const double s = (double)sxv[(size_t)tj * W + ti];
        if (s != 0.0) acc += s * foot_weight(foot_of(H, W + 1, ti + 0.5, tj - 0.5), j, i);
      }
    }
    ubar[o] = (T)acc;
  }
  if (j <= H && i < W) {
    const size_t o = (size_t)j * W + i;
    double acc = WINDOW ? window_gather(yv, pxv, pyv, H + 1, W, j, i, R) : (double)vbar[o];
    acc += (double)syv[o];
    // u faces (tj, ti) sample v at (ti - 1/2, tj + 1/2): rows tj, tj+1, columns ti-1, ti (clamped)
#pragma unroll
    for (int dj = -1; dj <= 0; ++dj) {
      const int tj = j + dj;
      if (tj < 0 || tj >= H) continue;
#pragma unroll
      for (int di = 0; di < 2; ++di) {
        const int ti = i + di;
        if (ti > W) continue;
        const double s = (double)syu[(size_t)tj * (W + 1) + ti];
        if (s != 0.0) acc += s * foot_weight(foot_of(H + 1, W, ti - 0.5, tj + 0.5), j, i);
      }
    }
    vbar[o] = (T)acc;
  }
}

// ------------------------------------------------------------ workspace
inline size_t al256(size_t n) { return (n + 255) & ~(size_t)255; }

template <typename T>
struct ScalarWs {
  T *ym, *gx, *gy;
  double *px, *py;
  int* rmax;
};

template <typename T>
struct VelWs {
  T *yu, *sxu, *syu, *yv, *sxv, *syv;
  double *pxu, *pyu, *pxv, *pyv;
  int* rmax;
};

size_t scalar_ws_bytes(int H, int W, size_t ts) {
  const size_t n = (size_t)H * W;
  return 3 * al256(n * ts) + 2 * al256(n * 8) + 256;
}

size_t vel_ws_bytes(int H, int W, size_t ts) {
  const size_t nu = (size_t)H * (W + 1), nv = (size_t)(H + 1) * W;
  return 3 * al256(nu * ts) + 3 * al256(nv * ts) + 2 * al256(nu * 8) + 2 * al256(nv * 8) + 256;
}

template <typename P>
P* carve(unsigned char*& cur, size_t n) {
  P* p = reinterpret_cast<P*>(cur);
  cur += al256(n * sizeof(P));
  return p;
}

template <typename T>
ScalarWs<T> carve_scalar(void* ws, int H, int W) {
  unsigned char* c = static_cast<unsigned char*>(ws);
  const size_t n = (size_t)H * W;
  ScalarWs<T> s;
  s.ym = carve<T>(c, n);
  s.gx = carve<T>(c, n);
  s.gy = carve<T>(c, n);
  s.px = carve<double>(c, n);
  s.py = carve<double>(c, n);
  s.rmax = reinterpret_cast<int*>(c);
  return s;
}

template <typename T>
VelWs<T> carve_vel(void* ws, int H, int W) {
  unsigned char* c = static_cast<unsigned char*>(ws);
  const size_t nu = (size_t)H * (W + 1), nv = (size_t)(H + 1) * W;
  VelWs<T> s;
  s.yu = carve<T>(c, nu);
  s.sxu = carve<T>(c, nu);
  s.syu = carve<T>(c, nu);
  s.yv = carve<T>(c, nv);
  s.sxv = carve<T>(c, nv);
  s.syv = carve<T>(c, nv);
  s.pxu = carve<double>(c, nu);
  s.pyu = carve<double>(c, nu);
  s.pxv = carve<double>(c, nv);
  s.pyv = carve<double>(c, nv);
  s.rmax = reinterpret_cast<int*>(c);
  return s;
}

template <typename T>
void fill(T* p, size_t n, T val, hipStream_t s) {
  const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill<T>, dim3(blocks ? blocks : 1), dim3(256), 0, s, p, n, val);
}

template <typename T>
int scalar_adjoint(const T* dbar, const T* src, const T* u, const T* v, const uint8_t* fluid,
                   int H, int W, double dt, double inv, T* sbar, T* ubar, T* vbar, int mode,
                   void* ws, hipStream_t s) {
  ScalarWs<T> w = carve_scalar<T>(ws, H, W);
  const bool scatter = mode == SMK_ADJ_SCATTER;
  fill<int>(w.rmax, 1, 0, s);
  if (scatter) fill<T>(sbar, (size_t)H * W, (T)0, s);
  hipLaunchKernelGGL(k_adj_scalar_sens<T>, agrid(H, W), dim3(AX, AY), 0, s, dbar, fluid, src, u,
                     v, H, W, dt, inv, w.ym, w.px, w.py, w.gx, w.gy, w.rmax, scatter ? 1 : 0,
                     sbar);
  int rc = check_launch("k_adj_scalar_sens");
  if (rc) return rc;
  if (!scatter || ubar != nullptr || vbar != nullptr) {
    auto g = agrid(H + 1, W + 1);
    if (scatter)
      hipLaunchKernelGGL((k_adj_scalar_gather<T, false>), g, dim3(AX, AY), 0, s, w.ym, w.px, w.py,
                         w.gx, w.gy, H, W, w.rmax, sbar, ubar, vbar);
    else
      hipLaunchKernelGGL((k_adj_scalar_gather<T, true>), g, dim3(AX, AY), 0, s, w.ym, w.px, w.py,
                         w.gx, w.gy, H, W, w.rmax, sbar, ubar, vbar);
    rc = check_launch("k_adj_scalar_gather");
  }
  return rc;
}

template <typename T>
int velocity_adjoint(const T* udbar, const T* vdbar, const T* us, const T* vs, const uint8_t* fmu,
                     const uint8_t* fmv, int H, int W, double dt, double inv, T* usbar, T* vsbar,
                     int mode, void* ws, hipStream_t s) {
  VelWs<T> w = carve_vel<T>(ws, H, W);
  const bool scatter = mode == SMK_ADJ_SCATTER;
  fill<int>(w.rmax, 1, 0, s);
  if (scatter) {
    fill<T>(usbar, (size_t)H * (W + 1), (T)0, s);
    fill<T>(vsbar, (size_t)(H + 1) * W, (T)0, s);
  }
  auto g = agrid(H + 1, W + 1);
  hipLaunchKernelGGL(k_adj_vel_sens<T>, g, dim3(AX, AY), 0, s, udbar, vdbar, fmu, fmv, us, vs, H,
                     W, dt, inv, w.yu, w.pxu, w.pyu, w.sxu, w.syu, w.yv, w.pxv, w.pyv, w.sxv,
                     w.syv, w.rmax, scatter ? 1 : 0, usbar, vsbar);
  int rc = check_launch("k_adj_vel_sens");
  if (rc) return rc;
  if (scatter)
    hipLaunchKernelGGL((k_adj_vel_gather<T, false>), g, dim3(AX, AY), 0, s, w.yu, w.pxu, w.pyu,
                       w.sxu, w.syu, w.yv, w.pxv, w.pyv, w.sxv, w.syv, H, W, w.rmax, usbar, vsbar);
  else
    hipLaunchKernelGGL((k_adj_vel_gather<T, true>), g, dim3(AX, AY), 0, s, w.yu, w.pxu, w.pyu,
                       w.sxu, w.syu, w.yv, w.pxv, w.pyv, w.sxv, w.syv, H, W, w.rmax, usbar, vsbar);
  return check_launch("k_adj_vel_gather");
}

bool clip_box(int H, int W, int& x0, int& y0, int& x1, int& y1) {
  x0 = std::max(x0, 0);
  y0 = std::max(y0, 0);
  x1 = std::min(x1, W);
  y1 = std::min(y1, H);
  return x1 > x0 && y1 > y0;
}

}  // namespace
}  // namespace smk

using namespace smk;

#define SMK_ADJ_DISPATCH(dtype, CALL_F64, CALL_F32)                           \
  do {                                                                       \
    if ((dtype) == SMK_F64) {                                                \
      CALL_F64;                                                              \
    } else if ((dtype) == SMK_F32) {                                         \
      CALL_F32;                                                              \
    } else {                                                                 \
      return fail(SMK_EINVAL, "dtype must be SMK_F64 (0) or SMK_F32 (1), got %d", (dtype)); \
    }                                                                        \
  } while (0)

extern "C" {

size_t smk_adjoint_workspace_bytes(int H, int W, int dtype) {
  if (H <= 0 || W <= 0 || (dtype != SMK_F64 && dtype != SMK_F32)) return 0;
  const size_t ts = dtype == SMK_F64 ? 8 : 4;
  return std::max(scalar_ws_bytes(H, W, ts), vel_ws_bytes(H, W, ts));
}

int smk_box_set_indexed(void* field, int H, int W, int x0, int y0, int x1, int y1,
                        const void* values, int index, int dtype, void* stream) {
  SMK_REQUIRE(field && values && H > 0 && W > 0 && index >= 0,
              "smk_box_set_indexed: bad arguments");
  if (!clip_box(H, W, x0, y0, x1, y1)) return SMK_OK;
  auto g = dim3(cdiv(x1 - x0, AX), cdiv(y1 - y0, AY));
  auto s = as_stream(stream);
  SMK_ADJ_DISPATCH(dtype,
                   hipLaunchKernelGGL(k_box_set_indexed<double>, g, dim3(AX, AY), 0, s,
                                      (double*)field, W, x0, y0, x1, y1, (const double*)values,
                                      index),
                   hipLaunchKernelGGL(k_box_set_indexed<float>, g, dim3(AX, AY), 0, s,
                                      (float*)field, W, x0, y0, x1, y1, (const float*)values,
                                      index));
  return check_launch("k_box_set_indexed");
}

int smk_box_set_adjoint(void* field_bar, int H, int W, int x0, int y0, int x1, int y1,
                        void* values_bar, int index, int dtype, void* stream) {
  SMK_REQUIRE(field_bar && values_bar && H > 0 && W > 0 && index >= 0,
              "smk_box_set_adjoint: bad arguments");
  if (!clip_box(H, W, x0, y0, x1, y1)) return SMK_OK;
  auto s = as_stream(stream);
  SMK_ADJ_DISPATCH(dtype,
                   hipLaunchKernelGGL((k_box_set_adj<256, double>), dim3(1), dim3(256), 0, s,
                                      (double*)field_bar, W, x0, y0, x1, y1, (double*)values_bar,
                                      index),
                   hipLaunchKernelGGL((k_box_set_adj<256, float>), dim3(1), dim3(256), 0, s,
                                      (float*)field_bar, W, x0, y0, x1, y1, (float*)values_bar,
                                      index));
  return check_launch("k_box_set_adj");
}

int smk_apply_buoyancy(void* v, const void* rho, int H, int W, double dt, double gravity,
                       double density_coefficient, int dtype, void* stream) {
  SMK_REQUIRE(v && rho && H > 0 && W > 0, "smk_apply_buoyancy: bad arguments");
  auto g = agrid(H, W);
  auto s = as_stream(stream);
  SMK_ADJ_DISPATCH(dtype,
                   hipLaunchKernelGGL(k_buoyancy<double>, g, dim3(AX, AY), 0, s, (double*)v,
                                      (const double*)rho, H, W, dt, gravity, density_coefficient),
                   hipLaunchKernelGGL(k_buoyancy<float>, g, dim3(AX, AY), 0, s, (float*)v,
                                      (const float*)rho, H, W, dt, gravity, density_coefficient));
  return check_launch("k_buoyancy");
}

int smk_apply_buoyancy_adjoint(const void* v_bar, void* rho_bar, int H, int W, double dt,
                               double gravity, double density_coefficient, int dtype,
                               void* stream) {
  SMK_REQUIRE(v_bar && rho_bar && H > 0 && W > 0, "smk_apply_buoyancy_adjoint: bad arguments");
  const double scale = dt * (-density_coefficient * gravity);
  auto g = agrid(H, W);
  auto s = as_stream(stream);
  SMK_ADJ_DISPATCH(dtype,
                   hipLaunchKernelGGL(k_buoyancy_adj<double>, g, dim3(AX, AY), 0, s,
                                      (const double*)v_bar, (double*)rho_bar, H, W, scale),
                   hipLaunchKernelGGL(k_buoyancy_adj<float>, g, dim3(AX, AY), 0, s,
                                      (const float*)v_bar, (float*)rho_bar, H, W, scale));
  return check_launch("k_buoyancy_adj");
}

int smk_advect_scalar_adjoint(const void* dst_bar, const void* src, const void* u, const void* v,
                              const uint8_t* is_fluid, int H, int W, double dt, double cell_size,
                              void* src_bar, void* u_bar, void* v_bar, int mode, void* workspace,
                              size_t workspace_bytes, int dtype, void* stream) {
  SMK_REQUIRE(dst_bar && src && u && v && is_fluid && src_bar && workspace,
              "smk_advect_scalar_adjoint: null pointer");
  SMK_REQUIRE(H > 0 && W > 0, "smk_advect_scalar_adjoint: bad shape");
  SMK_REQUIRE((u_bar == nullptr) == (v_bar == nullptr),
              "smk_advect_scalar_adjoint: u_bar and v_bar must both be given or both NULL");
  SMK_REQUIRE(mode == SMK_ADJ_GATHER || mode == SMK_ADJ_SCATTER,
              "smk_advect_scalar_adjoint: mode must be SMK_ADJ_GATHER or SMK_ADJ_SCATTER");
  SMK_REQUIRE(cell_size != 0.0, "smk_advect_scalar_adjoint: cell_size must be non-zero");
  SMK_REQUIRE(workspace_bytes >= smk_adjoint_workspace_bytes(H, W, dtype),
              "smk_advect_scalar_adjoint: workspace too small (%zu < %zu bytes)", workspace_bytes,
              smk_adjoint_workspace_bytes(H, W, dtype));
  const double inv = 1.0 / cell_size;
  auto s = as_stream(stream);
  SMK_ADJ_DISPATCH(dtype,
                   return scalar_adjoint<double>((const double*)dst_bar, (const double*)src,
                                                 (const double*)u, (const double*)v, is_fluid, H,
                                                 W, dt, inv, (double*)src_bar, (double*)u_bar,
                                                 (double*)v_bar, mode, workspace, s),
                   return scalar_adjoint<float>((const float*)dst_bar, (const float*)src,
                                                (const float*)u, (const float*)v, is_fluid, H, W,
                                                dt, inv, (float*)src_bar, (float*)u_bar,
                                                (float*)v_bar, mode, workspace, s));
}

int smk_advect_velocity_adjoint(const void* ud_bar, const void* vd_bar, const void* us,
                                const void* vs, const uint8_t* fmu, const uint8_t* fmv, int H,
                                int W, double dt, double cell_size, void* us_bar, void* vs_bar,
                                int mode, void* workspace, size_t workspace_bytes, int dtype,
                                void* stream) {
  SMK_REQUIRE(ud_bar && vd_bar && us && vs && fmu && fmv && us_bar && vs_bar && workspace,
              "smk_advect_velocity_adjoint: null pointer");
  SMK_REQUIRE(H > 0 && W > 0, "smk_advect_velocity_adjoint: bad shape");
  SMK_REQUIRE(mode == SMK_ADJ_GATHER || mode == SMK_ADJ_SCATTER,
              "smk_advect_velocity_adjoint: mode must be SMK_ADJ_GATHER or SMK_ADJ_SCATTER");
  SMK_REQUIRE(cell_size != 0.0, "smk_advect_velocity_adjoint: cell_size must be non-zero");
  SMK_REQUIRE(workspace_bytes >= smk_adjoint_workspace_bytes(H, W, dtype),
              "smk_advect_velocity_adjoint: workspace too small (%zu < %zu bytes)",
              workspace_bytes, smk_adjoint_workspace_bytes(H, W, dtype));
  const double inv = 1.0 / cell_size;
  auto s = as_stream(stream);
  SMK_ADJ_DISPATCH(dtype,
                   return velocity_adjoint<double>((const double*)ud_bar, (const double*)vd_bar,
                                                   (const double*)us, (const double*)vs, fmu, fmv,
                                                   H, W, dt, inv, (double*)us_bar,
                                                   (double*)vs_bar, mode, workspace, s),
                   return velocity_adjoint<float>((const float*)ud_bar, (const float*)vd_bar,
                                                  (const float*)us, (const float*)vs, fmu, fmv,
                                                  H, W, dt, inv, (float*)us_bar, (float*)vs_bar,
                                                  mode, workspace, s));
}

}  // extern "C"
