// smk_ops.hip -- pointwise / stencil / gather operators of the smoke step.
//
// One thread per output element on 64x4 blocks: a wavefront covers 64
// consecutive columns of one row, so every field read and write is a
// coalesced row segment.  These kernels are HBM-bound (0.1-0.5 flop/byte);
// no LDS staging is needed because each input element is reused only by
// its 4-neighbours, which the L2 already serves.
#include <cmath>
#include <cstdarg>
#include <mutex>

#include "smk_batch.hpp"
#include "smk_common.hpp"

namespace smk {

static thread_local std::string g_last_error = "no error";

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return fail((int)e, "launch of %s failed: %s", what, hipGetErrorString(e));
  return SMK_OK;
}

constexpr int BX = 64, BY = 4;

inline dim3 grid2d(int rows, int cols) { return dim3(cdiv(cols, BX), cdiv(rows, BY)); }

// --------------------------------------------------------------- masks
__global__ void k_face_masks(const uint8_t* __restrict__ solid, int H, int W,
                             uint8_t* __restrict__ fmu, uint8_t* __restrict__ fmv) {
  int bx_, by_;
  xcd_tile(bx_, by_);
  const int i = bx_ * BX + threadIdx.x, j = by_ * BY + threadIdx.y;
  if (j < H && i <= W) {
    uint8_t m = 0;
    if (i >= 1 && i <= W - 1) m = !(solid[(size_t)j * W + i - 1] || solid[(size_t)j * W + i]);
    fmu[(size_t)j * (W + 1) + i] = m;
  }
  if (j <= H && i < W) {
    uint8_t m = 0;
    if (j >= 1 && j <= H - 1) m = !(solid[(size_t)(j - 1) * W + i] || solid[(size_t)j * W + i]);
    fmv[(size_t)j * W + i] = m;
  }
}

__global__ void k_is_fluid(const uint8_t* __restrict__ fmu, const uint8_t* __restrict__ fmv,
                           int H, int W, uint8_t* __restrict__ out) {
  int bx_, by_;
  xcd_tile(bx_, by_);
  const int i = bx_ * BX + threadIdx.x, j = by_ * BY + threadIdx.y;
  if (j >= H || i >= W) return;
  const size_t u0 = (size_t)j * (W + 1) + i;
  out[(size_t)j * W + i] = (fmu[u0] || fmu[u0 + 1] || fmv[(size_t)j * W + i] ||
                            fmv[(size_t)(j + 1) * W + i]) ? 1 : 0;
}

// The flags of cell (i, j) from the face masks, both fluid bits set alike.
__device__ __forceinline__ uint8_t cell_flags_at(const uint8_t* __restrict__ fmu,
                                                 const uint8_t* __restrict__ fmv, int H, int W,
                                                 int i, int j) {
  const size_t u0 = (size_t)j * (W + 1) + i;
  const uint8_t ml = fmu[u0], mr = fmu[u0 + 1];
  const uint8_t md = fmv[(size_t)j * W + i], mu = fmv[(size_t)(j + 1) * W + i];
  uint8_t f = (ml || mr || md || mu) ? (uint8_t)(F_FLUID | F_LFLUID) : (uint8_t)0;
  if (ml && i > 0) f |= F_LEFT;
  if (mr && i < W - 1) f |= F_RIGHT;
  if (md && j > 0) f |= F_DOWN;
  if (mu && j < H - 1) f |= F_UP;
  return f;
}

__global__ void k_cell_flags(const uint8_t* __restrict__ fmu, const uint8_t* __restrict__ fmv,
                             int H, int W, uint8_t* __restrict__ flags) {
  int bx_, by_;
  xcd_tile(bx_, by_);
  const int i = bx_ * BX + threadIdx.x, j = by_ * BY + threadIdx.y;
  if (j >= H || i >= W) return;
  flags[(size_t)j * W + i] = cell_flags_at(fmu, fmv, H, W, i, j);
}

// Refresh the smoother's bits (F_FLUID and the neighbour bits) from the
// current face masks and keep the level's cached F_LFLUID: the state of a
// cached multigrid level 0 whose face-mask arrays changed in place
// (solver.py:42-47 keeps the arrays by reference and is_fluid by value).
__global__ void k_refresh_flags(const uint8_t* __restrict__ fmu, const uint8_t* __restrict__ fmv,
                                int H, int W, uint8_t* __restrict__ flags) {
  int bx_, by_;
  xcd_tile(bx_, by_);
  const int i = bx_ * BX + threadIdx.x, j = by_ * BY + threadIdx.y;
  if (j >= H || i >= W) return;
  const size_t o = (size_t)j * W + i;
  const uint8_t fresh = cell_flags_at(fmu, fmv, H, W, i, j);
  flags[o] = (uint8_t)((fresh & (uint8_t)~F_LFLUID) | (flags[o] & F_LFLUID));
}

__global__ void k_coarsen_solid(const uint8_t* __restrict__ sf, int H, int W,
                                uint8_t* __restrict__ sc) {
  const int Hc = H / 2, Wc = W / 2;
  const int I = blockIdx.x * BX + threadIdx.x, J = blockIdx.y * BY + threadIdx.y;
  if (J >= Hc || I >= Wc) return;
  const size_t a = (size_t)(2 * J) * W + 2 * I;
  sc[(size_t)J * Wc + I] = (sf[a] || sf[a + 1] || sf[a + W] || sf[a + W + 1]) ? 1 : 0;
}

// ----------------------------------------------------------- advection
template <typename T>
__global__ void k_advect_scalar(const T* __restrict__ src, T* __restrict__ dst, int H, int W,
                                const T* __restrict__ u, int uh, int uw,
                                const T* __restrict__ v, int vh, int vw,
                                const uint8_t* __restrict__ fluid, double dt, double inv,
                                double scale, int apply_scale) {
  int bx_, by_;
  xcd_tile(bx_, by_);
  const int i = bx_ * BX + threadIdx.x, j = by_ * BY + threadIdx.y;
  if (j >= H || i >= W) return;
  const size_t o = (size_t)j * W + i;
  // the mask is read with the velocity samples, not before them: every cell
  // is traced (all sample indices are clamped into the fields) and the store
  // selects, so the loads of a lane start together
  const bool fl = fluid[o] != 0;
  const double px = i + 0.5, py = j + 0.5;
  // u at (i + 1/2, j) and v at (i, j + 1/2): floors and ceilings known
  const double uu = sample_known(u, uh, uw, px, py - 0.5, i, i + 1, j, j);
  const double vv = sample_known(v, vh, vw, px - 0.5, py, i, i, j, j + 1);
  const double x = px - (uu * dt * inv);
  const double y = py - (vv * dt * inv);
  T r = (T)sample_field_rows(src, H, W, x - 0.5, y - 0.5);
  if (apply_scale) r = r * (T)scale;
  dst[o] = fl ? r : (T)0;
}

template <typename T>
__device__ __forceinline__ void advect_u_at(const T* __restrict__ us, T* __restrict__ ud, int uh,
                                            int uw, const T* __restrict__ v, int vh, int vw,
                                            const uint8_t* __restrict__ fmu, double dt,
                                            double inv, int j, int i) {
  const size_t o = (size_t)j * uw + i;
  // as in k_advect_scalar: every face is traced, the mask selects the store
  const bool open = fmu[o] != 0;
  const double px = (double)i, py = j + 0.5;
  const double uu = (double)us[o];
  const double vv = sample_known(v, vh, vw, px - 0.5, py, i - 1, i, j, j + 1);
  const double x = px - (uu * dt * inv);
  const double y = py - (vv * dt * inv);
  const T r = (T)sample_field(us, uh, uw, x, y - 0.5);
  ud[o] = open ? r : (T)0;
}

template <typename T>
__device__ __forceinline__ void advect_v_at(const T* __restrict__ vs, T* __restrict__ vd, int vh,
                                            int vw, const T* __restrict__ u, int uh, int uw,
                                            const uint8_t* __restrict__ fmv, double dt,
                                            double inv, int j, int i) {
  const size_t o = (size_t)j * vw + i;
  const bool open = fmv[o] != 0;
  const double px = i + 0.5, py = (double)j;
  const double uu = sample_known(u, uh, uw, px, py - 0.5, i, i + 1, j - 1, j);
  const double vv = (double)vs[o];
  const double x = px - (uu * dt * inv);
  const double y = py - (vv * dt * inv);
  const T r = (T)sample_field(vs, vh, vw, x - 0.5, y);
  vd[o] = open ? r : (T)0;
}

// The u face (j, i) and the v face (j, i) of one thread in straight-line
// code: both faces' first loads are issued before either face computes and
// both traced samples before either store, so a thread waits on memory
// twice, not four times.  The caller guarantees j <= H and i <= W; a thread
// on the top row (no u face) or right column (no v face) traces a clamped
// face and does not store it.  Per face, the same operations as advect_u_at
// and advect_v_at.
template <typename T>
__device__ __forceinline__ void advect_uv_at(const T* __restrict__ us, T* __restrict__ ud,
                                             const T* __restrict__ vs, T* __restrict__ vd,
                                             const uint8_t* __restrict__ fmu,
                                             const uint8_t* __restrict__ fmv, int H, int W,
                                             double dt, double inv, int j, int i) {
  const bool su = j < H, sv = i < W;
  const int ju = su ? j : H - 1, iv = sv ? i : W - 1;
  const int uw = W + 1, vh = H + 1;
  const size_t ou = (size_t)ju * uw + i, ov = (size_t)j * W + iv;
  const bool open_u = fmu[ou] != 0, open_v = fmv[ov] != 0;
  const double pxu = (double)i, pyu = ju + 0.5;  // u face
  const double pxv = iv + 0.5, pyv = (double)j;  // v face
  const double uu_u = (double)us[ou];
  const double vv_u = sample_known(vs, vh, W, pxu - 0.5, pyu, i - 1, i, ju, ju + 1);
  const double uu_v = sample_known(us, H, uw, pxv, pyv - 0.5, iv, iv + 1, j - 1, j);
  const double vv_v = (double)vs[ov];
  const double xu = pxu - (uu_u * dt * inv), yu = pyu - (vv_u * dt * inv);
  const double xv = pxv - (uu_v * dt * inv), yv = pyv - (vv_v * dt * inv);
  const T ru = (T)sample_field(us, H, uw, xu, yu - 0.5);
  const T rv = (T)sample_field(vs, vh, W, xv - 0.5, yv);
  if (su) ud[ou] = open_u ? ru : (T)0;
  if (sv) vd[ov] = open_v ? rv : (T)0;
}

template <typename T>
__global__ void k_advect_u(const T* __restrict__ us, T* __restrict__ ud, int uh, int uw,
                           const T* __restrict__ v, int vh, int vw,
                           const uint8_t* __restrict__ fmu, double dt, double inv) {
  int bx_, by_;
  xcd_tile(bx_, by_);
  const int i = bx_ * BX + threadIdx.x, j = by_ * BY + threadIdx.y;
  if (j < uh && i < uw) advect_u_at(us, ud, uh, uw, v, vh, vw, fmu, dt, inv, j, i);
}

template <typename T>
__global__ void k_advect_v(const T* __restrict__ vs, T* __restrict__ vd, int vh, int vw,
                           const T* __restrict__ u, int uh, int uw,
                           const uint8_t* __restrict__ fmv, double dt, double inv) {
  int bx_, by_;
  xcd_tile(bx_, by_);
  const int i = bx_ * BX + threadIdx.x, j = by_ * BY + threadIdx.y;
  if (j < vh && i < vw) advect_v_at(vs, vd, vh, vw, u, uh, uw, fmv, dt, inv, j, i);
}

// u and v faces of an H x W MAC grid in one launch over (H+1) x (W+1).
// A wavefront is one row segment of 64 consecutive columns, so the
// staggered averages take the neighbouring column from the adjacent lane
// (DPP): a thread loads u at rows jr and jm and v at rows j and jr + 1 of
// its own column, and lanes 0 and 63 load the one column beyond the
// wavefront themselves.  Per face these are the values, weights and lerps
// of advect_uv_at (the corner indices below are sample_known's clamped
// ones wherever the face is stored).  Every lane of the padded grid stays
// active, as DPP needs: indices are clamped into the grid and only faces
// inside it are stored.
template <typename T>
__global__ __launch_bounds__(BX * BY) void k_advect_velocity(
    const T* __restrict__ us, T* __restrict__ ud, const T* __restrict__ vs, T* __restrict__ vd,
    const uint8_t* __restrict__ fmu, const uint8_t* __restrict__ fmv, int H, int W, double dt,
    double inv) {
  static_assert(BX == 64, "a wavefront is one 64-column row segment");
  int bx_, by_;
  xcd_tile(bx_, by_);
  const int lane = threadIdx.x;
  const int gi = bx_ * BX + lane, gj = by_ * BY + threadIdx.y;
  const bool su = gj < H && gi <= W;       // u face (gj, gi) exists
  const bool sv = gj <= H && gi < W;       // v face (gj, gi) exists
  const int i = gi < W ? gi : W;           // u column, in [0, W]
  const int j = gj < H ? gj : H;           // v row, in [0, H]
  const int jr = j < H ? j : H - 1;        // u row of the u face
  const int jm = j > 0 ? j - 1 : 0;        // lower u row of the v face's average
  const int ic = i < W ? i : W - 1;        // v column of the v face
  const int uw = W + 1;
  const size_t ou = (size_t)jr * uw + i, ov = (size_t)j * W + ic;
  const T A = us[ou];                           // u(jr, i)
  const T D = us[(size_t)jm * uw + i];          // u(jm, i)
  const T B = vs[ov];                           // v(j, ic) = v(jr, ic) where su
  const T C = vs[(size_t)(jr + 1) * W + ic];    // v(jr + 1, ic)
  const bool open_u = fmu[ou] != 0, open_v = fmv[ov] != 0;
  // v at column i - 1 (clamped) for the u face, u at column i + 1 for the v
  // face: loaded by the wavefront's edge lanes (issued with the loads above,
  // before anything waits), from the neighbouring lane elsewhere
  T Be = (T)0, Ce = (T)0, Ae = (T)0, De = (T)0;
  const bool left_load = lane == 0 && i > 0, right_load = lane == 63 && i < W;
  if (left_load) {
    Be = vs[(size_t)jr * W + i - 1];
    Ce = vs[(size_t)(jr + 1) * W + i - 1];
  }
  if (right_load) {
    Ae = us[(size_t)jr * uw + i + 1];
    De = us[(size_t)jm * uw + i + 1];
  }
  const T Bn = from_left(B), Cn = from_left(C), An = from_right(A), Dn = from_right(D);
  const T Bl = left_load ? Be : (lane == 0 ? B : Bn);
  const T Cl = left_load ? Ce : (lane == 0 ? C : Cn);
  const T Ar = right_load ? Ae : An;
  const T Dr = right_load ? De : Dn;
  const double pxu = (double)i, pyu = jr + 0.5;  // u face
  const double pxv = ic + 0.5, pyv = (double)j;  // v face
  const int x0u = i > 0 ? i - 1 : 0;
  const double vv_u = lerp4(Bl, B, Cl, C, (pxu - 0.5) - (double)x0u, pyu - (double)jr);
  const double uu_v = lerp4(D, Dr, A, Ar, pxv - (double)ic, (pyv - 0.5) - (double)jm);
  const double uu_u = (double)A, vv_v = (double)B;
  const double xu = pxu - (uu_u * dt * inv), yu = pyu - (vv_u * dt * inv);
  const double xv = pxv - (uu_v * dt * inv), yv = pyv - (vv_v * dt * inv);
  const T ru = (T)sample_field_rows(us, H, uw, xu, yu - 0.5);
  const T rv = (T)sample_field_rows(vs, H + 1, W, xv - 0.5, yv);
  if (su) ud[ou] = open_u ? ru : (T)0;
  if (sv) vd[ov] = open_v ? rv : (T)0;
}

// --------------------------------------------------------------- forces
// forces.py:19-46 / :48-76 / :105.  The force buffers are float32 in the
// reference, so each force is rounded to float, summed in float, then
// widened: v += dt * (float)(bf + tf).
template <typename T>
__global__ void k_apply_forces(T* __restrict__ v, int vh, int vw,
                               const T* __restrict__ rho, int H, int W,
                               const T* __restrict__ temp, int Ht, int Wt, double dt,
                               double g, double bc, double dc, double Ta) {
  int bx_, by_;
  xcd_tile(bx_, by_);
  const int i = bx_ * BX + threadIdx.x, j = by_ * BY + threadIdx.y;
  if (j >= vh || i >= vw) return;
  // every operand loaded (rows clamped into the fields) before the edge
  // cases select, so a lane's loads start together
  const size_t o = (size_t)j * vw + i;
  const T vo = v[o];
  float bf = 0.0f, tf = 0.0f;
  const int ic = i < W ? i : W - 1;
  const T r0 = rho[(size_t)(j - 1 < 0 ? 0 : (j - 1 < H ? j - 1 : H - 1)) * W + ic];
  const T r1 = rho[(size_t)(j < H ? j : H - 1) * W + ic];
  if (j < H && i < W) {
    double d;
    if (j == 0 || j == H - 1) {
      d = (double)r1;
    } else {
      const T s = r0 + r1;
      d = 0.5 * (double)s;
    }
    bf = (float)(-dc * d * g);
  }
  if (j <= Ht && i < Wt) {
    double t;
    if (temp == nullptr) {
      t = 0.0;
    } else {
      const T t0 = temp[(size_t)(j - 1 < 0 ? 0 : (j - 1 < Ht ? j - 1 : Ht - 1)) * Wt + i];
      const T t1 = temp[(size_t)(j < Ht ? j : Ht - 1) * Wt + i];
      if (j == 0) {
        t = (double)t1;
      } else if (j == Ht) {
        t = (double)t0;
      } else {
        const T s = t0 + t1;
        t = 0.5 * (double)s;
      }
    }
    tf = (float)(-bc * (t - Ta));
  }
  const float s = bf + tf;
  v[o] = (T)((double)vo + dt * (double)s);
}

template <typename T>
__global__ void k_cool(T* __restrict__ t, long n, double ambient, double decay) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  t[k] = (T)(ambient + ((double)t[k] - ambient) * decay);
}

template <typename T>
__global__ void k_scale(T* __restrict__ x, long n, T f) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) x[k] = x[k] * f;
}

// ----------------------------------------------------------- projection
// Numba typing: velocity*mask and the difference stay in the array type, the
// 1/h product and the sum are float64 (inv_cell_size is a Python float).
template <typename T>
__global__ void k_divergence(const T* __restrict__ u, const T* __restrict__ v,
                             T* __restrict__ div, const uint8_t* __restrict__ fmu,
                             const uint8_t* __restrict__ fmv, int H, int W, double inv) {
  int bx_, by_;
  xcd_tile(bx_, by_);
  const int i = bx_ * BX + threadIdx.x, j = by_ * BY + threadIdx.y;
  if (j >= H || i >= W) return;
  const size_t a = (size_t)j * (W + 1) + i;
  const size_t b = (size_t)j * W + i;
  const T ddu = (u[a + 1] * (T)fmu[a + 1]) - (u[a] * (T)fmu[a]);
  const T ddv = (v[b + W] * (T)fmv[b + W]) - (v[b] * (T)fmv[b]);
  const double du = (double)ddu * inv;
  const double dv = (double)ddv * inv;
  div[b] = (T)(du + dv);
}

// pressure_gradient.py:4-37 (+ boundaries.py:7-39 when solid != nullptr).
template <typename T>
__global__ void k_pressure_gradient(T* __restrict__ u, T* __restrict__ v,
                                    const T* __restrict__ p,
                                    const uint8_t* __restrict__ fmu,
                                    const uint8_t* __restrict__ fmv,
                                    const uint8_t* __restrict__ solid, int H, int W,
                                    double scale) {
  int bx_, by_;
  xcd_tile(bx_, by_);
  const int i = bx_ * BX + threadIdx.x, j = by_ * BY + threadIdx.y;
  // Both faces of the thread in straight-line code: every operand of the u
  // face (j, i) and the v face (j, i) is loaded, at indices clamped into the
  // arrays, before any is tested, so a lane waits on memory once; masks and
  // walls select, and a face outside the grid (u on the top row, v on the
  // right column) is computed on clamped indices and not stored.  The two
  // faces share the pressure p(j, i) and the wall flag solid(j, i).
  if (j > H || i > W) return;
  const int jc = j < H ? j : H - 1;          // row of the u face
  const int ic = i < W ? i : W - 1;          // column of the v face
  const int il = i - 1 < 0 ? W - 1 : i - 1;  // index -1 wraps (Python)
  const int jb = j - 1 < 0 ? H - 1 : j - 1;
  const size_t ou = (size_t)jc * (W + 1) + i, ov = (size_t)j * W + ic;
  const bool open_u = fmu[ou] != 0, open_v = fmv[ov] != 0;
  const T uo = u[ou], vo = v[ov];
  const T pc = p[(size_t)jc * W + ic];       // p(j, ir) of the u face = p(jt, i) of the v face
  const T pl = p[(size_t)jc * W + il];
  const T pb = p[(size_t)jb * W + ic];
  bool wall_u = false, wall_v = false;
  if (solid != nullptr) {
    const uint8_t sc = solid[(size_t)jc * W + ic];
    const uint8_t sl = solid[(size_t)jc * W + (i > 0 ? i - 1 : 0)];
    const uint8_t sb = solid[(size_t)(j > 0 ? j - 1 : 0) * W + ic];
    wall_u = (i > 0 && sl == 1) || (i < W && sc == 1);
    wall_v = (j > 0 && sb == 1) || (j < H && sc == 1);
  }
  const T dpu = pc - pl, dpv = pc - pb;
  T val_u = open_u ? (T)((double)uo - scale * (double)dpu) : (T)0;
  T val_v = open_v ? (T)((double)vo - scale * (double)dpv) : (T)0;
  if (wall_u) val_u = (T)0;
  if (wall_v) val_v = (T)0;
  if (j < H) u[ou] = val_u;
  if (i < W) v[ov] = val_v;
}

template <typename T>
__global__ void k_enforce_boundaries(T* __restrict__ u, T* __restrict__ v,
                                     const uint8_t* __restrict__ solid, int H, int W) {
  int bx_, by_;
  xcd_tile(bx_, by_);
  const int i = bx_ * BX + threadIdx.x, j = by_ * BY + threadIdx.y;
  if (j < H && i <= W) {
    const bool ls = i > 0 ? solid[(size_t)j * W + i - 1] == 1 : false;
    const bool rs = i < W ? solid[(size_t)j * W + i] == 1 : false;
    if (ls || rs) u[(size_t)j * (W + 1) + i] = (T)0;
  }
  if (j <= H && i < W) {
    const bool ts = j > 0 ? solid[(size_t)(j - 1) * W + i] == 1 : false;
    const bool bs = j < H ? solid[(size_t)j * W + i] == 1 : false;
    if (ts || bs) v[(size_t)j * W + i] = (T)0;
  }
}

// residual.py:6-42, reading the three masks directly.
template <typename T>
__global__ void k_residual_masks(const T* __restrict__ p, const T* __restrict__ div,
                                 const uint8_t* __restrict__ fluid,
                                 const uint8_t* __restrict__ fmu,
                                 const uint8_t* __restrict__ fmv, int H, int W,
                                 double alpha, double beta, T* __restrict__ r) {
  int bx_, by_;
  xcd_tile(bx_, by_);
  const int i = bx_ * BX + threadIdx.x, j = by_ * BY + threadIdx.y;
  if (j >= H || i >= W) return;
  const size_t o = (size_t)j * W + i;
  const size_t a = (size_t)j * (W + 1) + i;
  // every operand loaded (clamped indices) before the masks are tested
  const bool fl = fluid[o] != 0;
  const uint8_t mu0 = fmu[a], mu1 = fmu[a + 1], mv0 = fmv[o], mv1 = fmv[o + W];
  const T c = p[o];
  const T pl = p[(size_t)j * W + (i - 1 > 0 ? i - 1 : 0)];
  const T pr = p[(size_t)j * W + (i + 1 < W - 1 ? i + 1 : W - 1)];
  const T pd = p[(size_t)(j - 1 > 0 ? j - 1 : 0) * W + i];
  const T pu = p[(size_t)(j + 1 < H - 1 ? j + 1 : H - 1) * W + i];
  const T dv = div[o];
  const T left = mu0 ? pl : c;
  const T right = mu1 ? pr : c;
  const T down = mv0 ? pd : c;
  const T up = mv1 ? pu : c;
  // neighbour sum in the array type, alpha/beta terms in float64 (Numba typing)
  const T nsum = left + right + down + up;
  r[o] = fl ? (T)(((double)nsum + alpha * (double)dv) - beta * (double)c) : (T)0;
}

// The step's metrics pass (simulation.py:251-265): compute_divergence of the
// projected velocity, then compute_residual of p against that divergence, in
// one launch.  A cell's residual needs only its own divergence, so each
// thread forms div exactly as k_divergence does, stores it and uses the
// stored value exactly as k_residual_masks does: both outputs equal the two
// separate launches', bit for bit.
template <typename T>
__global__ void k_div_residual(const T* __restrict__ u, const T* __restrict__ v,
                               const T* __restrict__ p, const uint8_t* __restrict__ fluid,
                               const uint8_t* __restrict__ fmu, const uint8_t* __restrict__ fmv,
                               int H, int W, double inv, double alpha, double beta,
                               T* __restrict__ div, T* __restrict__ r) {
  int bx_, by_;
  xcd_tile(bx_, by_);
  const int i = bx_ * BX + threadIdx.x, j = by_ * BY + threadIdx.y;
  if (j >= H || i >= W) return;
  const size_t a = (size_t)j * (W + 1) + i;
  const size_t o = (size_t)j * W + i;
  const uint8_t mu0 = fmu[a], mu1 = fmu[a + 1], mv0 = fmv[o], mv1 = fmv[o + W];
  const T ddu = (u[a + 1] * (T)mu1) - (u[a] * (T)mu0);
  const T ddv = (v[o + W] * (T)mv1) - (v[o] * (T)mv0);
  const T d = (T)((double)ddu * inv + (double)ddv * inv);
  div[o] = d;
  // the pressures are loaded (clamped indices) with everything else, before
  // the masks are tested, so a lane's loads start together
  const bool fl = fluid[o] != 0;
  const T c = p[o];
  const T pl = p[(size_t)j * W + (i - 1 > 0 ? i - 1 : 0)];
  const T pr = p[(size_t)j * W + (i + 1 < W - 1 ? i + 1 : W - 1)];
  const T pd = p[(size_t)(j - 1 > 0 ? j - 1 : 0) * W + i];
  const T pu = p[(size_t)(j + 1 < H - 1 ? j + 1 : H - 1) * W + i];
  const T left = mu0 ? pl : c;
  const T right = mu1 ? pr : c;
  const T down = mv0 ? pd : c;
  const T up = mv1 ? pu : c;
  const T nsum = left + right + down + up;
  r[o] = fl ? (T)(((double)nsum + alpha * (double)d) - beta * (double)c) : (T)0;
}

// ------------------------------------------------------------- emitters
// emitters.py:339-411 falloffs.
__device__ __forceinline__ double falloff(double d, double r, int approach) {
  switch (approach) {
    case 0: return d < r ? 1.0 : 0.0;
    case 1: return d >= r ? 0.0 : 1.0 - (d / r);
    case 2: {
      if (d >= r) return 0.0;
      const double q = d / r;
      return 1.0 - q * q;
    }
    case 3: {
      if (d >= r) return 0.0;
      const double sigma = r / 3.0;
      const double q = d / sigma;
      return exp(-0.5 * (q * q));
    }
    default: return 0.0;
  }
}

template <typename T>
__global__ void k_emit(T* __restrict__ f, int H, int W, int x0, int y0, int x1, int y1,
                       double value, double dt, double px, double py, double radius,
                       int approach) {
  const int x = x0 + blockIdx.x * BX + threadIdx.x, y = y0 + blockIdx.y * BY + threadIdx.y;
  if (x >= x1 || y >= y1) return;
  const double dx = x - px, dy = y - py;
  const double d = sqrt(dx * dx + dy * dy);
  const double fo = falloff(d, radius, approach);
  const size_t o = (size_t)y * W + x;
  f[o] = (T)((double)f[o] + value * fo * dt);
}

template <typename T>
__global__ void k_box(T* __restrict__ f, int W, int x0, int y0, int x1, int y1, T value,
                      int mode) {
  const int x = x0 + blockIdx.x * BX + threadIdx.x, y = y0 + blockIdx.y * BY + threadIdx.y;
  if (x >= x1 || y >= y1) return;
  const size_t o = (size_t)y * W + x;
  f[o] = mode == 0 ? value : f[o] + value;
}

// Up to BOX_MAX pairwise-disjoint boxes in one launch: blockIdx.z is the
// box, each block a 64 x 4 piece of it (the grid spans the largest box).
constexpr int BOX_MAX = 8;
struct BoxList {
  int x0[BOX_MAX], y0[BOX_MAX], x1[BOX_MAX], y1[BOX_MAX], mode[BOX_MAX];
  double value[BOX_MAX];
};

template <typename T>
__global__ void k_boxes(T* __restrict__ f, int W, const BoxList b) {
  const int k = blockIdx.z;
  const int x = b.x0[k] + blockIdx.x * BX + threadIdx.x;
  const int y = b.y0[k] + blockIdx.y * BY + threadIdx.y;
  if (x >= b.x1[k] || y >= b.y1[k]) return;
  const size_t o = (size_t)y * W + x;
  const T value = (T)b.value[k];
  f[o] = b.mode[k] == 0 ? value : f[o] + value;
}

// ------------------------------------------------------ batched transport
// The same per-face / per-cell arithmetic as the kernels above, for many
// scenes in one persistent launch: workgroups walk a packed job list of
// 64 x 4 tiles ({scene, tile x, tile y, -}); a thread owns one position of
// the tile.  Face jobs tile the (H+1) x (W+1) face superset of their scene.

__device__ __forceinline__ void job_pos(const int4& jb, int& i, int& j) {
  i = jb.y * BT_X + (int)(threadIdx.x & (BT_X - 1));
  j = jb.z * BT_Y + (int)(threadIdx.x / BT_X);
}

__device__ __forceinline__ bool in_box(const BatchScene& d, int j, int i) {
  return i >= d.box_x0 && i < d.box_x1 && j >= d.box_y0 && j < d.box_y1;
}

// Box source (density := value in the box) fused with forces.py:78-105: the
// force reads the density with the box applied, so no tile depends on
// another tile's writes.
__global__ __launch_bounds__(BT_NT) void k_b_box_forces(const BatchScene* __restrict__ sc,
                                                        const int4* __restrict__ jobs, int njobs,
                                                        const int* __restrict__ parity,
                                                        BatchPhys ph) {
  const int par = *parity;
  for (int q = blockIdx.x; q < njobs; q += gridDim.x) {
    const int4 jb = jobs[q];
    const BatchScene& d = sc[jb.x];
    const int H = d.H, W = d.W;
    int i, j;
    job_pos(jb, i, j);
    double* rho = d.rho[par];
    if (j < H && i < W && in_box(d, j, i)) rho[(size_t)j * W + i] = d.box_value;
    if (j > H || i >= W) continue;
    auto rho_at = [&](int r) -> double {
      return in_box(d, r, i) ? d.box_value : rho[(size_t)r * W + i];
    };
    float bf = 0.0f;
    if (j < H) {
      double dd;
      if (j == 0 || j == H - 1) {
        dd = rho_at(j);
      } else {
        const double s2 = rho_at(j - 1) + rho_at(j);
        dd = 0.5 * s2;
      }
      bf = (float)(-ph.density_coefficient * dd * ph.gravity);
    }
    const float tf = (float)(-ph.buoyancy * (0.0 - ph.ambient_temperature));
    const float s = bf + tf;
    const size_t o = (size_t)j * W + i;
    d.v[o] = d.v[o] + ph.dt * (double)s;
  }
}

__global__ __launch_bounds__(BT_NT) void k_b_advect_velocity(const BatchScene* __restrict__ sc,
                                                             const int4* __restrict__ jobs,
                                                             int njobs, BatchPhys ph) {
  for (int q = blockIdx.x; q < njobs; q += gridDim.x) {
    const int4 jb = jobs[q];
    const BatchScene& d = sc[jb.x];
    const int H = d.H, W = d.W;
    const double inv = 1.0 / d.h;
    int i, j;
    job_pos(jb, i, j);
    if (j <= H && i <= W)
      advect_uv_at(d.u, d.u_tmp, d.v, d.v_tmp, d.fmu, d.fmv, H, W, ph.dt, inv, j, i);
  }
}

__global__ __launch_bounds__(BT_NT) void k_b_advect_density(const BatchScene* __restrict__ sc,
                                                            const int4* __restrict__ jobs,
                                                            int njobs,
                                                            const int* __restrict__ parity,
                                                            BatchPhys ph) {
  const int par = *parity;
  for (int q = blockIdx.x; q < njobs; q += gridDim.x) {
    const int4 jb = jobs[q];
    const BatchScene& d = sc[jb.x];
    const int H = d.H, W = d.W;
    int i, j;
    job_pos(jb, i, j);
    if (j >= H || i >= W) continue;
    const size_t o = (size_t)j * W + i;
    const double* src = d.rho[par];
    double* dst = d.rho[par ^ 1];
    if (!d.fluid[o]) {
      dst[o] = 0.0;
      continue;
    }
    const double inv = 1.0 / d.h;
    const double px = i + 0.5, py = j + 0.5;
    const double uu = sample_known(d.u_tmp, H, W + 1, px, py - 0.5, i, i + 1, j, j);
    const double vv = sample_known(d.v_tmp, H + 1, W, px - 0.5, py, i, i, j, j + 1);
    const double x = px - (uu * ph.dt * inv);
    const double y = py - (vv * ph.dt * inv);
    dst[o] = sample_field(src, H, W, x - 0.5, y - 0.5);
  }
}

// divergence.py:4-25 of the advected velocity, p := 0 (simulation.py:310)
// and the ||alpha*div||^2 partial of the tile (the tile is the 64 x 4 block
// of the single-scene norm kernel, so the partials are the same numbers).
__global__ __launch_bounds__(BT_NT) void k_b_divergence_norm(const BatchScene* __restrict__ sc,
                                                             const int4* __restrict__ jobs,
                                                             int njobs,
                                                             double* __restrict__ partial) {
  __shared__ double red[BT_NT / 64];
  for (int q = blockIdx.x; q < njobs; q += gridDim.x) {
    const int4 jb = jobs[q];
    const BatchScene& d = sc[jb.x];
    const int H = d.H, W = d.W;
    const double inv = 1.0 / d.h;
    int i, j;
    job_pos(jb, i, j);
    double acc = 0.0;
    if (j < H && i < W) {
      const size_t a = (size_t)j * (W + 1) + i;
      const size_t b = (size_t)j * W + i;
      const double ddu = (d.u_tmp[a + 1] * (double)d.fmu[a + 1]) - (d.u_tmp[a] * (double)d.fmu[a]);
      const double ddv = (d.v_tmp[b + W] * (double)d.fmv[b + W]) - (d.v_tmp[b] * (double)d.fmv[b]);
      const double du = ddu * inv;
      const double dv = ddv * inv;
      const double dvg = du + dv;
      d.div[b] = dvg;
      d.p[b] = 0.0;
      if (d.flags0[b] & F_LFLUID) {
        const double bb = d.alpha0 * dvg;
        acc = bb * bb;
      }
    }
    const double tot = block_sum<BT_NT>(acc, red);
    if (threadIdx.x == 0) partial[d.part_base + jb.z * d.nbx + jb.y] = tot;
  }
}

// pressure_gradient.py:4-37 + boundaries.py:7-39 from the advected velocity
// into the state; flips the density parity for the next step.
__global__ __launch_bounds__(BT_NT) void k_b_gradient(const BatchScene* __restrict__ sc,
                                                      const int4* __restrict__ jobs, int njobs,
                                                      int* __restrict__ parity, BatchPhys ph) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *parity ^= 1;
  for (int q = blockIdx.x; q < njobs; q += gridDim.x) {
    const int4 jb = jobs[q];
    const BatchScene& d = sc[jb.x];
    const int H = d.H, W = d.W;
    const double scale = ph.dt / (ph.fluid_density * d.h);
    int i, j;
    job_pos(jb, i, j);
    const double* p = d.p;
    if (j < H && i <= W) {
      const size_t o = (size_t)j * (W + 1) + i;
      double val = 0.0;
      if (d.fmu[o]) {
        const int ir = i < W ? i : W - 1;
        const int il = i - 1 < 0 ? W - 1 : i - 1;
        const double dp = p[(size_t)j * W + ir] - p[(size_t)j * W + il];
        val = d.u_tmp[o] - scale * dp;
      }
      const bool ls = i > 0 ? d.solid[(size_t)j * W + i - 1] == 1 : false;
      const bool rs = i < W ? d.solid[(size_t)j * W + i] == 1 : false;
      d.u[o] = (ls || rs) ? 0.0 : val;
    }
    if (j <= H && i < W) {
      const size_t o = (size_t)j * W + i;
      double val = 0.0;
      if (d.fmv[o]) {
        const int jt = j < H ? j : H - 1;
        const int jb2 = j - 1 < 0 ? H - 1 : j - 1;
        const double dp = p[(size_t)jt * W + i] - p[(size_t)jb2 * W + i];
        val = d.v_tmp[o] - scale * dp;
      }
      const bool ts = j > 0 ? d.solid[(size_t)(j - 1) * W + i] == 1 : false;
      const bool bs = j < H ? d.solid[(size_t)j * W + i] == 1 : false;
      d.v[o] = (ts || bs) ? 0.0 : val;
    }
  }
}

inline unsigned batch_grid(int njobs, int blocks) {
  const int b = blocks > 0 ? blocks : 2048;
  return (unsigned)(njobs < b ? (njobs > 0 ? njobs : 1) : b);
}

int batch_box_forces(const BatchScene* sc, const int4* jobs, int njobs, const int* parity,
                     const BatchPhys& ph, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_b_box_forces, dim3(batch_grid(njobs, blocks)), dim3(BT_NT), 0, s, sc, jobs,
                     njobs, parity, ph);
  return check_launch("k_b_box_forces");
}

int batch_advect_velocity(const BatchScene* sc, const int4* jobs, int njobs, const BatchPhys& ph,
                          int blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_b_advect_velocity, dim3(batch_grid(njobs, blocks)), dim3(BT_NT), 0, s, sc,
                     jobs, njobs, ph);
  return check_launch("k_b_advect_velocity");
}

int batch_advect_density(const BatchScene* sc, const int4* jobs, int njobs, const int* parity,
                         const BatchPhys& ph, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_b_advect_density, dim3(batch_grid(njobs, blocks)), dim3(BT_NT), 0, s, sc,
                     jobs, njobs, parity, ph);
  return check_launch("k_b_advect_density");
}

int batch_divergence_rhs_norm(const BatchScene* sc, const int4* cell_jobs, int njobs,
                              double* partial, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_b_divergence_norm, dim3(batch_grid(njobs, blocks)), dim3(BT_NT), 0, s, sc,
                     cell_jobs, njobs, partial);
  return check_launch("k_b_divergence_norm");
}

int batch_gradient(const BatchScene* sc, const int4* jobs, int njobs, int* parity,
                   const BatchPhys& ph, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_b_gradient, dim3(batch_grid(njobs, blocks)), dim3(BT_NT), 0, s, sc, jobs,
                     njobs, parity, ph);
  return check_launch("k_b_gradient");
}

int refresh_cell_flags(const uint8_t* fmu, const uint8_t* fmv, int H, int W, uint8_t* flags,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_refresh_flags, grid2d(H, W), dim3(BX, BY), 0, s, fmu, fmv, H, W, flags);
  return check_launch("k_refresh_flags");
}

}  // namespace smk

using namespace smk;

#define SMK_DISPATCH(dtype, KERNEL_CALL_F64, KERNEL_CALL_F32)                 \
  do {                                                                       \
    if ((dtype) == SMK_F64) {                                                \
      KERNEL_CALL_F64;                                                       \
    } else if ((dtype) == SMK_F32) {                                         \
      KERNEL_CALL_F32;                                                       \
    } else {                                                                 \
      return fail(SMK_EINVAL, "dtype must be SMK_F64 (0) or SMK_F32 (1), got %d", (dtype)); \
    }                                                                        \
  } while (0)

extern "C" {

const char* smk_last_error(void) { return g_last_error.c_str(); }
int smk_abi_version(void) { return 6; }

int smk_face_masks_from_solid(const uint8_t* solid, int H, int W, uint8_t* fmu,
                              uint8_t* fmv, void* stream) {
  SMK_REQUIRE(solid && fmu && fmv && H > 0 && W > 0, "smk_face_masks_from_solid: bad arguments");
  hipLaunchKernelGGL(k_face_masks, grid2d(H + 1, W + 1), dim3(BX, BY), 0, as_stream(stream),
                     solid, H, W, fmu, fmv);
  return check_launch("k_face_masks");
}

int smk_is_fluid(const uint8_t* fmu, const uint8_t* fmv, int H, int W, uint8_t* out,
                 void* stream) {
  SMK_REQUIRE(fmu && fmv && out && H > 0 && W > 0, "smk_is_fluid: bad arguments");
  hipLaunchKernelGGL(k_is_fluid, grid2d(H, W), dim3(BX, BY), 0, as_stream(stream), fmu, fmv,
                     H, W, out);
  return check_launch("k_is_fluid");
}

int smk_cell_flags(const uint8_t* fmu, const uint8_t* fmv, int H, int W, uint8_t* flags,
                   void* stream) {
  SMK_REQUIRE(fmu && fmv && flags && H > 0 && W > 0, "smk_cell_flags: bad arguments");
  hipLaunchKernelGGL(k_cell_flags, grid2d(H, W), dim3(BX, BY), 0, as_stream(stream), fmu, fmv,
                     H, W, flags);
  return check_launch("k_cell_flags");
}

int smk_coarsen_solid(const uint8_t* sf, int H, int W, uint8_t* sc, void* stream) {
  SMK_REQUIRE(sf && sc && H > 1 && W > 1, "smk_coarsen_solid: bad arguments");
  hipLaunchKernelGGL(k_coarsen_solid, grid2d(H / 2, W / 2), dim3(BX, BY), 0, as_stream(stream),
                     sf, H, W, sc);
  return check_launch("k_coarsen_solid");
}

int smk_advect_scalar(const void* src, void* dst, int H, int W, const void* u, int uh, int uw,
                      const void* v, int vh, int vw, const uint8_t* is_fluid, double dt,
                      double cell_size, double scale, int dtype, void* stream) {
  SMK_REQUIRE(src && dst && u && v && is_fluid, "smk_advect_scalar: null pointer");
  SMK_REQUIRE(H > 0 && W > 0 && uh > 0 && uw > 0 && vh > 0 && vw > 0,
              "smk_advect_scalar: bad shape");
  SMK_REQUIRE(cell_size != 0.0, "smk_advect_scalar: cell_size must be non-zero");
  const double inv = 1.0 / cell_size;
  const int aps = scale != 1.0;
  auto g = grid2d(H, W);
  auto s = as_stream(stream);
  SMK_DISPATCH(dtype,
               hipLaunchKernelGGL(k_advect_scalar<double>, g, dim3(BX, BY), 0, s,
                                  (const double*)src, (double*)dst, H, W, (const double*)u, uh,
                                  uw, (const double*)v, vh, vw, is_fluid, dt, inv, scale, aps),
               hipLaunchKernelGGL(k_advect_scalar<float>, g, dim3(BX, BY), 0, s,
                                  (const float*)src, (float*)dst, H, W, (const float*)u, uh, uw,
                                  (const float*)v, vh, vw, is_fluid, dt, inv, scale, aps));
  return check_launch("k_advect_scalar");
}

int smk_advect_u(const void* us, void* ud, int uh, int uw, const void* v, int vh, int vw,
                 const uint8_t* fmu, double dt, double cell_size, int dtype, void* stream) {
  SMK_REQUIRE(us && ud && v && fmu, "smk_advect_u: null pointer");
  SMK_REQUIRE(uh > 0 && uw > 0 && vh > 0 && vw > 0, "smk_advect_u: bad shape");
  SMK_REQUIRE(cell_size != 0.0, "smk_advect_u: cell_size must be non-zero");
  const double inv = 1.0 / cell_size;
  auto g = grid2d(uh, uw);
  auto s = as_stream(stream);
  SMK_DISPATCH(dtype,
               hipLaunchKernelGGL(k_advect_u<double>, g, dim3(BX, BY), 0, s, (const double*)us,
                                  (double*)ud, uh, uw, (const double*)v, vh, vw, fmu, dt, inv),
               hipLaunchKernelGGL(k_advect_u<float>, g, dim3(BX, BY), 0, s, (const float*)us,
                                  (float*)ud, uh, uw, (const float*)v, vh, vw, fmu, dt, inv));
  return check_launch("k_advect_u");
}

int smk_advect_v(const void* vs, void* vd, int vh, int vw, const void* u, int uh, int uw,
                 const uint8_t* fmv, double dt, double cell_size, int dtype, void* stream) {
  SMK_REQUIRE(vs && vd && u && fmv, "smk_advect_v: null pointer");
  SMK_REQUIRE(uh > 0 && uw > 0 && vh > 0 && vw > 0, "smk_advect_v: bad shape");
  SMK_REQUIRE(cell_size != 0.0, "smk_advect_v: cell_size must be non-zero");
  const double inv = 1.0 / cell_size;
  auto g = grid2d(vh, vw);
  auto s = as_stream(stream);
  SMK_DISPATCH(dtype,
               hipLaunchKernelGGL(k_advect_v<double>, g, dim3(BX, BY), 0, s, (const double*)vs,
                                  (double*)vd, vh, vw, (const double*)u, uh, uw, fmv, dt, inv),
               hipLaunchKernelGGL(k_advect_v<float>, g, dim3(BX, BY), 0, s, (const float*)vs,
                                  (float*)vd, vh, vw, (const float*)u, uh, uw, fmv, dt, inv));
  return check_launch("k_advect_v");
}

int smk_advect_velocity(const void* us, void* ud, const void* vs, void* vd, const uint8_t* fmu,
                        const uint8_t* fmv, int H, int W, double dt, double cell_size, int dtype,
                        void* stream) {
  SMK_REQUIRE(us && ud && vs && vd && fmu && fmv, "smk_advect_velocity: null pointer");
  SMK_REQUIRE(H > 0 && W > 0, "smk_advect_velocity: bad shape");
  SMK_REQUIRE(cell_size != 0.0, "smk_advect_velocity: cell_size must be non-zero");
  const double inv = 1.0 / cell_size;
  auto g = grid2d(H + 1, W + 1);
  auto s = as_stream(stream);
  SMK_DISPATCH(dtype,
               hipLaunchKernelGGL(k_advect_velocity<double>, g, dim3(BX, BY), 0, s,
                                  (const double*)us, (double*)ud, (const double*)vs,
                                  (double*)vd, fmu, fmv, H, W, dt, inv),
               hipLaunchKernelGGL(k_advect_velocity<float>, g, dim3(BX, BY), 0, s,
                                  (const float*)us, (float*)ud, (const float*)vs, (float*)vd,
                                  fmu, fmv, H, W, dt, inv));
  return check_launch("k_advect_velocity");
}

int smk_apply_forces(void* v, int vh, int vw, const void* density, int H, int W,
                     const void* temperature, int Ht, int Wt, double dt, double gravity,
                     double buoyancy_coefficient, double density_coefficient,
                     double ambient_temperature, int dtype, void* stream) {
  SMK_REQUIRE(v && density, "smk_apply_forces: null pointer");
  SMK_REQUIRE(vh > 0 && vw > 0 && H > 0 && W > 0, "smk_apply_forces: bad shape");
  // forces.py:101-105 indexes both (H+1, W) force buffers with v's indices
  SMK_REQUIRE(vh <= H + 1 && vw <= W, "smk_apply_forces: v larger than the force buffer");
  if (temperature == nullptr) { Ht = H; Wt = W; }
  SMK_REQUIRE(vh <= Ht + 1 && vw <= Wt, "smk_apply_forces: v larger than the temperature force");
  auto g = grid2d(vh, vw);
  auto s = as_stream(stream);
  SMK_DISPATCH(dtype,
               hipLaunchKernelGGL(k_apply_forces<double>, g, dim3(BX, BY), 0, s, (double*)v, vh,
                                  vw, (const double*)density, H, W, (const double*)temperature,
                                  Ht, Wt, dt, gravity, buoyancy_coefficient, density_coefficient,
                                  ambient_temperature),
               hipLaunchKernelGGL(k_apply_forces<float>, g, dim3(BX, BY), 0, s, (float*)v, vh,
                                  vw, (const float*)density, H, W, (const float*)temperature, Ht,
                                  Wt, dt, gravity, buoyancy_coefficient, density_coefficient,
                                  ambient_temperature));
  return check_launch("k_apply_forces");
}

int smk_cool_temperature(void* T, int H, int W, double ambient, double decay, int dtype,
                         void* stream) {
  SMK_REQUIRE(T && H > 0 && W > 0, "smk_cool_temperature: bad arguments");
  const long n = (long)H * W;
  auto s = as_stream(stream);
  SMK_DISPATCH(dtype,
               hipLaunchKernelGGL(k_cool<double>, dim3(cdiv(n, 256)), dim3(256), 0, s,
                                  (double*)T, n, ambient, decay),
               hipLaunchKernelGGL(k_cool<float>, dim3(cdiv(n, 256)), dim3(256), 0, s, (float*)T,
                                  n, ambient, decay));
  return check_launch("k_cool");
}

int smk_scale(void* x, long n, double factor, int dtype, void* stream) {
  SMK_REQUIRE(x && n >= 0, "smk_scale: bad arguments");
  if (n == 0) return SMK_OK;
  auto s = as_stream(stream);
  SMK_DISPATCH(dtype,
               hipLaunchKernelGGL(k_scale<double>, dim3(cdiv(n, 256)), dim3(256), 0, s,
                                  (double*)x, n, factor),
               hipLaunchKernelGGL(k_scale<float>, dim3(cdiv(n, 256)), dim3(256), 0, s,
                                  (float*)x, n, (float)factor));
  return check_launch("k_scale");
}

int smk_divergence(const void* u, const void* v, void* div, const uint8_t* fmu,
                   const uint8_t* fmv, int H, int W, double cell_size, int dtype, void* stream) {
  SMK_REQUIRE(u && v && div && fmu && fmv, "smk_divergence: null pointer");
  SMK_REQUIRE(H > 0 && W > 0 && cell_size != 0.0, "smk_divergence: bad arguments");
  auto g = grid2d(H, W);
  auto s = as_stream(stream);
  SMK_DISPATCH(dtype,
               hipLaunchKernelGGL(k_divergence<double>, g, dim3(BX, BY), 0, s, (const double*)u,
                                  (const double*)v, (double*)div, fmu, fmv, H, W,
                                  1.0 / cell_size),
               hipLaunchKernelGGL(k_divergence<float>, g, dim3(BX, BY), 0, s, (const float*)u,
                                  (const float*)v, (float*)div, fmu, fmv, H, W,
                                  1.0 / cell_size));
  return check_launch("k_divergence");
}

int smk_divergence_residual(const void* u, const void* v, const void* p,
                            const uint8_t* is_fluid, const uint8_t* fmu, const uint8_t* fmv,
                            int H, int W, double cell_size, double alpha,
                            double reciprocal_beta, void* div, void* r, int dtype,
                            void* stream) {
  SMK_REQUIRE(u && v && p && is_fluid && fmu && fmv && div && r,
              "smk_divergence_residual: null pointer");
  SMK_REQUIRE(H > 0 && W > 0 && cell_size != 0.0 && reciprocal_beta != 0.0,
              "smk_divergence_residual: bad arguments");
  const double inv = 1.0 / cell_size, beta = 1.0 / reciprocal_beta;
  auto g = grid2d(H, W);
  auto s = as_stream(stream);
  SMK_DISPATCH(dtype,
               hipLaunchKernelGGL(k_div_residual<double>, g, dim3(BX, BY), 0, s,
                                  (const double*)u, (const double*)v, (const double*)p, is_fluid,
                                  fmu, fmv, H, W, inv, alpha, beta, (double*)div, (double*)r),
               hipLaunchKernelGGL(k_div_residual<float>, g, dim3(BX, BY), 0, s,
                                  (const float*)u, (const float*)v, (const float*)p, is_fluid,
                                  fmu, fmv, H, W, inv, alpha, beta, (float*)div, (float*)r));
  return check_launch("k_div_residual");
}

int smk_pressure_gradient(void* u, void* v, const void* p, const uint8_t* fmu,
                          const uint8_t* fmv, const uint8_t* solid, int H, int W,
                          double fluid_density, double cell_size, double dt, int dtype,
                          void* stream) {
  SMK_REQUIRE(u && v && p && fmu && fmv, "smk_pressure_gradient: null pointer");
  SMK_REQUIRE(H > 0 && W > 0, "smk_pressure_gradient: bad shape");
  const double scale = dt / (fluid_density * cell_size);
  auto g = grid2d(H + 1, W + 1);
  auto s = as_stream(stream);
  SMK_DISPATCH(dtype,
               hipLaunchKernelGGL(k_pressure_gradient<double>, g, dim3(BX, BY), 0, s, (double*)u,
                                  (double*)v, (const double*)p, fmu, fmv, solid, H, W, scale),
               hipLaunchKernelGGL(k_pressure_gradient<float>, g, dim3(BX, BY), 0, s, (float*)u,
                                  (float*)v, (const float*)p, fmu, fmv, solid, H, W,
                                  scale));
  return check_launch("k_pressure_gradient");
}

int smk_enforce_boundaries(void* u, void* v, const uint8_t* solid, int H, int W, int dtype,
                           void* stream) {
  SMK_REQUIRE(u && v && solid && H > 0 && W > 0, "smk_enforce_boundaries: bad arguments");
  auto g = grid2d(H + 1, W + 1);
  auto s = as_stream(stream);
  SMK_DISPATCH(dtype,
               hipLaunchKernelGGL(k_enforce_boundaries<double>, g, dim3(BX, BY), 0, s,
                                  (double*)u, (double*)v, solid, H, W),
               hipLaunchKernelGGL(k_enforce_boundaries<float>, g, dim3(BX, BY), 0, s,
                                  (float*)u, (float*)v, solid, H, W));
  return check_launch("k_enforce_boundaries");
}

int smk_residual(const void* p, const void* div, const uint8_t* is_fluid, const uint8_t* fmu,
                 const uint8_t* fmv, int H, int W, double alpha, double reciprocal_beta, void* r,
                 int dtype, void* stream) {
  SMK_REQUIRE(p && div && is_fluid && fmu && fmv && r, "smk_residual: null pointer");
  SMK_REQUIRE(H > 0 && W > 0 && reciprocal_beta != 0.0, "smk_residual: bad arguments");
  const double beta = 1.0 / reciprocal_beta;
  auto g = grid2d(H, W);
  auto s = as_stream(stream);
  SMK_DISPATCH(dtype,
               hipLaunchKernelGGL(k_residual_masks<double>, g, dim3(BX, BY), 0, s,
                                  (const double*)p, (const double*)div, is_fluid, fmu, fmv, H, W,
                                  alpha, beta, (double*)r),
               hipLaunchKernelGGL(k_residual_masks<float>, g, dim3(BX, BY), 0, s,
                                  (const float*)p, (const float*)div, is_fluid, fmu, fmv, H, W,
                                  alpha, beta, (float*)r));
  return check_launch("k_residual_masks");
}

int smk_emit_scalar(void* field, int H, int W, double value, double dt, double pos_x,
                    double pos_y, double radius, int approach, int dtype, void* stream) {
  SMK_REQUIRE(field && H > 0 && W > 0, "smk_emit_scalar: bad arguments");
  SMK_REQUIRE(approach >= 0 && approach <= 3, "smk_emit_scalar: unknown approach %d", approach);
  // Cells with distance >= radius get falloff 0 and are left unchanged, so
  // only the disc's bounding box is visited (emitters.py:52-56 does the same).
  double fx0 = floor(pos_x - radius) - 1.0, fx1 = ceil(pos_x + radius) + 2.0;
  double fy0 = floor(pos_y - radius) - 1.0, fy1 = ceil(pos_y + radius) + 2.0;
  if (!(fx0 == fx0) || !(fy0 == fy0) || !(fx1 == fx1) || !(fy1 == fy1)) return SMK_OK;
  const int x0 = (int)fmax(0.0, fmin((double)W, fx0)), x1 = (int)fmax(0.0, fmin((double)W, fx1));
  const int y0 = (int)fmax(0.0, fmin((double)H, fy0)), y1 = (int)fmax(0.0, fmin((double)H, fy1));
  if (x1 <= x0 || y1 <= y0) return SMK_OK;
  dim3 g(cdiv(x1 - x0, BX), cdiv(y1 - y0, BY));
  auto s = as_stream(stream);
  SMK_DISPATCH(dtype,
               hipLaunchKernelGGL(k_emit<double>, g, dim3(BX, BY), 0, s, (double*)field, H, W, x0,
                                  y0, x1, y1, value, dt, pos_x, pos_y, radius, approach),
               hipLaunchKernelGGL(k_emit<float>, g, dim3(BX, BY), 0, s, (float*)field, H, W, x0,
                                  y0, x1, y1, value, dt, pos_x, pos_y, radius, approach));
  return check_launch("k_emit");
}

int smk_box_source(void* field, int H, int W, int x0, int y0, int x1, int y1, double value,
                   int mode, int dtype, void* stream) {
  SMK_REQUIRE(field && H > 0 && W > 0, "smk_box_source: bad arguments");
  SMK_REQUIRE(mode == 0 || mode == 1, "smk_box_source: mode must be 0 (set) or 1 (add)");
  x0 = x0 < 0 ? 0 : x0; y0 = y0 < 0 ? 0 : y0;
  x1 = x1 > W ? W : x1; y1 = y1 > H ? H : y1;
  if (x1 <= x0 || y1 <= y0) return SMK_OK;
  dim3 g(cdiv(x1 - x0, BX), cdiv(y1 - y0, BY));
  auto s = as_stream(stream);
  SMK_DISPATCH(dtype,
               hipLaunchKernelGGL(k_box<double>, g, dim3(BX, BY), 0, s, (double*)field, W, x0, y0,
                                  x1, y1, value, mode),
               hipLaunchKernelGGL(k_box<float>, g, dim3(BX, BY), 0, s, (float*)field, W, x0, y0,
                                  x1, y1, (float)value, mode));
  return check_launch("k_box");
}

int smk_box_sources(void* field, int H, int W, int n, const int* rects, const double* values,
                    const int* modes, int dtype, void* stream) {
  SMK_REQUIRE(field && H > 0 && W > 0 && n >= 0, "smk_box_sources: bad arguments");
  SMK_REQUIRE(n == 0 || (rects && values && modes), "smk_box_sources: null box list");
  BoxList b{};
  int m = 0, maxw = 0, maxh = 0;
  bool fused = n <= BOX_MAX;
  for (int k = 0; k < n; ++k) {
    SMK_REQUIRE(modes[k] == 0 || modes[k] == 1,
                "smk_box_sources: mode must be 0 (set) or 1 (add)");
    const int x0 = rects[4 * k] < 0 ? 0 : rects[4 * k];
    const int y0 = rects[4 * k + 1] < 0 ? 0 : rects[4 * k + 1];
    const int x1 = rects[4 * k + 2] > W ? W : rects[4 * k + 2];
    const int y1 = rects[4 * k + 3] > H ? H : rects[4 * k + 3];
    if (x1 <= x0 || y1 <= y0 || !fused) continue;
    for (int q = 0; q < m && fused; ++q)  // overlap with a box already listed
      if (x0 < b.x1[q] && b.x0[q] < x1 && y0 < b.y1[q] && b.y0[q] < y1) fused = false;
    if (!fused) continue;
    b.x0[m] = x0;
    b.y0[m] = y0;
    b.x1[m] = x1;
    b.y1[m] = y1;
    b.mode[m] = modes[k];
    b.value[m] = values[k];
    maxw = x1 - x0 > maxw ? x1 - x0 : maxw;
    maxh = y1 - y0 > maxh ? y1 - y0 : maxh;
    ++m;
  }
  if (!fused) {  // overlapping boxes: one after another, in list order
    for (int k = 0; k < n; ++k) {
      int rc = smk_box_source(field, H, W, rects[4 * k], rects[4 * k + 1], rects[4 * k + 2],
                              rects[4 * k + 3], values[k], modes[k], dtype, stream);
      if (rc) return rc;
    }
    return SMK_OK;
  }
  if (m == 0) return SMK_OK;
  dim3 g(cdiv(maxw, BX), cdiv(maxh, BY), m);
  auto s = as_stream(stream);
  SMK_DISPATCH(dtype,
               hipLaunchKernelGGL(k_boxes<double>, g, dim3(BX, BY), 0, s, (double*)field, W, b),
               hipLaunchKernelGGL(k_boxes<float>, g, dim3(BX, BY), 0, s, (float*)field, W, b));
  return check_launch("k_boxes");
}

}  // extern "C"
