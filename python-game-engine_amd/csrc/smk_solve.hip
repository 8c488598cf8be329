// smk_solve.hip -- pressure solvers: Jacobi, red-black SOR, geometric
// multigrid (reference scaling or rediscretised), FP32-in-FP64 defect
// correction.
//
// Core kernel: k_smooth_tiled.  A workgroup owns a TX x TY tile of the
// output, stages the tile plus an R-cell halo of p, rhs and the packed cell
// flags in LDS once, then runs `npass` colour passes (RB) or iterations
// (Jacobi) entirely in LDS.  After pass k only cells at distance >= k+1 from
// the staged region's edge are exact, so R >= npass keeps the tile interior
// exact: npass sweeps cost one HBM read and one write of p instead of npass.
// The kernel reads p_in and writes p_out (ping-pong), so neighbouring tiles
// always stage the same old iterate and results do not depend on tile order
// or size.  Optional fused stages, all computed from the LDS tile:
//   prolong-in   p += bilinear(coarse e) before the passes
//                (transfer.py:38-72 + solver.py:134)
//   restrict-out residual of the interior, 2x2 masked mean / coarse_div into
//                the coarse rhs (residual.py:6-42, transfer.py:5-36,
//                solver.py:115-125); needs R >= npass + 1
//   norm-out     per-workgroup sum of r^2 over fluid interior cells
//                (deterministic: fixed per-thread order, fixed tree)
//
// All arithmetic follows the reference expression order with
// -ffp-contract=off: neighbour sum ((l + r) + d) + u in the storage type,
// then (sum + alpha*div) * 1/4 and the SOR blend in float64 (Numba types the
// Python-float coefficients as float64).
#include <cmath>
#include <cstring>
#include <vector>

#include "smk_common.hpp"

namespace smk {

enum : int { MODE_RB = 0, MODE_JACOBI = 1 };

template <typename T>
struct SmoothArgs {
  const T* p_in;
  T* p_out;
  const T* rhs;
  const uint8_t* flags;
  int H, W;
  double alpha;    // -h^2 rho / dt of this level
  double omega;    // SOR blend (RB mode)
  int npass;       // colour passes (RB) or iterations (Jacobi)
  int first_colour;
  int R;           // halo width staged
  int zero_init;   // treat p_in as all zero (coarse levels, solver.py:127)
  // prolong-in
  const T* coarse_e;
  int Hc, Wc;
  // restrict-out
  T* coarse_rhs;
  double coarse_div;
  // norm-out
  double* norm_partial;
};

template <typename T>
__device__ __forceinline__ double poisson_new(T l, T r, T d, T u, T rhs, double alpha) {
  const T s = l + r + d + u;
  return ((double)s + alpha * (double)rhs) * 0.25;
}

template <typename T>
__device__ __forceinline__ double poisson_residual(T l, T r, T d, T u, T c, T rhs,
                                                   double alpha) {
  const T s = l + r + d + u;
  return ((double)s + alpha * (double)rhs) - 4.0 * (double)c;
}

// Region = RH rows x 64 columns staged in LDS; lane x of every wavefront
// owns region column x, the NW waves split the rows.  R is the halo, so the
// output tile is (64 - 2R) x (RH - 2R).  Red-black passes walk row pairs:
// in rows (2m, 2m+1) lane x updates the one cell of the pass's colour, so
// every lane is busy and no index needs a division.
template <typename T, int MODE, int RH, int NT>
__global__ __launch_bounds__(NT) void k_smooth_tile64(const SmoothArgs<T> a) {
  constexpr int RW = 64;
  constexpr int NW = NT / 64;
  __shared__ T sp[RH * RW];
  __shared__ T sr[RH * RW];
  __shared__ T sq[MODE == MODE_JACOBI ? RH * RW : 1];
  __shared__ uint8_t sf[RH * RW];
  __shared__ double red[NW];

  const int R = a.R;
  const int TX = RW - 2 * R, TY = RH - 2 * R;
  const int H = a.H, W = a.W;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int gx0 = blockIdx.x * TX - R, gy0 = blockIdx.y * TY - R;
  const int gx = gx0 + lane;
  const bool colin = gx >= 0 && gx < W;

  // ---- stage p (+ prolonged correction), rhs, flags
  for (int ly = wave; ly < RH; ly += NW) {
    const int gy = gy0 + ly;
    const int o = ly * RW + lane;
    T pv = (T)0, rv = (T)0;
    uint8_t fv = 0;
    if (colin && gy >= 0 && gy < H) {
      const size_t g = (size_t)gy * W + gx;
      fv = a.flags[g];
      rv = a.rhs[g];
      pv = a.zero_init ? (T)0 : a.p_in[g];
      if (a.coarse_e != nullptr) {
        T corr = (T)0;
        if (fv & F_FLUID) {
          const double xc = (gx + 0.5) * 0.5, yc = (gy + 0.5) * 0.5;
          corr = (T)sample_field(a.coarse_e, a.Hc, a.Wc, xc - 0.5, yc - 0.5);
        }
        pv = pv + corr;
      }
    }
    sp[o] = pv;
    sr[o] = rv;
    sf[o] = fv;
  }
  __syncthreads();

  // ---- passes
  T* cur = sp;
  for (int pass = 0; pass < a.npass; ++pass) {
    const int lo = pass + 1;
    const bool xin = lane >= lo && lane < RW - lo;
    if (MODE == MODE_RB) {
      const int colour = (a.first_colour + pass) & 1;
      const int par = (lane + colour) & 1;  // region origin has even parity
      for (int pr = wave; pr < RH / 2; pr += NW) {
        const int ly = 2 * pr + par;
        if (!xin || ly < lo || ly >= RH - lo) continue;
        const int o = ly * RW + lane;
        const uint8_t f = sf[o];
        if (!(f & F_FLUID)) continue;
        const T ce = cur[o];
        const T l = (f & F_LEFT) ? cur[o - 1] : ce;
        const T rr = (f & F_RIGHT) ? cur[o + 1] : ce;
        const T d = (f & F_DOWN) ? cur[o - RW] : ce;
        const T u = (f & F_UP) ? cur[o + RW] : ce;
        const double nw = poisson_new(l, rr, d, u, sr[o], a.alpha);
        cur[o] = (T)((1.0 - a.omega) * (double)ce + a.omega * nw);
      }
    } else {
      T* nxt = (cur == sp) ? sq : sp;
      for (int ly = wave; ly < RH; ly += NW) {
        if (!xin || ly < lo || ly >= RH - lo) continue;
        const int o = ly * RW + lane;
        const uint8_t f = sf[o];
        const T ce = cur[o];
        if (!(f & F_FLUID)) { nxt[o] = ce; continue; }
        const T l = (f & F_LEFT) ? cur[o - 1] : ce;
        const T rr = (f & F_RIGHT) ? cur[o + 1] : ce;
        const T d = (f & F_DOWN) ? cur[o - RW] : ce;
        const T u = (f & F_UP) ? cur[o + RW] : ce;
        nxt[o] = (T)poisson_new(l, rr, d, u, sr[o], a.alpha);
      }
      cur = nxt;
    }
    __syncthreads();
  }

  // ---- write the tile interior
  const bool xout = lane >= R && lane < R + TX && colin;
  for (int ly = R + wave; ly < R + TY; ly += NW) {
    const int gy = gy0 + ly;
    if (xout && gy < H) a.p_out[(size_t)gy * W + gx] = cur[ly * RW + lane];
  }

  // ---- fused residual -> restriction into the coarse rhs
  if (a.coarse_rhs != nullptr) {
    const int CX = TX / 2, CY = TY / 2;
    const int Wc = W / 2, Hc = H / 2;
    const int cx = lane & 31;
    for (int cyb = 2 * wave; cyb < CY; cyb += 2 * NW) {
      const int cy = cyb + (lane >> 5);
      if (cy >= CY || cx >= CX) continue;
      const int Jc = blockIdx.y * CY + cy, Ic = blockIdx.x * CX + cx;
      if (Jc >= Hc || Ic >= Wc) continue;
      double total = 0.0;
      int count = 0;
#pragma unroll
      for (int dj = 0; dj < 2; ++dj)
#pragma unroll
        for (int di = 0; di < 2; ++di) {
          const int o = (R + 2 * cy + dj) * RW + R + 2 * cx + di;
          const uint8_t f = sf[o];
          if (!(f & F_FLUID)) continue;
          const T ce = cur[o];
          const T l = (f & F_LEFT) ? cur[o - 1] : ce;
          const T rr = (f & F_RIGHT) ? cur[o + 1] : ce;
          const T d = (f & F_DOWN) ? cur[o - RW] : ce;
          const T u = (f & F_UP) ? cur[o + RW] : ce;
          // the residual is stored in the level's dtype (solver.py:116-117),
          // then summed in float64 (transfer.py:26-33)
          const T rv = (T)poisson_residual(l, rr, d, u, ce, sr[o], a.alpha);
          total += (double)rv;
          count += 1;
        }
      const T mean = count > 0 ? (T)(total / (double)count) : (T)0;
      a.coarse_rhs[(size_t)Jc * Wc + Ic] = (T)((double)mean / a.coarse_div);
    }
  }

  // ---- fused residual norm over the fluid interior cells
  if (a.norm_partial != nullptr) {
    double acc = 0.0;
    for (int ly = R + wave; ly < R + TY; ly += NW) {
      const int gy = gy0 + ly;
      if (!xout || gy >= H) continue;
      const int o = ly * RW + lane;
      const uint8_t f = sf[o];
      if (!(f & F_FLUID)) continue;
      const T ce = cur[o];
      const T l = (f & F_LEFT) ? cur[o - 1] : ce;
      const T rr = (f & F_RIGHT) ? cur[o + 1] : ce;
      const T d = (f & F_DOWN) ? cur[o - RW] : ce;
      const T u = (f & F_UP) ? cur[o + RW] : ce;
      const double rv = poisson_residual(l, rr, d, u, ce, sr[o], a.alpha);
      acc += rv * rv;
    }
    const double tot = block_sum<NT>(acc, red);
    if (threadIdx.x == 0) a.norm_partial[blockIdx.y * gridDim.x + blockIdx.x] = tot;
  }
}

// The small levels of a V-cycle (top .. coarsest) in one workgroup with every
// level's p, rhs and flags resident in LDS: pre-smooth and residual-restrict
// going down, the coarsest smoothing, prolong-correct and post-smooth coming
// up (solver.py:94-139), one launch instead of two per level.
constexpr int SMALL_MAX = 8;

template <typename T>
struct SmallArgs {
  int nl;
  int H[SMALL_MAX], W[SMALL_MAX];
  const uint8_t* flags[SMALL_MAX];
  double alpha[SMALL_MAX], omega[SMALL_MAX], coarse_div[SMALL_MAX];
  const T* rhs_top;
  const T* p_top_in;
  T* p_top_out;
  int zero_top;
  int npass_smooth;    // per pre/post smoothing (RB: 2*sweeps)
  int npass_coarsest;  // coarsest level
};

template <typename T, int MODE, int NT>
__device__ void small_smooth(T* p, T* tmp, const T* rhs, const uint8_t* fl, int H, int W,
                             double alpha, double omega, int npass) {
  const int N = H * W;
  T* cur = p;
  for (int pass = 0; pass < npass; ++pass) {
    if (MODE == MODE_RB) {
      const int colour = pass & 1;
      const int nh = (W + 1) >> 1;
      for (int k = threadIdx.x; k < H * nh; k += NT) {
        const int j = k / nh, c = k - j * nh;
        const int i = ((colour + j) & 1) + 2 * c;
        if (i >= W) continue;
        const int o = j * W + i;
        const uint8_t f = fl[o];
        if (!(f & F_FLUID)) continue;
        const T ce = cur[o];
        const T l = (f & F_LEFT) ? cur[o - 1] : ce;
        const T rr = (f & F_RIGHT) ? cur[o + 1] : ce;
        const T d = (f & F_DOWN) ? cur[o - W] : ce;
        const T u = (f & F_UP) ? cur[o + W] : ce;
        const double nw = poisson_new(l, rr, d, u, rhs[o], alpha);
        cur[o] = (T)((1.0 - omega) * (double)ce + omega * nw);
      }
    } else {
      T* nxt = (cur == p) ? tmp : p;
      for (int o = threadIdx.x; o < N; o += NT) {
        const uint8_t f = fl[o];
        const T ce = cur[o];
        if (!(f & F_FLUID)) { nxt[o] = ce; continue; }
        const T l = (f & F_LEFT) ? cur[o - 1] : ce;
        const T rr = (f & F_RIGHT) ? cur[o + 1] : ce;
        const T d = (f & F_DOWN) ? cur[o - W] : ce;
        const T u = (f & F_UP) ? cur[o + W] : ce;
        nxt[o] = (T)poisson_new(l, rr, d, u, rhs[o], alpha);
      }
      cur = nxt;
    }
    __syncthreads();
  }
  if (cur != p) {  // Jacobi with an odd count: copy back (jacobi.py:51-52)
    for (int o = threadIdx.x; o < N; o += NT) p[o] = cur[o];
    __syncthreads();
  }
}

template <typename T, int MODE, int NT>
__global__ __launch_bounds__(NT) void k_vcycle_small(const SmallArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* P[SMALL_MAX];
  T* B[SMALL_MAX];
  uint8_t* F[SMALL_MAX];
  unsigned char* cursor = smem;
  int nmax = 0;
  for (int l = 0; l < a.nl; ++l) {
    const int n = a.H[l] * a.W[l];
    nmax = n > nmax ? n : nmax;
    P[l] = reinterpret_cast<T*>(cursor);
    cursor += ((size_t)n * sizeof(T) + 15) & ~(size_t)15;
    B[l] = reinterpret_cast<T*>(cursor);
    cursor += ((size_t)n * sizeof(T) + 15) & ~(size_t)15;
    F[l] = cursor;
    cursor += ((size_t)n + 15) & ~(size_t)15;
  }
  T* tmp = reinterpret_cast<T*>(cursor);  // Jacobi ping-pong partner
  (void)nmax;

  // stage the top level
  {
    const int n = a.H[0] * a.W[0];
    for (int o = threadIdx.x; o < n; o += NT) {
      P[0][o] = a.zero_top ? (T)0 : a.p_top_in[o];
      B[0][o] = a.rhs_top[o];
      F[0][o] = a.flags[0][o];
    }
    for (int l = 1; l < a.nl; ++l) {
      const int nl_ = a.H[l] * a.W[l];
      for (int o = threadIdx.x; o < nl_; o += NT) {
        F[l][o] = a.flags[l][o];
        P[l][o] = (T)0;  // solver.py:127
      }
    }
  }
  __syncthreads();

  const int last = a.nl - 1;
  // down
  for (int l = 0; l < last; ++l) {
    const int H = a.H[l], W = a.W[l];
    small_smooth<T, MODE, NT>(P[l], tmp, B[l], F[l], H, W, a.alpha[l], a.omega[l],
                              a.npass_smooth);
    const int Hc = a.H[l + 1], Wc = a.W[l + 1];
    for (int k = threadIdx.x; k < Hc * Wc; k += NT) {
      const int J = k / Wc, I = k - J * Wc;
      double total = 0.0;
      int count = 0;
      for (int dj = 0; dj < 2; ++dj)
        for (int di = 0; di < 2; ++di) {
          const int o = (2 * J + dj) * W + 2 * I + di;
          const uint8_t f = F[l][o];
          if (!(f & F_FLUID)) continue;
          const T* p = P[l];
          const T ce = p[o];
          const T lf = (f & F_LEFT) ? p[o - 1] : ce;
          const T rt = (f & F_RIGHT) ? p[o + 1] : ce;
          const T dn = (f & F_DOWN) ? p[o - W] : ce;
          const T up = (f & F_UP) ? p[o + W] : ce;
          const T rv = (T)poisson_residual(lf, rt, dn, up, ce, B[l][o], a.alpha[l]);
          total += (double)rv;
          count += 1;
        }
      const T mean = count > 0 ? (T)(total / (double)count) : (T)0;
      B[l + 1][k] = (T)((double)mean / a.coarse_div[l]);
    }
    __syncthreads();
  }
  // coarsest
  small_smooth<T, MODE, NT>(P[last], tmp, B[last], F[last], a.H[last], a.W[last],
                            a.alpha[last], a.omega[last], a.npass_coarsest);
  // up
  for (int l = last - 1; l >= 0; --l) {
    const int H = a.H[l], W = a.W[l];
    const int Hc = a.H[l + 1], Wc = a.W[l + 1];
    for (int o = threadIdx.x; o < H * W; o += NT) {
      const int j = o / W, i = o - j * W;
      T corr = (T)0;
      if (F[l][o] & F_FLUID)
        corr = (T)sample_field(P[l + 1], Hc, Wc, (i + 0.5) * 0.5 - 0.5, (j + 0.5) * 0.5 - 0.5);
      P[l][o] = P[l][o] + corr;
    }
    __syncthreads();
    small_smooth<T, MODE, NT>(P[l], tmp, B[l], F[l], H, W, a.alpha[l], a.omega[l],
                              a.npass_smooth);
  }
  const int n0 = a.H[0] * a.W[0];
  for (int o = threadIdx.x; o < n0; o += NT) a.p_top_out[o] = P[0][o];
}

constexpr int NB_X = 64, NB_Y = 4, NB_T = NB_X * NB_Y;

// Residual norm^2 partials per 64x4 block (fixed order).  If `defect` is
// non-null the residual r is also written as defect[k] = r / alpha in float32:
// the right-hand side of the FP32 correction solve (4 - N) e = alpha * defect.
template <typename T>
__global__ __launch_bounds__(NB_T) void k_residual_norm(const T* __restrict__ p,
                                                        const T* __restrict__ rhs,
                                                        const uint8_t* __restrict__ flags,
                                                        int H, int W, double alpha,
                                                        double* __restrict__ partial,
                                                        float* __restrict__ defect) {
  __shared__ double red[NB_T / 64];
  const int i = blockIdx.x * NB_X + (threadIdx.x & (NB_X - 1));
  const int j = blockIdx.y * NB_Y + (threadIdx.x / NB_X);
  double acc = 0.0;
  if (i < W && j < H) {
    const size_t o = (size_t)j * W + i;
    const uint8_t f = flags[o];
    double rv = 0.0;
    if (f & F_FLUID) {
      const T ce = p[o];
      const T l = (f & F_LEFT) ? p[o - 1] : ce;
      const T rr = (f & F_RIGHT) ? p[o + 1] : ce;
      const T d = (f & F_DOWN) ? p[o - W] : ce;
      const T u = (f & F_UP) ? p[o + W] : ce;
      rv = poisson_residual(l, rr, d, u, ce, rhs[o], alpha);
      acc = rv * rv;
    }
    if (defect != nullptr) defect[o] = (f & F_FLUID) ? (float)(rv / alpha) : 0.0f;
  }
  const double tot = block_sum<NB_T>(acc, red);
  if (threadIdx.x == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = tot;
}

// ||alpha*rhs||^2 partials over fluid cells (the p = 0 residual).
template <typename T>
__global__ __launch_bounds__(NB_T) void k_rhs_norm(const T* __restrict__ rhs,
                                                   const uint8_t* __restrict__ flags, int H,
                                                   int W, double alpha,
                                                   double* __restrict__ partial) {
  __shared__ double red[NB_T / 64];
  const int i = blockIdx.x * NB_X + (threadIdx.x & (NB_X - 1));
  const int j = blockIdx.y * NB_Y + (threadIdx.x / NB_X);
  double acc = 0.0;
  if (i < W && j < H) {
    const size_t o = (size_t)j * W + i;
    if (flags[o] & F_FLUID) {
      const double b = alpha * (double)rhs[o];
      acc = b * b;
    }
  }
  const double tot = block_sum<NB_T>(acc, red);
  if (threadIdx.x == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = tot;
}

// Sum n partials in a fixed order into out[slot].
__global__ __launch_bounds__(1024) void k_sum_partials(const double* __restrict__ partial, int n,
                                                       double* __restrict__ out, int slot) {
  __shared__ double red[1024 / 64];
  double acc = 0.0;
  for (int k = threadIdx.x; k < n; k += 1024) acc += partial[k];
  const double tot = block_sum<1024>(acc, red);
  if (threadIdx.x == 0) out[slot] = tot;
}

template <typename T>
__global__ void k_add_correction(T* __restrict__ p, const float* __restrict__ e, long n) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = p[k] + (T)e[k];
}

__global__ void k_solid_from_flags(const uint8_t* __restrict__ flags, long n,
                                   uint8_t* __restrict__ solid) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) solid[k] = (flags[k] & F_FLUID) ? 0 : 1;
}

// ------------------------------------------------------------------ host side
constexpr int TILE_NT = 256;
constexpr int TILE_RH = 48;
constexpr int MAX_R = 8;
constexpr int SMALL_NT = 256;

inline int tile_tx(int R) { return 64 - 2 * R; }
inline int tile_ty(int R) { return TILE_RH - 2 * R; }

template <typename T>
int launch_tiled(int mode, const SmoothArgs<T>& a, hipStream_t s) {
  if (a.R < 1 || a.R > MAX_R) return fail(SMK_EINVAL, "halo %d outside [1, %d]", a.R, MAX_R);
  dim3 grid(cdiv(a.W, tile_tx(a.R)), cdiv(a.H, tile_ty(a.R)));
  if (mode == MODE_RB)
    hipLaunchKernelGGL((k_smooth_tile64<T, MODE_RB, TILE_RH, TILE_NT>), grid, dim3(TILE_NT), 0,
                       s, a);
  else
    hipLaunchKernelGGL((k_smooth_tile64<T, MODE_JACOBI, 32, TILE_NT>),
                       dim3(cdiv(a.W, tile_tx(a.R)), cdiv(a.H, 32 - 2 * a.R)), dim3(TILE_NT), 0,
                       s, a);
  return check_launch("k_smooth_tile64");
}

// partial-sum slots one tiled launch can produce
inline int tiled_blocks(int mode, int H, int W, int R) {
  const int ty = mode == MODE_RB ? tile_ty(R) : 32 - 2 * R;
  return (int)(cdiv(W, tile_tx(R)) * cdiv(H, ty));
}

inline size_t small_lds_bytes(const int* Hs, const int* Ws, int nl, size_t es, int mode) {
  size_t b = 0, nmax = 0;
  for (int l = 0; l < nl; ++l) {
    const size_t n = (size_t)Hs[l] * Ws[l];
    nmax = n > nmax ? n : nmax;
    b += 2 * ((n * es + 15) & ~(size_t)15) + ((n + 15) & ~(size_t)15);
  }
  if (mode == MODE_JACOBI) b += nmax * es;
  return b;
}

static size_t g_lds_limit = 0;

// Largest dynamic LDS a workgroup may use on this device (opt-in limit when
// the runtime reports one), and raise the kernels' attribute to it.
template <typename T>
static size_t small_lds_limit() {
  static size_t lim = 0;
  if (lim) return lim;
  int dev = 0, a = 0, b = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&a, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  (void)hipDeviceGetAttribute(&b, hipDeviceAttributeSharedMemPerBlockOptin, dev);
  size_t want = (size_t)(a > b ? a : b);
  if (want > 160 * 1024) want = 160 * 1024;
  if (want < 64 * 1024) want = 64 * 1024;
  if (want > 64 * 1024) {
    hipError_t e1 = hipFuncSetAttribute(
        reinterpret_cast<const void*>(&k_vcycle_small<T, MODE_RB, SMALL_NT>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)want);
    hipError_t e2 = hipFuncSetAttribute(
        reinterpret_cast<const void*>(&k_vcycle_small<T, MODE_JACOBI, SMALL_NT>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)want);
    if (e1 != hipSuccess || e2 != hipSuccess) {
      (void)hipGetLastError();
      want = 64 * 1024;
    }
  }
  lim = want;
  g_lds_limit = want;
  return lim;
}

template <typename T>
int launch_small(int mode, const SmallArgs<T>& a, size_t lds, hipStream_t s) {
  if (mode == MODE_RB)
    hipLaunchKernelGGL((k_vcycle_small<T, MODE_RB, SMALL_NT>), dim3(1), dim3(SMALL_NT), lds, s,
                       a);
  else
    hipLaunchKernelGGL((k_vcycle_small<T, MODE_JACOBI, SMALL_NT>), dim3(1), dim3(SMALL_NT), lds,
                       s, a);
  return check_launch("k_vcycle_small");
}

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

inline int norm_blocks(int H, int W) { return (int)(cdiv(W, NB_X) * cdiv(H, NB_Y)); }

// Workspace of the single-grid solvers: flags | p scratch | partials | results
struct SolverWs {
  uint8_t* flags;
  void* ptmp;
  double* partial;
  double* result;  // [0] = ||r||^2, [1] = ||alpha b||^2
  size_t bytes;
};

inline size_t solver_ws_bytes(int H, int W, int dtype) {
  const size_t n = (size_t)H * W;
  const size_t es = dtype == SMK_F64 ? 8 : 4;
  return align256(n) + align256(n * es) + align256((size_t)norm_blocks(H, W) * 8) + 256;
}

inline SolverWs carve_ws(void* ws, int H, int W, int dtype) {
  SolverWs w;
  const size_t n = (size_t)H * W;
  const size_t es = dtype == SMK_F64 ? 8 : 4;
  unsigned char* b = reinterpret_cast<unsigned char*>(ws);
  w.flags = b;
  b += align256(n);
  w.ptmp = b;
  b += align256(n * es);
  w.partial = reinterpret_cast<double*>(b);
  b += align256((size_t)norm_blocks(H, W) * 8);
  w.result = reinterpret_cast<double*>(b);
  w.bytes = solver_ws_bytes(H, W, dtype);
  return w;
}

// Runs `npass_total` passes from src into the buffer `dst_final`, using
// `other` as the ping-pong partner; chunked into launches of at most
// MAX_R passes.  The chunking keeps the launch count's parity such that the
// last launch writes dst_final (src may equal dst_final).
template <typename T>
int run_passes(int mode, T* dst_final, T* other, const T* src, const T* rhs,
               const uint8_t* flags, int H, int W, double alpha, double omega,
               int npass_total, int first_colour, hipStream_t s) {
  if (npass_total <= 0) {
    if (src != dst_final)
      SMK_CHECK_HIP(hipMemcpyAsync(dst_final, src, (size_t)H * W * sizeof(T),
                                   hipMemcpyDeviceToDevice, s));
    return SMK_OK;
  }
  std::vector<int> chunks;
  for (int left = npass_total; left > 0;) {
    const int c = left < MAX_R ? left : MAX_R;
    chunks.push_back(c);
    left -= c;
  }
  // launches alternate buffers: src -> X -> Y -> ... ; the last must be dst_final.
  // With src == dst_final an even launch count ends in dst_final; with src ==
  // other an odd count does.  Split a chunk to fix the parity, or copy.
  const bool src_is_final = (src == dst_final);
  const bool need_even = src_is_final || (src != other);
  // Generic source (neither buffer): first launch writes `other` if the count
  // is even, dst_final if odd -- so any count works when src is foreign.
  const bool foreign = (src != dst_final && src != other);
  if (!foreign) {
    const bool even = (chunks.size() % 2) == 0;
    if (need_even != even) {
      int k = -1;
      for (size_t q = 0; q < chunks.size(); ++q)
        if (chunks[q] >= 2) { k = (int)q; break; }
      if (k >= 0) {
        const int c = chunks[k];
        chunks[k] = c / 2;
        chunks.insert(chunks.begin() + k + 1, c - c / 2);
      } else {
        // single 1-pass chunk with the wrong parity: copy first
        T* to = src_is_final ? other : dst_final;
        SMK_CHECK_HIP(hipMemcpyAsync(to, src, (size_t)H * W * sizeof(T),
                                     hipMemcpyDeviceToDevice, s));
        src = to;
      }
    }
  }
  const int L = (int)chunks.size();
  // destination of launch q: counted back from the last, which is dst_final
  const T* in = src;
  int colour = first_colour;
  for (int q = 0; q < L; ++q) {
    T* out = (((L - 1 - q) % 2) == 0) ? dst_final : other;
    SmoothArgs<T> a{};
    a.p_in = in;
    a.p_out = out;
    a.rhs = rhs;
    a.flags = flags;
    a.H = H;
    a.W = W;
    a.alpha = alpha;
    a.omega = omega;
    a.npass = chunks[q];
    a.first_colour = colour;
    a.R = chunks[q];
    a.zero_init = 0;
    a.coarse_div = 1.0;
    int rc = launch_tiled<T>(mode, a, s);
    if (rc) return rc;
    colour = (colour + chunks[q]) & 1;
    in = out;
  }
  return SMK_OK;
}

template <typename T>
int residual_norm2_dev(const T* p, const T* rhs, const uint8_t* flags, int H, int W,
                       double alpha, double* partial, double* result, int slot, float* defect,
                       hipStream_t s) {
  dim3 g(cdiv(W, NB_X), cdiv(H, NB_Y));
  hipLaunchKernelGGL(k_residual_norm<T>, g, dim3(NB_T), 0, s, p, rhs, flags, H, W, alpha,
                     partial, defect);
  int rc = check_launch("k_residual_norm");
  if (rc) return rc;
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, partial, (int)(g.x * g.y),
                     result, slot);
  return check_launch("k_sum_partials");
}

template <typename T>
int rhs_norm2_dev(const T* rhs, const uint8_t* flags, int H, int W, double alpha,
                  double* partial, double* result, int slot, hipStream_t s) {
  dim3 g(cdiv(W, NB_X), cdiv(H, NB_Y));
  hipLaunchKernelGGL(k_rhs_norm<T>, g, dim3(NB_T), 0, s, rhs, flags, H, W, alpha, partial);
  int rc = check_launch("k_rhs_norm");
  if (rc) return rc;
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, partial, (int)(g.x * g.y),
                     result, slot);
  return check_launch("k_sum_partials");
}

inline double calc_alpha(double h, double rho, double dt) { return (-(h * h) * rho) / dt; }

inline double optimal_omega(int size) {
  const double w = 2.0 / (1.0 + sin(M_PI / (double)size));
  return w < 1.95 ? w : 1.95;
}

template <typename T>
int rbsor_solve_t(T* p, const T* div, const uint8_t* fmu, const uint8_t* fmv, int H, int W,
                  double rho, double h, double dt, int iterations, double omega, double tol,
                  int check_every, void* ws, smk_solve_info* info, hipStream_t s) {
  SolverWs w = carve_ws(ws, H, W, sizeof(T) == 8 ? SMK_F64 : SMK_F32);
  int rc = smk_cell_flags(fmu, fmv, H, W, w.flags, s);
  if (rc) return rc;
  const double alpha = calc_alpha(h, rho, dt);
  if (omega <= 0.0 || omega > 2.0) omega = optimal_omega(H > W ? H : W);
  T* tmp = reinterpret_cast<T*>(w.ptmp);
  if (tol <= 0.0) {
    rc = run_passes<T>(MODE_RB, p, tmp, p, div, w.flags, H, W, alpha, omega, 2 * iterations, 0,
                       s);
    if (rc) return rc;
    if (info) {
      info->iterations = iterations;
      info->converged = 0;
      info->rel_residual = -1.0;
    }
    return SMK_OK;
  }
  if (check_every < 1) check_every = 1;
  rc = rhs_norm2_dev<T>(div, w.flags, H, W, alpha, w.partial, w.result, 1, s);
  if (rc) return rc;
  rc = residual_norm2_dev<T>(p, div, w.flags, H, W, alpha, w.partial, w.result, 0, nullptr, s);
  if (rc) return rc;
  double host[2];
  SMK_CHECK_HIP(hipMemcpyAsync(host, w.result, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  SMK_CHECK_HIP(hipStreamSynchronize(s));
  const double b2 = host[1];
  double rel = b2 > 0.0 ? sqrt(host[0] / b2) : 0.0;
  int done = 0;
  while (rel > tol && done < iterations) {
    int k = check_every;
    if (done + k > iterations) k = iterations - done;
    rc = run_passes<T>(MODE_RB, p, tmp, p, div, w.flags, H, W, alpha, omega, 2 * k, 0, s);
    if (rc) return rc;
    done += k;
    rc = residual_norm2_dev<T>(p, div, w.flags, H, W, alpha, w.partial, w.result, 0, nullptr,
                               s);
    if (rc) return rc;
    SMK_CHECK_HIP(hipMemcpyAsync(host, w.result, sizeof(double), hipMemcpyDeviceToHost, s));
    SMK_CHECK_HIP(hipStreamSynchronize(s));
    rel = b2 > 0.0 ? sqrt(host[0] / b2) : 0.0;
  }
  if (info) {
    info->iterations = done;
    info->converged = rel <= tol;
    info->rel_residual = rel;
  }
  return SMK_OK;
}

// ------------------------------------------------------------------ multigrid
struct Level {
  int H, W;
  uint8_t* flags;   // owned
  uint8_t* solid;   // owned
  uint8_t* fmu;     // owned (coarse levels)
  uint8_t* fmv;     // owned (coarse levels)
  void* p[2];       // ping-pong; level 0: p[1] owned, p[0] = caller's p
  void* rhs;        // owned (coarse levels); level 0: caller's div
  int cur;          // which p holds the current iterate
  // FP32 hierarchy for mixed precision (allocated on first use)
  void* p32[2];
  void* rhs32;
  int cur32;
};

}  // namespace smk

struct smk_mg {
  int H, W, dtype, nl;
  std::vector<smk::Level> lv;
  double* partial;
  double* result;
  int npartial;
  int small_from[2];  // first level handled by k_vcycle_small: [native dtype, float32]
  int small_mode[2];  // smoother mode small_from was computed for
};

namespace smk {

static int mg_alloc(void** ptr, size_t bytes) {
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) return fail((int)e, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  return SMK_OK;
}

struct MgRun {
  int mode;
  double omega_override;
  int si, ci;
  int rediscretize;
  double rho, dt;
};

inline double level_omega(const MgRun& run, const Level& L) {
  if (run.mode != MODE_RB) return 1.0;
  return run.omega_override > 0.0 ? run.omega_override : optimal_omega(L.H > L.W ? L.H : L.W);
}

// First level from which the rest of the hierarchy fits one workgroup's LDS.
template <typename T>
int mg_small_from(smk_mg* mg, int mode) {
  const size_t lim = small_lds_limit<T>();
  int Hs[64], Ws[64];
  for (int l = 0; l < mg->nl; ++l) {
    Hs[l] = mg->lv[l].H;
    Ws[l] = mg->lv[l].W;
  }
  for (int from = 0; from < mg->nl; ++from) {
    if (mg->nl - from > SMALL_MAX) continue;
    if (small_lds_bytes(Hs + from, Ws + from, mg->nl - from, sizeof(T), mode) <= lim) return from;
  }
  return -1;  // the coarsest level alone does not fit: tiled kernels only
}

// Smooth level l with npass passes (tiled kernels), optionally prolonging a
// coarse correction in the first launch and restricting the residual or
// reducing the residual norm in the last.  Flips lv.cur per launch.
template <typename T>
int mg_smooth(std::vector<Level>& lv, int l, bool use32, int npass, bool zero_init,
              const T* coarse_e, int Hc, int Wc, T* coarse_rhs, double coarse_div,
              double* norm_partial, int* nparts, double alpha, double omega, int mode,
              hipStream_t s) {
  Level& L = lv[l];
  void** pb = use32 ? L.p32 : L.p;
  int& cur = use32 ? L.cur32 : L.cur;
  const T* rhs = reinterpret_cast<const T*>(use32 ? L.rhs32 : L.rhs);
  const bool tail = coarse_rhs != nullptr || norm_partial != nullptr;
  const int extra = tail ? 1 : 0;
  std::vector<int> chunks;
  for (int left = npass; left > 0;) {
    const int c = left < (MAX_R - extra) ? left : (MAX_R - extra);
    chunks.push_back(c);
    left -= c;
  }
  if (chunks.empty()) chunks.push_back(0);
  int colour = 0;
  for (size_t q = 0; q < chunks.size(); ++q) {
    const bool first = q == 0, last = q + 1 == chunks.size();
    SmoothArgs<T> a{};
    a.p_in = reinterpret_cast<const T*>(pb[cur]);
    a.p_out = reinterpret_cast<T*>(pb[cur ^ 1]);
    a.rhs = rhs;
    a.flags = L.flags;
    a.H = L.H;
    a.W = L.W;
    a.alpha = alpha;
    a.omega = omega;
    a.npass = chunks[q];
    a.first_colour = colour;
    a.R = chunks[q] + ((last && tail) ? 1 : 0);
    if (a.R < 1) a.R = 1;
    a.zero_init = (first && zero_init) ? 1 : 0;
    if (first && coarse_e != nullptr) {
      a.coarse_e = coarse_e;
      a.Hc = Hc;
      a.Wc = Wc;
    }
    a.coarse_div = 1.0;
    if (last && coarse_rhs != nullptr) {
      a.coarse_rhs = coarse_rhs;
      a.coarse_div = coarse_div;
    }
    if (last && norm_partial != nullptr) {
      a.norm_partial = norm_partial;
      *nparts = tiled_blocks(mode, L.H, L.W, a.R);
    }
    int rc = launch_tiled<T>(mode, a, s);
    if (rc) return rc;
    cur ^= 1;
    colour = (colour + chunks[q]) & 1;
  }
  return SMK_OK;
}

// Levels l .. nl-1 in one k_vcycle_small launch (in place on p[cur]).
template <typename T>
int mg_small(smk_mg* mg, int l, bool use32, const MgRun& run, double h, bool zero_init,
             hipStream_t s) {
  std::vector<Level>& lv = mg->lv;
  SmallArgs<T> a{};
  a.nl = mg->nl - l;
  double hl = h;
  for (int k = 0; k < a.nl; ++k) {
    const Level& L = lv[l + k];
    a.H[k] = L.H;
    a.W[k] = L.W;
    a.flags[k] = L.flags;
    a.alpha[k] = calc_alpha(hl, run.rho, run.dt);
    a.omega[k] = level_omega(run, L);
    a.coarse_div[k] = run.rediscretize ? a.alpha[k] : calc_alpha(hl * 2.0, run.rho, run.dt);
    hl *= 2.0;
  }
  Level& T0 = lv[l];
  void** pb = use32 ? T0.p32 : T0.p;
  const int cur = use32 ? T0.cur32 : T0.cur;
  a.rhs_top = reinterpret_cast<const T*>(use32 ? T0.rhs32 : T0.rhs);
  a.p_top_in = reinterpret_cast<const T*>(pb[cur]);
  a.p_top_out = reinterpret_cast<T*>(pb[cur]);
  a.zero_top = zero_init ? 1 : 0;
  a.npass_smooth = run.mode == MODE_RB ? 2 * run.si : run.si;
  a.npass_coarsest = run.mode == MODE_RB ? 2 * run.ci : run.ci;
  const size_t lds = small_lds_bytes(a.H, a.W, a.nl, sizeof(T), run.mode);
  return launch_small<T>(run.mode, a, lds, s);
}

// solver.py:94-139.  norm_partial (level 0 only): reduce ||r||^2 of the
// result in the last post-smooth launch; *nparts receives the slot count.
template <typename T>
int mg_vcycle(smk_mg* mg, int l, bool use32, const MgRun& run, double h, bool zero_init,
              double* norm_partial, int* nparts, hipStream_t s) {
  std::vector<Level>& lv = mg->lv;
  const int from = mg->small_from[use32 ? 1 : 0];
  if (from >= 0 && l >= from) {
    int rc = mg_small<T>(mg, l, use32, run, h, zero_init, s);
    if (rc || norm_partial == nullptr) return rc;
    *nparts = -1;  // caller computes the norm separately
    return SMK_OK;
  }
  Level& L = lv[l];
  const bool coarsest = (l == mg->nl - 1);
  const int n_iter = coarsest ? run.ci : run.si;
  const int npass = run.mode == MODE_RB ? 2 * n_iter : n_iter;
  const double alpha = calc_alpha(h, run.rho, run.dt);
  const double omega = level_omega(run, L);
  if (coarsest)
    return mg_smooth<T>(lv, l, use32, npass, zero_init, nullptr, 0, 0, nullptr, 1.0,
                        norm_partial, nparts, alpha, omega, run.mode, s);
  Level& C = lv[l + 1];
  T* crhs = reinterpret_cast<T*>(use32 ? C.rhs32 : C.rhs);
  const double coarse_div = run.rediscretize ? alpha : calc_alpha(h * 2.0, run.rho, run.dt);
  int rc = mg_smooth<T>(lv, l, use32, npass, zero_init, nullptr, 0, 0, crhs, coarse_div, nullptr,
                        nullptr, alpha, omega, run.mode, s);
  if (rc) return rc;
  rc = mg_vcycle<T>(mg, l + 1, use32, run, h * 2.0, true, nullptr, nullptr, s);
  if (rc) return rc;
  const T* ce = reinterpret_cast<const T*>(use32 ? C.p32[C.cur32] : C.p[C.cur]);
  return mg_smooth<T>(lv, l, use32, npass, false, ce, C.H, C.W, nullptr, 1.0, norm_partial,
                      nparts, alpha, omega, run.mode, s);
}

inline MgRun make_run(const smk_mg_params* prm, double rho, double dt) {
  MgRun run{};
  run.mode = prm->smoother == SMK_SMOOTH_JACOBI ? MODE_JACOBI : MODE_RB;
  run.omega_override = prm->smoother_omega;
  run.si = prm->smoother_iterations;
  run.ci = prm->coarsest_iterations;
  run.rediscretize = prm->coarse_scale == SMK_COARSE_REDISCRETIZED;
  run.rho = rho;
  run.dt = dt;
  return run;
}

template <typename T>
void mg_prepare_small(smk_mg* mg, int slot, int mode) {
  if (mg->small_mode[slot] != mode) {
    mg->small_from[slot] = mg_small_from<T>(mg, mode);
    mg->small_mode[slot] = mode;
  }
}

template <typename T>
int read_norm(smk_mg* mg, int nparts, const T* p, const T* div, double alpha, double* host,
              int count, hipStream_t s) {
  Level& L0 = mg->lv[0];
  int rc;
  if (nparts < 0) {
    rc = residual_norm2_dev<T>(p, div, L0.flags, mg->H, mg->W, alpha, mg->partial, mg->result, 0,
                               nullptr, s);
  } else {
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, mg->partial, nparts,
                       mg->result, 0);
    rc = check_launch("k_sum_partials");
  }
  if (rc) return rc;
  SMK_CHECK_HIP(hipMemcpyAsync(host, mg->result, count * sizeof(double), hipMemcpyDeviceToHost, s));
  SMK_CHECK_HIP(hipStreamSynchronize(s));
  return SMK_OK;
}

template <typename T>
int mg_solve_native(smk_mg* mg, T* p, const T* div, double rho, double h, double dt,
                    const smk_mg_params* prm, smk_solve_info* info, hipStream_t s) {
  Level& L0 = mg->lv[0];
  L0.p[0] = p;
  L0.rhs = const_cast<T*>(div);
  L0.cur = 0;
  const MgRun run = make_run(prm, rho, dt);
  mg_prepare_small<T>(mg, 0, run.mode);
  const double alpha = calc_alpha(h, rho, dt);
  int rc;
  int cycles = 0;
  double rel = -1.0;
  bool converged = false;
  if (prm->tolerance > 0.0) {
    rc = rhs_norm2_dev<T>(div, L0.flags, mg->H, mg->W, alpha, mg->partial, mg->result, 1, s);
    if (rc) return rc;
    double host[2];
    rc = read_norm<T>(mg, -1, p, div, alpha, host, 2, s);
    if (rc) return rc;
    const double b2 = host[1];
    rel = b2 > 0.0 ? sqrt(host[0] / b2) : 0.0;
    while (rel > prm->tolerance && cycles < prm->v_cycles) {
      int nparts = -1;
      rc = mg_vcycle<T>(mg, 0, false, run, h, false, mg->partial, &nparts, s);
      if (rc) return rc;
      cycles++;
      const T* pc = reinterpret_cast<const T*>(L0.p[L0.cur]);
      rc = read_norm<T>(mg, nparts, pc, div, alpha, host, 1, s);
      if (rc) return rc;
      rel = b2 > 0.0 ? sqrt(host[0] / b2) : 0.0;
    }
    converged = rel <= prm->tolerance;
  } else {
    for (int c = 0; c < prm->v_cycles; ++c) {
      rc = mg_vcycle<T>(mg, 0, false, run, h, false, nullptr, nullptr, s);
      if (rc) return rc;
      cycles++;
    }
  }
  if (L0.cur != 0) {
    SMK_CHECK_HIP(hipMemcpyAsync(p, L0.p[L0.cur], (size_t)mg->H * mg->W * sizeof(T),
                                 hipMemcpyDeviceToDevice, s));
    L0.cur = 0;
  }
  if (info) {
    info->iterations = cycles;
    info->converged = converged ? 1 : 0;
    info->rel_residual = rel;
    info->inner_cycles = 0;
    info->refinements = 0;
  }
  return SMK_OK;
}

static int mg_ensure32(smk_mg* mg) {
  for (int l = 0; l < mg->nl; ++l) {
    Level& L = mg->lv[l];
    if (L.rhs32 != nullptr) continue;
    const size_t n = (size_t)L.H * L.W * sizeof(float);
    int rc = mg_alloc(&L.p32[0], n);
    if (!rc) rc = mg_alloc(&L.p32[1], n);
    if (!rc) rc = mg_alloc(&L.rhs32, n);
    if (rc) return rc;
    L.cur32 = 0;
  }
  return SMK_OK;
}

// FP32 V-cycles inside FP64 defect correction: r = alpha*div - A p in FP64
// (with its norm), defect = r/alpha in FP32, e = k FP32 V-cycles from zero,
// p += e in FP64; repeat until ||r||/||alpha*div|| <= tolerance.
int mg_solve_mixed(smk_mg* mg, double* p, const double* div, double rho, double h, double dt,
                   const smk_mg_params* prm, smk_solve_info* info, hipStream_t s) {
  int rc = mg_ensure32(mg);
  if (rc) return rc;
  Level& L0 = mg->lv[0];
  const MgRun run = make_run(prm, rho, dt);
  mg_prepare_small<float>(mg, 1, run.mode);
  const double alpha = calc_alpha(h, rho, dt);
  const int H = mg->H, W = mg->W;
  const long n = (long)H * W;
  const int inner = prm->inner_cycles > 0 ? prm->inner_cycles : 1;
  const double tol = prm->tolerance;
  rc = rhs_norm2_dev<double>(div, L0.flags, H, W, alpha, mg->partial, mg->result, 1, s);
  if (rc) return rc;
  int refinements = 0, inner_total = 0;
  double rel = -1.0, b2 = 0.0;
  double host[2];
  for (;;) {
    // defect in FP32 into level-0 rhs32, norm^2 of r into result[0]
    rc = residual_norm2_dev<double>(p, div, L0.flags, H, W, alpha, mg->partial, mg->result, 0,
                                    reinterpret_cast<float*>(L0.rhs32), s);
    if (rc) return rc;
    SMK_CHECK_HIP(hipMemcpyAsync(host, mg->result, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    SMK_CHECK_HIP(hipStreamSynchronize(s));
    b2 = host[1];
    rel = b2 > 0.0 ? sqrt(host[0] / b2) : 0.0;
    if (tol > 0.0 ? rel <= tol : refinements >= prm->v_cycles) break;
    if (inner_total >= prm->v_cycles) break;
    for (int c = 0; c < inner; ++c) {
      rc = mg_vcycle<float>(mg, 0, true, run, h, c == 0, nullptr, nullptr, s);
      if (rc) return rc;
      inner_total++;
    }
    const float* e = reinterpret_cast<const float*>(L0.p32[L0.cur32]);
    hipLaunchKernelGGL(k_add_correction<double>, dim3(cdiv(n, 256)), dim3(256), 0, s, p, e, n);
    rc = check_launch("k_add_correction");
    if (rc) return rc;
    refinements++;
  }
  if (info) {
    info->iterations = inner_total;
    info->converged = (tol > 0.0 && rel <= tol) ? 1 : 0;
    info->rel_residual = rel;
    info->inner_cycles = inner_total;
    info->refinements = refinements;
  }
  return SMK_OK;
}

}  // namespace smk

using namespace smk;

extern "C" {

size_t smk_solver_workspace_bytes(int H, int W, int dtype) {
  if (H <= 0 || W <= 0) return 0;
  return solver_ws_bytes(H, W, dtype);
}

int smk_jacobi_solve(void* p, const void* div, const uint8_t* fmu, const uint8_t* fmv, int H,
                     int W, double rho, double h, double dt, int iterations, void* ws,
                     size_t ws_bytes, int dtype, void* stream) {
  SMK_REQUIRE(p && div && fmu && fmv && ws, "smk_jacobi_solve: null pointer");
  SMK_REQUIRE(H > 0 && W > 0 && iterations >= 0, "smk_jacobi_solve: bad arguments");
  SMK_REQUIRE(dtype == SMK_F64 || dtype == SMK_F32, "smk_jacobi_solve: bad dtype %d", dtype);
  SMK_REQUIRE(ws_bytes >= solver_ws_bytes(H, W, dtype),
              "smk_jacobi_solve: workspace too small (%zu < %zu)", ws_bytes,
              solver_ws_bytes(H, W, dtype));
  hipStream_t s = as_stream(stream);
  SolverWs w = carve_ws(ws, H, W, dtype);
  int rc = smk_cell_flags(fmu, fmv, H, W, w.flags, stream);
  if (rc) return rc;
  const double alpha = calc_alpha(h, rho, dt);
  if (dtype == SMK_F64)
    return run_passes<double>(MODE_JACOBI, (double*)p, (double*)w.ptmp, (const double*)p,
                              (const double*)div, w.flags, H, W, alpha, 1.0, iterations, 0, s);
  return run_passes<float>(MODE_JACOBI, (float*)p, (float*)w.ptmp, (const float*)p,
                           (const float*)div, w.flags, H, W, alpha, 1.0, iterations, 0, s);
}

int smk_rbsor_solve(void* p, const void* div, const uint8_t* fmu, const uint8_t* fmv, int H,
                    int W, double rho, double h, double dt, int iterations, double omega,
                    double tol, int check_every, void* ws, size_t ws_bytes,
                    smk_solve_info* info, int dtype, void* stream) {
  SMK_REQUIRE(p && div && fmu && fmv && ws, "smk_rbsor_solve: null pointer");
  SMK_REQUIRE(H > 0 && W > 0 && iterations >= 0, "smk_rbsor_solve: bad arguments");
  SMK_REQUIRE(dtype == SMK_F64 || dtype == SMK_F32, "smk_rbsor_solve: bad dtype %d", dtype);
  SMK_REQUIRE(ws_bytes >= solver_ws_bytes(H, W, dtype),
              "smk_rbsor_solve: workspace too small (%zu < %zu)", ws_bytes,
              solver_ws_bytes(H, W, dtype));
  hipStream_t s = as_stream(stream);
  if (dtype == SMK_F64)
    return rbsor_solve_t<double>((double*)p, (const double*)div, fmu, fmv, H, W, rho, h, dt,
                                 iterations, omega, tol, check_every, ws, info, s);
  return rbsor_solve_t<float>((float*)p, (const float*)div, fmu, fmv, H, W, rho, h, dt,
                              iterations, omega, tol, check_every, ws, info, s);
}

int smk_residual_norm2(const void* p, const void* div, const uint8_t* fmu, const uint8_t* fmv,
                       int H, int W, double alpha, void* ws, size_t ws_bytes, double* norm2,
                       int dtype, void* stream) {
  SMK_REQUIRE(p && div && fmu && fmv && ws && norm2, "smk_residual_norm2: null pointer");
  SMK_REQUIRE(H > 0 && W > 0, "smk_residual_norm2: bad shape");
  SMK_REQUIRE(dtype == SMK_F64 || dtype == SMK_F32, "smk_residual_norm2: bad dtype %d", dtype);
  SMK_REQUIRE(ws_bytes >= solver_ws_bytes(H, W, dtype), "smk_residual_norm2: workspace too small");
  hipStream_t s = as_stream(stream);
  SolverWs w = carve_ws(ws, H, W, dtype);
  int rc = smk_cell_flags(fmu, fmv, H, W, w.flags, stream);
  if (rc) return rc;
  if (dtype == SMK_F64)
    rc = residual_norm2_dev<double>((const double*)p, (const double*)div, w.flags, H, W, alpha,
                                    w.partial, w.result, 0, nullptr, s);
  else
    rc = residual_norm2_dev<float>((const float*)p, (const float*)div, w.flags, H, W, alpha,
                                   w.partial, w.result, 0, nullptr, s);
  if (rc) return rc;
  SMK_CHECK_HIP(hipMemcpyAsync(norm2, w.result, sizeof(double), hipMemcpyDeviceToHost, s));
  SMK_CHECK_HIP(hipStreamSynchronize(s));
  return SMK_OK;
}

int smk_mg_destroy(smk_mg* mg) {
  if (mg == nullptr) return SMK_OK;
  for (int l = 0; l < mg->nl; ++l) {
    Level& L = mg->lv[l];
    (void)hipFree(L.flags);
    (void)hipFree(L.solid);
    if (l > 0) {
      (void)hipFree(L.fmu);
      (void)hipFree(L.fmv);
      (void)hipFree(L.p[0]);
      (void)hipFree(L.rhs);
    }
    (void)hipFree(L.p[1]);
    (void)hipFree(L.p32[0]);
    (void)hipFree(L.p32[1]);
    (void)hipFree(L.rhs32);
  }
  (void)hipFree(mg->partial);
  (void)hipFree(mg->result);
  delete mg;
  return SMK_OK;
}

int smk_mg_create(const uint8_t* fmu, const uint8_t* fmv, int H, int W, int num_levels,
                  int dtype, void* stream, smk_mg** out) {
  SMK_REQUIRE(fmu && fmv && out, "smk_mg_create: null pointer");
  SMK_REQUIRE(H > 0 && W > 0, "smk_mg_create: bad shape");
  SMK_REQUIRE(dtype == SMK_F64 || dtype == SMK_F32, "smk_mg_create: bad dtype %d", dtype);
  *out = nullptr;
  hipStream_t s = as_stream(stream);
  smk_mg* mg = new smk_mg();
  mg->small_from[0] = mg->small_from[1] = -1;
  mg->small_mode[0] = mg->small_mode[1] = -1;
  mg->H = H;
  mg->W = W;
  mg->dtype = dtype;
  const size_t es = dtype == SMK_F64 ? 8 : 4;
  if (num_levels < 1) num_levels = 1;
  // solver.py:34-73: halve while both sizes are even, up to num_levels
  std::vector<std::pair<int, int>> dims{{H, W}};
  for (int l = 1; l < num_levels; ++l) {
    auto d = dims.back();
    if (d.first % 2 != 0 || d.second % 2 != 0) break;
    dims.push_back({d.first / 2, d.second / 2});
  }
  mg->nl = (int)dims.size();
  mg->lv.assign(mg->nl, Level{});
  int rc = SMK_OK;
  for (int l = 0; l < mg->nl && !rc; ++l) {
    Level& L = mg->lv[l];
    L.H = dims[l].first;
    L.W = dims[l].second;
    const size_t n = (size_t)L.H * L.W;
    rc = mg_alloc((void**)&L.flags, n);
    if (!rc) rc = mg_alloc((void**)&L.solid, n);
    if (!rc) rc = mg_alloc(&L.p[1], n * es);
    if (!rc && l > 0) {
      rc = mg_alloc(&L.p[0], n * es);
      if (!rc) rc = mg_alloc(&L.rhs, n * es);
      if (!rc) rc = mg_alloc((void**)&L.fmu, (size_t)L.H * (L.W + 1));
      if (!rc) rc = mg_alloc((void**)&L.fmv, (size_t)(L.H + 1) * L.W);
    }
  }
  mg->npartial = norm_blocks(H, W);
  if (!rc) rc = mg_alloc((void**)&mg->partial, (size_t)mg->npartial * sizeof(double));
  if (!rc) rc = mg_alloc((void**)&mg->result, 4 * sizeof(double));
  // level 0 flags from the caller's masks; solid = ~is_fluid (solver.py:38-40)
  if (!rc) rc = smk_cell_flags(fmu, fmv, H, W, mg->lv[0].flags, stream);
  if (!rc) {
    const long n = (long)H * W;
    hipLaunchKernelGGL(k_solid_from_flags, dim3(cdiv(n, 256)), dim3(256), 0, s,
                       mg->lv[0].flags, n, mg->lv[0].solid);
    rc = check_launch("k_solid_from_flags");
  }
  for (int l = 1; l < mg->nl && !rc; ++l) {
    Level& F = mg->lv[l - 1];
    Level& C = mg->lv[l];
    rc = smk_coarsen_solid(F.solid, F.H, F.W, C.solid, stream);
    if (!rc) rc = smk_face_masks_from_solid(C.solid, C.H, C.W, C.fmu, C.fmv, stream);
    if (!rc) rc = smk_cell_flags(C.fmu, C.fmv, C.H, C.W, C.flags, stream);
  }
  if (!rc) {
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = fail((int)e, "smk_mg_create: %s", hipGetErrorString(e));
  }
  if (rc) {
    smk_mg_destroy(mg);
    return rc;
  }
  *out = mg;
  return SMK_OK;
}

int smk_mg_num_levels(const smk_mg* mg) { return mg ? mg->nl : 0; }

int smk_mg_solve(smk_mg* mg, void* p, const void* div, double rho, double h, double dt,
                 const smk_mg_params* prm, smk_solve_info* info, void* stream) {
  SMK_REQUIRE(mg && p && div && prm, "smk_mg_solve: null pointer");
  SMK_REQUIRE(prm->smoother == SMK_SMOOTH_JACOBI || prm->smoother == SMK_SMOOTH_RBSOR,
              "smk_mg_solve: smoother must be SMK_SMOOTH_JACOBI or SMK_SMOOTH_RBSOR");
  SMK_REQUIRE(prm->v_cycles >= 0 && prm->smoother_iterations >= 0 &&
                  prm->coarsest_iterations >= 0,
              "smk_mg_solve: negative iteration count");
  SMK_REQUIRE(prm->coarse_scale == SMK_COARSE_REFERENCE ||
                  prm->coarse_scale == SMK_COARSE_REDISCRETIZED,
              "smk_mg_solve: bad coarse_scale");
  hipStream_t s = as_stream(stream);
  if (prm->precision == SMK_PREC_MIXED) {
    SMK_REQUIRE(mg->dtype == SMK_F64, "smk_mg_solve: mixed precision needs a float64 handle");
    return mg_solve_mixed(mg, (double*)p, (const double*)div, rho, h, dt, prm, info, s);
  }
  SMK_REQUIRE(prm->precision == SMK_PREC_NATIVE, "smk_mg_solve: bad precision %d", prm->precision);
  if (mg->dtype == SMK_F64)
    return mg_solve_native<double>(mg, (double*)p, (const double*)div, rho, h, dt, prm, info, s);
  return mg_solve_native<float>(mg, (float*)p, (const float*)div, rho, h, dt, prm, info, s);
}

}  // extern "C"
