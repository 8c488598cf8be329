// smk_solve.hip -- pressure solvers: Jacobi, red-black SOR, geometric
// multigrid (reference scaling or rediscretised), FP32-in-FP64 defect
// correction.
//
// Core kernels: the temporally blocked tile smoothers (k_smooth_tile64 and
// its variants: k_smooth_fine / _r / _d for the finest level, _r / _d for
// the others; the variants differ in region height, workgroup size and where
// p and the rhs live while the passes run).  A workgroup owns a TX x TY tile
// of the output, stages the tile plus an R-cell halo of p, rhs and the
// packed cell flags once, then runs `npass` colour passes (RB) or iterations
// (Jacobi) on chip.  After pass k only cells at distance >= k+1 from
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
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "smk_coarse_exact.hpp"
#include "smk_common.hpp"

namespace smk {

enum : int { MODE_RB = 0, MODE_JACOBI = 1, MODE_GS = 2 };

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
  // float fields only: the passes and residuals in float32 arithmetic (the
  // inner V-cycles of a mixed solve, which need only converge) instead of
  // the reference's float64 expression
  int fast_math;
  // the launch does nothing while *gate == 0 (device-side convergence)
  const int* gate;
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


// Branch-free stencil reads: the centre, the four neighbours and rhs are
// loaded unconditionally (index o must keep o-1, o+1, o-S, o+S in range),
// then the flags select which neighbour enters the sum, as the reference's
// masked conditional reads (residual.py:35-38) do.
template <typename T>
__device__ __forceinline__ void stencil_load(const T* cur, const T* rhs, int o, int S,
                                             uint8_t f, T& ce, T& nsum, T& b) {
  ce = cur[o];
  const T lv = cur[o - 1], rv = cur[o + 1], dv = cur[o - S], uv = cur[o + S];
  b = rhs[o];
  const T l = (f & F_LEFT) ? lv : ce;
  const T r = (f & F_RIGHT) ? rv : ce;
  const T d = (f & F_DOWN) ? dv : ce;
  const T u = (f & F_UP) ? uv : ce;
  nsum = l + r + d + u;
}

template <typename T>
__device__ __forceinline__ double relax_value(int mode, T ce, T nsum, T b, double alpha,
                                              double omega) {
  const double nw = ((double)nsum + alpha * (double)b) * 0.25;
  return mode == 0 ? (1.0 - omega) * (double)ce + omega * nw : nw;
}

template <typename T>
__device__ __forceinline__ double residual_value(T ce, T nsum, T b, double alpha) {
  return ((double)nsum + alpha * (double)b) - 4.0 * (double)ce;
}

// relax_value / residual_value in float32 arithmetic (SmoothArgs::fast_math)
__device__ __forceinline__ float relax_f(float ce, float ns, float b, float alpha, float omega) {
  const float nw = __builtin_fmaf(alpha, b, ns) * 0.25f;
  return __builtin_fmaf(omega, nw - ce, ce);
}
__device__ __forceinline__ float residual_f(float ce, float ns, float b, float alpha) {
  return __builtin_fmaf(-4.0f, ce, __builtin_fmaf(alpha, b, ns));
}

// Region = RH rows x 64 columns staged in LDS; lane x of every wavefront
// owns region column x, the NW waves split the rows.  R is the halo, so the
// output tile is (64 - 2R) x (RH - 2R).  Red-black passes walk row pairs:
// in rows (2m, 2m+1) lane x updates the one cell of the pass's colour, so
// every lane is busy and no index needs a division.
template <typename T, int MODE, int RH, int NT>
__device__ __forceinline__ void smooth_tile64_body(const SmoothArgs<T>& a, int bx, int by,
                                                   int part) {
  constexpr int RW = 64;
  constexpr int NW = NT / 64;
  __shared__ T sp[RH * RW];
  __shared__ T sr[RH * RW];
  __shared__ T sq[MODE == MODE_JACOBI ? RH * RW : 1];
  __shared__ uint8_t sf[RH * RW];
  __shared__ double red[NW];

  if (a.gate != nullptr && *a.gate == 0) return;
  const int R = a.R;
  const int TX = RW - 2 * R, TY = RH - 2 * R;
  const int H = a.H, W = a.W;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int gx0 = bx * TX - R, gy0 = by * TY - R;
  const int gx = gx0 + lane;
  const bool colin = gx >= 0 && gx < W;

  // ---- stage p (+ prolonged correction), rhs, flags.  Addresses are
  // clamped into the grid so every load is unconditional: all of a lane's
  // loads are issued back to back and the lane waits once, instead of once
  // per row behind a bounds branch.  Out-of-grid cells are then zeroed.
  static_assert(RH % (2 * NW) == 0, "row pairs must split evenly over the waves");
  constexpr int KR = RH / NW;
  // staged in two halves to bound the registers held by loads in flight
  constexpr int KH = (KR % 2 == 0) ? KR / 2 : KR;
  const int gxc = gx < 0 ? 0 : (gx >= W ? W - 1 : gx);
#pragma unroll
  for (int h0 = 0; h0 < KR; h0 += KH) {
    T pv[KH], rv[KH];
    uint8_t fv[KH];
#pragma unroll
    for (int k = 0; k < KH; ++k) {
      const int gy = gy0 + wave + (h0 + k) * NW;
      const int gyc = gy < 0 ? 0 : (gy >= H ? H - 1 : gy);
      const size_t g = (size_t)gyc * W + gxc;
      fv[k] = a.flags[g];
      rv[k] = a.rhs[g];
      pv[k] = a.zero_init ? (T)0 : a.p_in[g];
    }
    if (a.coarse_e != nullptr) {
      T corr[KH];
#pragma unroll
      for (int k = 0; k < KH; ++k) {
        const int gy = gy0 + wave + (h0 + k) * NW;
        const double xc = (gx + 0.5) * 0.5, yc = (gy + 0.5) * 0.5;
        corr[k] = (T)sample_field(a.coarse_e, a.Hc, a.Wc, xc - 0.5, yc - 0.5);
      }
#pragma unroll
      for (int k = 0; k < KH; ++k) pv[k] = pv[k] + ((fv[k] & F_LFLUID) ? corr[k] : (T)0);
    }
#pragma unroll
    for (int k = 0; k < KH; ++k) {
      const int ly = wave + (h0 + k) * NW;
      const int gy = gy0 + ly;
      const bool in = colin && gy >= 0 && gy < H;
      const int o = ly * RW + lane;
      sp[o] = in ? pv[k] : (T)0;
      sr[o] = in ? rv[k] : (T)0;
      sf[o] = in ? fv[k] : (uint8_t)0;
    }
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
      // Cells of one colour never neighbour each other: every new value of
      // the lane is computed before any is stored, and cells outside the
      // valid window read a safe interior index and are not stored.
      constexpr int K = RH / (2 * NW);
      T nv[K];
      int oo[K];
      bool st[K];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int ly = 2 * (wave + k * NW) + par;
        const bool valid = xin && ly >= lo && ly < RH - lo;
        const int o = valid ? ly * RW + lane : RW + 1;
        const uint8_t f = sf[o];
        T ce, ns, b;
        stencil_load(cur, sr, o, RW, f, ce, ns, b);
        if (sizeof(T) == 4 && a.fast_math)
          nv[k] = (T)relax_f((float)ce, (float)ns, (float)b, (float)a.alpha, (float)a.omega);
        else
          nv[k] = (T)relax_value(0, ce, ns, b, a.alpha, a.omega);
        oo[k] = o;
        st[k] = valid && (f & F_FLUID);
      }
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (st[k]) cur[oo[k]] = nv[k];
    } else {
      T* nxt = (cur == sp) ? sq : sp;
      constexpr int K = RH / NW;
      T nv[K];
      int oo[K];
      bool st[K];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int ly = wave + k * NW;
        const bool valid = xin && ly >= lo && ly < RH - lo;
        const int o = valid ? ly * RW + lane : RW + 1;
        const uint8_t f = sf[o];
        T ce, ns, b;
        stencil_load(cur, sr, o, RW, f, ce, ns, b);
        nv[k] = (f & F_FLUID) ? (T)relax_value(1, ce, ns, b, a.alpha, 1.0) : ce;
        oo[k] = o;
        st[k] = valid;
      }
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (st[k]) nxt[oo[k]] = nv[k];
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
      const int Jc = by * CY + cy, Ic = bx * CX + cx;
      if (Jc >= Hc || Ic >= Wc) continue;
      // the four children's residuals, stored in the level's dtype
      // (solver.py:116-117) and summed in float64 over the fluid children
      // (transfer.py:26-33); adding +0.0 for a non-fluid child is exact
      double total = 0.0;
      int count = 0;
#pragma unroll
      for (int dj = 0; dj < 2; ++dj)
#pragma unroll
        for (int di = 0; di < 2; ++di) {
          const int o = (R + 2 * cy + dj) * RW + R + 2 * cx + di;
          const uint8_t f = sf[o];
          T ce, ns, b;
          stencil_load(cur, sr, o, RW, f, ce, ns, b);
          const T rv = (sizeof(T) == 4 && a.fast_math && MODE == MODE_RB)
                           ? (T)residual_f((float)ce, (float)ns, (float)b, (float)a.alpha)
                           : (T)residual_value(ce, ns, b, a.alpha);
          const bool fl = f & F_LFLUID;
          total += fl ? (double)rv : 0.0;
          count += fl ? 1 : 0;
        }
      const T mean = count > 0 ? (T)(total / (double)count) : (T)0;
      a.coarse_rhs[(size_t)Jc * Wc + Ic] = (T)((double)mean / a.coarse_div);
    }
  }

  // ---- fused residual norm over the fluid interior cells
  if (a.norm_partial != nullptr) {
    double acc = 0.0;
#pragma unroll 4
    for (int ly = R + wave; ly < R + TY; ly += NW) {
      const int gy = gy0 + ly;
      const int o = ly * RW + (xout ? lane : R);
      const uint8_t f = sf[o];
      T ce, ns, b;
      stencil_load(cur, sr, o, RW, f, ce, ns, b);
      const double rv = (sizeof(T) == 4 && a.fast_math && MODE == MODE_RB)
                            ? (double)residual_f((float)ce, (float)ns, (float)b, (float)a.alpha)
                            : residual_value(ce, ns, b, a.alpha);
      acc += (xout && gy < H && (f & F_LFLUID)) ? rv * rv : 0.0;
    }
    const double tot = block_sum<NT>(acc, red);
    if (threadIdx.x == 0) a.norm_partial[part] = tot;
  }
}

template <typename T, int MODE, int RH, int NT>
__global__ __launch_bounds__(NT) void k_smooth_tile64(const SmoothArgs<T> a) {
  int bx, by;
  xcd_tile(bx, by);
  smooth_tile64_body<T, MODE, RH, NT>(a, bx, by, by * gridDim.x + bx);
}

template <typename T, int RH, int NT>
__global__ __launch_bounds__(NT) void k_smooth_fine(const SmoothArgs<T> a) {
  int bx, by;
  xcd_tile(bx, by);
  smooth_tile64_body<T, MODE_RB, RH, NT>(a, bx, by, by * gridDim.x + bx);
}

// Red-black tile smoothing with only p (and the flags, for the restriction)
// staged in LDS.  Lane x of wave w updates the cells of region column x in
// the row pairs m = w + k*NW (k < K), one cell per pair and pass, so the
// right-hand side and flags of those 2K cells stay in the lane's registers
// for the whole launch, indexed by colour: the colour-c cell of pair m is
// row 2m + ((c + x) & 1).  Half the LDS of smooth_tile64_body, so more
// workgroups share a CU, and two fewer LDS reads per cell update.  The
// per-cell arithmetic is the same expression in the same order.  Residuals
// for the restriction are computed by the owning lane and passed through the
// p array once the passes are done.
//
// These launches are bound by vector-ALU issue as much as by memory (a
// wave64 VALU instruction occupies its SIMD for four cycles), so whatever
// depends only on a row is computed once per wavefront: the wave index is
// made uniform, row indices, clamps and row base addresses then live in
// scalar registers, and a lane's loads and stores add only its fixed column
// offset to a scalar row base.
// The index of the calling thread's wavefront, uniform over it.
__device__ __forceinline__ int wave_id() {
  return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

// Element `off` bytes into a row whose base is uniform over the wavefront.
template <typename T>
__device__ __forceinline__ const T& row_at(const T* row, unsigned off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(row) + off);
}
template <typename T>
__device__ __forceinline__ T& row_at(T* row, unsigned off) {
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(row) + off);
}

__device__ __forceinline__ int clampi(int v, int n) { return v < 0 ? 0 : (v >= n ? n - 1 : v); }

// Per-tile state of a lane: the rhs of its 2K cells by colour, and the
// wave-uniform mask of pairs whose cells are fluid with all four faces open
// in every lane (the stencil then needs no selects: a scalar branch).  The
// flags of the other pairs are read from the LDS copy when needed (cells at
// walls and obstacles), so no lane masks stay live across the passes.
template <typename T, int RH, int NT>
struct RTile {
  static constexpr int RW = 64;
  static constexpr int NW = NT / 64;
  static_assert(RH % (2 * NW) == 0, "row pairs must split evenly over the waves");
  static constexpr int K = RH / (2 * NW);  // row pairs per lane
  T rb[2][K];  // rhs of the colour-c cell of pair k
  uint32_t open[2];
};

// p is staged in an LDS array of RH + 2 rows: the region's rows 0 .. RH-1
// and one row of padding above and below, so that every cell of the region
// can read its four neighbours in range.  The passes then update every cell
// of a row pair, not only those of the shrinking exact window: after pass k
// the cells at distance >= k from the region's edge are exact whatever the
// others hold (an exact cell reads only cells at distance >= k - 1 of the
// other colour), so the extra updates change nothing that is written out,
// and no lane needs a window test.
template <typename T, int RH>
struct SPad {
  static constexpr int N = (RH + 2) * 64;  // elements of the LDS array
  static constexpr int OFF = 64;           // element of region row 0, column 0
};

// The coarse-grid patch the bilinear prolongation of a 64 x RH region reads:
// region column gx samples the coarse grid at x = (gx + 0.5) / 2 - 0.5, so the
// region's samples touch coarse columns (gx0 - 1) >> 1 .. that + 33 (and rows
// likewise), clamped into the coarse grid exactly as sample_field clamps.
// Staged in LDS with the tile, the prolongation then costs no second trip to
// memory.
template <typename T, int RH, int NT>
struct CPatch {
  static constexpr int PW = 34, PH = RH / 2 + 2, N = PH * PW;
  static constexpr int PER = (N + NT - 1) / NT;  // values per thread
};

// Issue the patch loads (coarse columns cx0 .., rows cy0 ..) into registers.
template <typename T, int RH, int NT>
__device__ __forceinline__ void cpatch_issue(const SmoothArgs<T>& a, int bx, int by,
                                             T (&v)[CPatch<T, RH, NT>::PER]) {
  using CP = CPatch<T, RH, NT>;
  const int R = a.R, TX = 64 - 2 * R, TY = RH - 2 * R;
  const int cx0 = (bx * TX - R - 1) >> 1, cy0 = (by * TY - R - 1) >> 1;
#pragma unroll
  for (int i = 0; i < CP::PER; ++i) {
    const int idx = threadIdx.x + i * NT;
    const int py = idx / CP::PW, px = idx - py * CP::PW;
    const int gy = clampi(cy0 + py, a.Hc), gx = clampi(cx0 + px, a.Wc);
    v[i] = idx < CP::N ? a.coarse_e[(size_t)gy * a.Wc + gx] : (T)0;
  }
}

template <typename T, int RH, int NT>
__device__ __forceinline__ void cpatch_store(const T (&v)[CPatch<T, RH, NT>::PER], T* spatch) {
  using CP = CPatch<T, RH, NT>;
#pragma unroll
  for (int i = 0; i < CP::PER; ++i) {
    const int idx = threadIdx.x + i * NT;
    if (idx < CP::N) spatch[idx] = v[i];
  }
}

// The bilinear prolongation of the staged coarse patch, sample_field's value:
// the sample position of fine index g is g / 2 - 0.25, never an integer, so
// sample_field's floor and ceil are (g - 1) >> 1 and one more; with the same
// clamping, weights (x minus the clamped floor, exact) and lerps (the first
// two differences in T) the value is sample_field's.  The x part depends only
// on the lane's column (PCol, once per tile); the lerp along x of one coarse
// row (prolong_xlerp) is shared by the fine rows that sample that row: fine
// rows g and g + 1 of a row pair read coarse rows r, r + 1 and, for even g,
// r + 2 (r = (g - 1) >> 1), three x-lerps for two cells instead of four.
struct PCol {
  int c0, c1;  // patch columns of the lane's clamped coarse floor and ceil
  double tx;
};

template <typename T, int RH, int NT>
__device__ __forceinline__ PCol prolong_col(const SmoothArgs<T>& a, int bx) {
  const int R = a.R, TX = 64 - 2 * R;
  const int cx0 = (bx * TX - R - 1) >> 1;
  const int gx = bx * TX - R + (int)(threadIdx.x & 63);
  const int xf = (gx - 1) >> 1;
  const int x0 = clampi(xf, a.Wc), x1 = clampi(xf + 1, a.Wc);
  PCol c;
  c.c0 = x0 - cx0;
  c.c1 = x1 - cx0;
  c.tx = ((gx + 0.5) * 0.5 - 0.5) - (double)x0;
  return c;
}

template <typename T, int RH, int NT>
__device__ __forceinline__ double prolong_xlerp(const SmoothArgs<T>& a, const T* spatch,
                                                const PCol& pc, int cy0, int y) {
  const int r = (clampi(y, a.Hc) - cy0) * CPatch<T, RH, NT>::PW;  // uniform
  const T c00 = spatch[r + pc.c0], c10 = spatch[r + pc.c1];
  return (double)c00 + (double)(T)(c10 - c00) * pc.tx;
}

// y weight of fine row gy: its position minus the clamped coarse floor
__device__ __forceinline__ double prolong_ty(int gy, int Hc) {
  return ((gy + 0.5) * 0.5 - 0.5) - (double)clampi((gy - 1) >> 1, Hc);
}

// The prolonged corrections of fine rows gya and gya + 1 (gya uniform) at the
// lane's column.
template <typename T, int RH, int NT>
__device__ __forceinline__ void prolong_pair(const SmoothArgs<T>& a, const T* spatch,
                                             const PCol& pc, int cy0, int gya, T& ca, T& cb) {
  const int r0 = (gya - 1) >> 1;
  const double l0 = prolong_xlerp<T, RH, NT>(a, spatch, pc, cy0, r0);
  const double l1 = prolong_xlerp<T, RH, NT>(a, spatch, pc, cy0, r0 + 1);
  const double tya = prolong_ty(gya, a.Hc), tyb = prolong_ty(gya + 1, a.Hc);
  ca = (T)(l0 + (l1 - l0) * tya);
  if (gya & 1) {  // row gya + 1 samples coarse rows r0 and r0 + 1 too
    cb = (T)(l0 + (l1 - l0) * tyb);
  } else {
    const double l2 = prolong_xlerp<T, RH, NT>(a, spatch, pc, cy0, r0 + 2);
    cb = (T)(l1 + (l2 - l1) * tyb);
  }
}

// Raw loads of the tile's rows for the lane's pairs h0 .. h0+KH-1 (clamped
// into the grid, unconditional: a lane issues them back to back).
template <typename T, int RH, int NT, int KH>
__device__ __forceinline__ void rtile_issue(const SmoothArgs<T>& a, int bx, int by, int h0,
                                            T (&pv)[KH][2], T (&rv)[KH][2],
                                            uint32_t (&fv)[KH][2]) {
  using RT = RTile<T, RH, NT>;
  const int R = a.R;
  const int TX = RT::RW - 2 * R, TY = RH - 2 * R;
  const int lane = threadIdx.x & 63, wave = wave_id();
  const int gx = bx * TX - R + lane, gy0 = by * TY - R;
  const unsigned gxc = (unsigned)clampi(gx, a.W), gxo = gxc * (unsigned)sizeof(T);
#pragma unroll
  for (int k = 0; k < KH; ++k)
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int gy = gy0 + 2 * (wave + (h0 + k) * RT::NW) + q;
      const size_t row = (size_t)clampi(gy, a.H) * a.W;
      fv[k][q] = row_at(a.flags + row, gxc);
      rv[k][q] = row_at(a.rhs + row, gxo);
      pv[k][q] = a.zero_init ? (T)0 : row_at(a.p_in + row, gxo);
    }
}

// Apply the prolonged correction (spatch: the staged coarse patch, or null),
// zero what lies outside the grid, and place p and the flags in LDS and the
// rhs and flags in the lane's registers.
template <typename T, int RH, int NT, int KH>
__device__ __forceinline__ void rtile_commit(const SmoothArgs<T>& a, int bx, int by, int h0,
                                             T (&pv)[KH][2], const T (&rv)[KH][2],
                                             const uint32_t (&fv)[KH][2], T* sp, uint8_t* sf,
                                             RTile<T, RH, NT>& t, const T* spatch,
                                             const PCol& pc) {
  using RT = RTile<T, RH, NT>;
  constexpr uint32_t F_ALL = F_FLUID | F_LFLUID | F_LEFT | F_RIGHT | F_DOWN | F_UP;
  const int R = a.R;
  const int TX = RT::RW - 2 * R, TY = RH - 2 * R;
  const int lane = threadIdx.x & 63, wave = wave_id();
  const int px = lane & 1;
  const int gx = bx * TX - R + lane, gy0 = by * TY - R;
  const bool colin = gx >= 0 && gx < a.W;
  if (spatch != nullptr) {  // the prolonged correction, on fluid cells
    const int cy0 = (gy0 - 1) >> 1;
#pragma unroll
    for (int k = 0; k < KH; ++k) {
      T ca, cb;
      prolong_pair<T, RH, NT>(a, spatch, pc, cy0, gy0 + 2 * (wave + (h0 + k) * RT::NW), ca, cb);
      pv[k][0] = (fv[k][0] & F_LFLUID) ? pv[k][0] + ca : pv[k][0];
      pv[k][1] = (fv[k][1] & F_LFLUID) ? pv[k][1] + cb : pv[k][1];
    }
  }
#pragma unroll
  for (int k = 0; k < KH; ++k) {
    uint32_t f[2];
    T r[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int ly = 2 * (wave + (h0 + k) * RT::NW) + q;
      const int gy = gy0 + ly;
      const bool in = colin && gy >= 0 && gy < a.H;
      const int o = ly * RT::RW + lane;
      f[q] = in ? fv[k][q] : 0u;
      r[q] = in ? rv[k][q] : (T)0;
      sp[o] = in ? pv[k][q] : (T)0;
      sf[o] = (uint8_t)f[q];
    }
    // row 2m + q has colour (q + x) & 1
    const int kk = h0 + k;
    t.rb[0][kk] = px ? r[1] : r[0];
    t.rb[1][kk] = px ? r[0] : r[1];
    const uint32_t f0 = px ? f[1] : f[0], f1 = px ? f[0] : f[1];
    if (__ballot((f0 & F_ALL) == F_ALL) == ~0ull) t.open[0] |= 1u << kk;
    if (__ballot((f1 & F_ALL) == F_ALL) == ~0ull) t.open[1] |= 1u << kk;
  }
}

template <typename T, int RH, int NT>
__device__ __forceinline__ void rtile_reset(RTile<T, RH, NT>& t) {
  t.open[0] = t.open[1] = 0u;
}

// OPEN: every pair of the lane's colour-C cells is open in every lane (no
// selects), decided once per pass rather than per cell, so the pass is one
// straight run of code.  A cell's reads are its own value and cells of the
// other colour, which the pass does not write, so a pair's new values are
// stored before the next pair reads.  Every cell of a pair is updated
// (SPad), the fluid flag alone deciding the store.  With SKIP (launches of
// many passes, where the exact window shrinks far) a pair whose two rows
// both lie outside the pass's exact window (rows lo .. RH-lo-1) is skipped
// by the whole wavefront; the test costs more than it saves at 4 passes.
template <int C, bool OPEN, bool FM, bool SKIP, typename T, int RH, int NT>
__device__ __forceinline__ void rtile_sweep_k(const SmoothArgs<T>& a, const RTile<T, RH, NT>& t,
                                              T* sp, const uint8_t* sf, int lo) {
  using RT = RTile<T, RH, NT>;
  constexpr int K = RT::K, RW = RT::RW, NW = RT::NW;
  const int lane = threadIdx.x & 63, wave = wave_id();
  const int par = (lane + C) & 1;
  const int o0 = (2 * wave + par) * RW + lane;  // the lane's cell of pair 0
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int lyp = 2 * (wave + k * NW);  // uniform: the pair's first row
    if (SKIP && (lyp + 1 < lo || lyp >= RH - lo)) continue;
    const int o = o0 + 2 * k * NW * RW;
    const T ce = sp[o];
    const T lv = sp[o - 1], rv = sp[o + 1], dv = sp[o - RW], uv = sp[o + RW];
    T ns;
    bool fl;
    if (OPEN) {
      ns = lv + rv + dv + uv;
      fl = true;
    } else {
      const uint32_t f = sf[o];
      const T l = (f & F_LEFT) ? lv : ce;
      const T r = (f & F_RIGHT) ? rv : ce;
      const T d = (f & F_DOWN) ? dv : ce;
      const T u = (f & F_UP) ? uv : ce;
      ns = l + r + d + u;
      fl = (f & F_FLUID) != 0u;
    }
    T nv;
    if constexpr (FM && sizeof(T) == 4)
      nv = relax_f(ce, ns, t.rb[C][k], (float)a.alpha, (float)a.omega);
    else
      nv = (T)relax_value(0, ce, ns, t.rb[C][k], a.alpha, a.omega);
    if (fl) sp[o] = nv;
  }
}

template <int C, bool FM, bool SKIP, typename T, int RH, int NT>
__device__ __forceinline__ void rtile_sweep(const SmoothArgs<T>& a, const RTile<T, RH, NT>& t,
                                            T* sp, const uint8_t* sf, int lo) {
  constexpr uint32_t ALL = (1u << RTile<T, RH, NT>::K) - 1u;
  if ((t.open[C] & ALL) == ALL)  // wave-uniform
    rtile_sweep_k<C, true, FM, SKIP>(a, t, sp, sf, lo);
  else
    rtile_sweep_k<C, false, FM, SKIP>(a, t, sp, sf, lo);
}

template <bool FM, bool SKIP, typename T, int RH, int NT>
__device__ __forceinline__ void rtile_passes_m(const SmoothArgs<T>& a,
                                               const RTile<T, RH, NT>& t, T* sp,
                                               const uint8_t* sf) {
  for (int pass = 0; pass < a.npass; ++pass) {
    if (((a.first_colour + pass) & 1) == 0)
      rtile_sweep<0, FM, SKIP>(a, t, sp, sf, pass + 1);
    else
      rtile_sweep<1, FM, SKIP>(a, t, sp, sf, pass + 1);
    __syncthreads();
  }
}

// Launches of more passes than this skip the row pairs outside each pass's
// exact window (single-grid solves run up to 24 passes per launch).
constexpr int RTILE_SKIP_PASSES = 8;

// npass colour passes over the region (the exact window shrinks by one cell
// per pass).
template <typename T, int RH, int NT>
__device__ __forceinline__ void rtile_passes(const SmoothArgs<T>& a, const RTile<T, RH, NT>& t,
                                             T* sp, const uint8_t* sf) {
  const bool skip = a.npass > RTILE_SKIP_PASSES;
  if (sizeof(T) == 4 && a.fast_math) {
    if (skip)
      rtile_passes_m<true, true>(a, t, sp, sf);
    else
      rtile_passes_m<true, false>(a, t, sp, sf);
  } else {
    if (skip)
      rtile_passes_m<false, true>(a, t, sp, sf);
    else
      rtile_passes_m<false, false>(a, t, sp, sf);
  }
}

// Write the tile interior of p; then the residuals of the lane's own cells
// in the output window feed the restriction (through the p array) and the
// residual norm.
template <typename T, int RH, int NT>
__device__ __forceinline__ void rtile_epilogue(const SmoothArgs<T>& a, const RTile<T, RH, NT>& t,
                                               int bx, int by, int part, T* sp,
                                               const uint8_t* sf, double* red) {
  using RT = RTile<T, RH, NT>;
  constexpr int K = RT::K, RW = RT::RW, NW = RT::NW;
  const int R = a.R;
  const int TX = RW - 2 * R, TY = RH - 2 * R;
  const int H = a.H, W = a.W;
  const int lane = threadIdx.x & 63, wave = wave_id();
  const int gx = bx * TX - R + lane, gy0 = by * TY - R;
  const bool colin = gx >= 0 && gx < W;
  const bool xout = lane >= R && lane < R + TX && colin;
  const unsigned gxo = (unsigned)(colin ? gx : 0) * (unsigned)sizeof(T);
  const int lyend = min(R + TY, H - gy0);  // uniform: output rows end
  for (int ly = R + wave; ly < lyend; ly += NW)
    if (xout) row_at(a.p_out + (size_t)(gy0 + ly) * W, gxo) = sp[ly * RW + lane];
  const bool want_norm = a.norm_partial != nullptr;
  if (a.coarse_rhs == nullptr && !want_norm) return;

  T res[2][K];
  double acc = 0.0;
  constexpr uint32_t ALL = (1u << K) - 1u;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int par = (lane + c) & 1;
    // as in rtile_sweep_k: one wave-uniform decision for the colour's pairs
    const bool open_all = (t.open[c] & ALL) == ALL;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      // every cell of the region (SPad); only the output window's are used
      const int ly = 2 * (wave + k * NW) + par;
      const int o = ly * RW + lane;
      const T ce = sp[o];
      const T lv = sp[o - 1], rv = sp[o + 1], dv = sp[o - RW], uv = sp[o + RW];
      T ns;
      bool fl;
      if (open_all) {
        ns = lv + rv + dv + uv;
        fl = true;
      } else {
        const uint32_t f = sf[o];
        const T l = (f & F_LEFT) ? lv : ce;
        const T r = (f & F_RIGHT) ? rv : ce;
        const T d = (f & F_DOWN) ? dv : ce;
        const T u = (f & F_UP) ? uv : ce;
        ns = l + r + d + u;
        fl = (f & F_LFLUID) != 0u;
      }
      const double rr = (sizeof(T) == 4 && a.fast_math)
                            ? (double)residual_f((float)ce, (float)ns, (float)t.rb[c][k],
                                                 (float)a.alpha)
                            : residual_value(ce, ns, t.rb[c][k], a.alpha);
      res[c][k] = (T)rr;
      if (want_norm) {
        const bool win = xout && ly >= R && ly < lyend;
        acc += (win && fl) ? rr * rr : 0.0;
      }
    }
  }

  if (a.coarse_rhs != nullptr) {
    __syncthreads();  // every lane has read p
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int par = (lane + c) & 1;
#pragma unroll
      for (int k = 0; k < K; ++k) sp[(2 * (wave + k * NW) + par) * RW + lane] = res[c][k];
    }
    __syncthreads();
    const int CX = TX / 2, CY = TY / 2;
    const int Wc = W / 2, Hc = H / 2;
    const int cx = lane & 31;
    const int Ic = bx * CX + cx;
    const bool cxin = cx < CX && Ic < Wc;
    const unsigned co = (unsigned)(cxin ? Ic : 0) * (unsigned)sizeof(T);
    // one coarse row per half-wavefront: rows cyb and cyb + 1 of a wave
    const int cyend = min(CY, Hc - by * CY);  // uniform
    for (int cyb = 2 * wave; cyb < cyend; cyb += 2 * NW) {
      const int cy = cyb + (lane >> 5);
      if (cy >= cyend || !cxin) continue;
      // the children's residuals summed in float64 over the fluid children
      // in the (dj, di) order of smooth_tile64_body
      double total = 0.0;
      int count = 0;
#pragma unroll
      for (int dj = 0; dj < 2; ++dj)
#pragma unroll
        for (int di = 0; di < 2; ++di) {
          const int o = (R + 2 * cy + dj) * RW + R + 2 * cx + di;
          const bool fl = sf[o] & F_LFLUID;
          const T rv = sp[o];
          total += fl ? (double)rv : 0.0;
          count += fl ? 1 : 0;
        }
      // total / count, exact scalings for 4, 2 and 1 children (the quotient
      // by a power of two is the product by its reciprocal, both correctly
      // rounded); no children gives 0 as before
      const double mean_d = count == 3 ? total / 3.0
                                       : total * (count == 4 ? 0.25 : (count == 2 ? 0.5 : 1.0));
      const T mean = (T)mean_d;
      row_at(a.coarse_rhs + (size_t)(by * CY + cy) * Wc, co) = (T)((double)mean / a.coarse_div);
    }
  }

  if (want_norm) {
    const double tot = block_sum<NT>(acc, red);
    if (threadIdx.x == 0) a.norm_partial[part] = tot;
  }
}

// The prolonged correction added to the lane's fluid cells in LDS (the p of
// the region is staged; the patch is in LDS; a barrier follows).
template <typename T, int RH, int NT>
__device__ __forceinline__ void rtile_prolong(const SmoothArgs<T>& a, int bx, int by, T* sp,
                                              const uint8_t* sf, const T* spatch) {
  using RT = RTile<T, RH, NT>;
  constexpr int K = RT::K, RW = RT::RW, NW = RT::NW;
  const int R = a.R;
  const int TY = RH - 2 * R;
  const int lane = threadIdx.x & 63, wave = wave_id();
  const int gy0 = by * TY - R;
  const int cy0 = (gy0 - 1) >> 1;
  const PCol pc = prolong_col<T, RH, NT>(a, bx);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    T corr[2];
    prolong_pair<T, RH, NT>(a, spatch, pc, cy0, gy0 + 2 * (wave + k * NW), corr[0], corr[1]);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int ly = 2 * (wave + k * NW) + q;
      const int o = ly * RW + lane;
      sp[o] = (sf[o] & F_LFLUID) ? sp[o] + corr[q] : sp[o];
    }
  }
}

template <typename T, int RH, int NT>
__device__ __forceinline__ void smooth_tile64r_body(const SmoothArgs<T>& a, int bx, int by,
                                                    int part) {
  using RT = RTile<T, RH, NT>;
  using CP = CPatch<T, RH, NT>;
  __shared__ T sp_[SPad<T, RH>::N];
  T* sp = sp_ + SPad<T, RH>::OFF;
  __shared__ uint8_t sf[RH * RT::RW];
  __shared__ double red[RT::NW];
  __shared__ T spatch[CP::N];
  if (a.gate != nullptr && *a.gate == 0) return;
  RT t;
  rtile_reset(t);
  const bool prolong = a.coarse_e != nullptr;
  T cv[CP::PER];
  PCol pc{};
  if (prolong) {
    cpatch_issue<T, RH, NT>(a, bx, by, cv);  // overlaps the tile's loads
    pc = prolong_col<T, RH, NT>(a, bx);
  }
  // in two halves of the lane's pairs (quarters from eight pairs on), to
  // bound the registers of loads in flight
  constexpr int K = RT::K;
  constexpr int KH = K >= 8 ? K / 4 : ((K % 2 == 0) ? K / 2 : K);
#pragma unroll
  for (int h0 = 0; h0 < K; h0 += KH) {
    T pv[KH][2], rv[KH][2];
    uint32_t fv[KH][2];
    rtile_issue<T, RH, NT, KH>(a, bx, by, h0, pv, rv, fv);
    if (h0 == 0 && prolong) {  // the whole patch in LDS before any lane samples it
      cpatch_store<T, RH, NT>(cv, spatch);
      __syncthreads();
    }
    rtile_commit<T, RH, NT, KH>(a, bx, by, h0, pv, rv, fv, sp, sf, t,
                                prolong ? spatch : nullptr, pc);
  }
  t.open[0] = __builtin_amdgcn_readfirstlane(t.open[0]);
  t.open[1] = __builtin_amdgcn_readfirstlane(t.open[1]);
  __syncthreads();
  rtile_passes(a, t, sp, sf);
  rtile_epilogue(a, t, bx, by, part, sp, sf, red);
}

// The p rows of the region go to LDS by LDS-DMA (global_load_lds_dword): no
// register holds them in flight, so every row of the tile is requested at
// once and staging costs one memory latency.  One instruction moves 64
// dwords of one row (wave-uniform LDS destination, lane i -> dword i).
// Addresses are clamped into the grid, so a slot outside the grid holds the
// nearest edge value; such slots get zero flags, are never updated, and no
// in-grid cell reads them (the faces on the grid's edge are closed).
template <typename T, int RH, int NT>
__device__ __forceinline__ void rtile_dma_p(const SmoothArgs<T>& a, int bx, int by, T* sp) {
  constexpr int RW = 64, NW = NT / 64;
  constexpr int EPD = (int)sizeof(T) / 4;  // dwords per value
  constexpr int CPI = 64 / EPD;            // values per instruction
  const int lane = threadIdx.x & 63, wave = wave_id();
  const int R = a.R, TX = RW - 2 * R, TY = RH - 2 * R;
  const int gx0 = bx * TX - R, gy0 = by * TY - R;
#pragma unroll
  for (int h = 0; h < EPD; ++h) {
    const unsigned off =
        (unsigned)clampi(gx0 + h * CPI + lane / EPD, a.W) * (unsigned)sizeof(T) + 4u * (lane % EPD);
#pragma unroll
    for (int k = 0; k < RH / NW; ++k) {
      const int ly = wave + k * NW;
      const T* row = a.p_in + (size_t)clampi(gy0 + ly, a.H) * a.W;
      __builtin_amdgcn_global_load_lds(
          &row_at(row, off), (__attribute__((address_space(3))) void*)(sp + ly * RW + h * CPI),
          4, 0, 0);
    }
  }
}

// rhs and flags of the lane's cells into its registers (all pairs at once, or
// in groups of four pairs from eight pairs on, to bound the registers of loads
// in flight), the flags also into LDS for the restriction.
template <typename T, int RH, int NT>
__device__ __forceinline__ void rtile_stage_rf(const SmoothArgs<T>& a, int bx, int by,
                                               uint8_t* sf, RTile<T, RH, NT>& t) {
  using RT = RTile<T, RH, NT>;
  constexpr int K = RT::K;
  constexpr int KG = K >= 8 ? 4 : K;
  constexpr uint32_t F_ALL = F_FLUID | F_LFLUID | F_LEFT | F_RIGHT | F_DOWN | F_UP;
  const int R = a.R;
  const int TX = RT::RW - 2 * R, TY = RH - 2 * R;
  const int lane = threadIdx.x & 63, wave = wave_id();
  const int px = lane & 1;
  const int gx = bx * TX - R + lane, gy0 = by * TY - R;
  const bool colin = gx >= 0 && gx < a.W;
  const unsigned gxc = (unsigned)clampi(gx, a.W), gxo = gxc * (unsigned)sizeof(T);
#pragma unroll
  for (int k0 = 0; k0 < K; k0 += KG) {
    T rv[KG][2];
    uint32_t fv[KG][2];
#pragma unroll
    for (int j = 0; j < KG; ++j)
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int gy = gy0 + 2 * (wave + (k0 + j) * RT::NW) + q;
        const size_t row = (size_t)clampi(gy, a.H) * a.W;
        fv[j][q] = row_at(a.flags + row, gxc);
        rv[j][q] = row_at(a.rhs + row, gxo);
      }
#pragma unroll
    for (int j = 0; j < KG; ++j) {
      const int k = k0 + j;
      uint32_t f[2];
      T r[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int ly = 2 * (wave + k * RT::NW) + q;
        const int gy = gy0 + ly;
        const bool in = colin && gy >= 0 && gy < a.H;
        f[q] = in ? fv[j][q] : 0u;
        r[q] = in ? rv[j][q] : (T)0;
        sf[ly * RT::RW + lane] = (uint8_t)f[q];
      }
      t.rb[0][k] = px ? r[1] : r[0];
      t.rb[1][k] = px ? r[0] : r[1];
      const uint32_t f0 = px ? f[1] : f[0], f1 = px ? f[0] : f[1];
      if (__ballot((f0 & F_ALL) == F_ALL) == ~0ull) t.open[0] |= 1u << k;
      if (__ballot((f1 & F_ALL) == F_ALL) == ~0ull) t.open[1] |= 1u << k;
    }
  }
}

template <typename T, int RH, int NT>
__device__ __forceinline__ void smooth_tile64d_body(const SmoothArgs<T>& a, int bx, int by,
                                                    int part) {
  using RT = RTile<T, RH, NT>;
  using CP = CPatch<T, RH, NT>;
  constexpr int RW = RT::RW, NW = RT::NW;
  __shared__ T sp_[SPad<T, RH>::N];
  T* sp = sp_ + SPad<T, RH>::OFF;
  __shared__ uint8_t sf[RH * RW];
  __shared__ double red[NW];
  __shared__ T spatch[CP::N];
  if (a.gate != nullptr && *a.gate == 0) return;
  const int lane = threadIdx.x & 63, wave = wave_id();
  RT t;
  rtile_reset(t);
  if (a.zero_init) {
#pragma unroll
    for (int k = 0; k < RH / NW; ++k) sp[(wave + k * NW) * RW + lane] = (T)0;
  } else {
    rtile_dma_p<T, RH, NT>(a, bx, by, sp);
  }
  const bool prolong = a.coarse_e != nullptr;
  T cv[CP::PER];
  if (prolong) cpatch_issue<T, RH, NT>(a, bx, by, cv);
  rtile_stage_rf<T, RH, NT>(a, bx, by, sf, t);
  if (prolong) cpatch_store<T, RH, NT>(cv, spatch);
  t.open[0] = __builtin_amdgcn_readfirstlane(t.open[0]);
  t.open[1] = __builtin_amdgcn_readfirstlane(t.open[1]);
  __builtin_amdgcn_s_waitcnt(0);  // the LDS-DMA writes have landed
  __syncthreads();
  if (prolong) {
    rtile_prolong<T, RH, NT>(a, bx, by, sp, sf, spatch);
    __syncthreads();
  }
  rtile_passes(a, t, sp, sf);
  rtile_epilogue(a, t, bx, by, part, sp, sf, red);
}

// waves per SIMD the register allocation must allow: 4 when a float64 lane
// holds six or more row pairs (more would spill); 8 for the float32 64-row
// tiles (four 512-thread workgroups per CU); 6 for float64 lanes of four
// pairs (three 512-thread workgroups per CU, as the LDS of the 64-row tile
// allows); 5 otherwise
template <typename T, int RH, int NT>
constexpr int r_waves() {
  return sizeof(T) == 8 ? (RH / (2 * (NT / 64)) >= 6 ? 4 : (RH / (2 * (NT / 64)) == 4 ? 6 : 5))
                        : (RH == 64 ? 8 : 5);
}

template <typename T, int RH, int NT>
__global__ __launch_bounds__(NT, (r_waves<T, RH, NT>())) void k_smooth_fine_r(const SmoothArgs<T> a) {
  int bx, by;
  xcd_tile(bx, by);
  smooth_tile64r_body<T, RH, NT>(a, bx, by, by * gridDim.x + bx);
}

template <typename T, int RH, int NT>
__global__ __launch_bounds__(NT, (r_waves<T, RH, NT>())) void k_smooth_tile64_r(const SmoothArgs<T> a) {
  int bx, by;
  xcd_tile(bx, by);
  smooth_tile64r_body<T, RH, NT>(a, bx, by, by * gridDim.x + bx);
}

// the LDS-DMA variants: six waves per SIMD, four for a float64 lane with six
// or more row pairs
template <typename T, int RH, int NT>
constexpr int d_waves() {
  return (sizeof(T) == 8 && RH / (2 * (NT / 64)) >= 6) ? 4 : 6;
}

template <typename T, int RH, int NT>
__global__ __launch_bounds__(NT, (d_waves<T, RH, NT>())) void k_smooth_fine_d(const SmoothArgs<T> a) {
  int bx, by;
  xcd_tile(bx, by);
  smooth_tile64d_body<T, RH, NT>(a, bx, by, by * gridDim.x + bx);
}

template <typename T, int RH, int NT>
__global__ __launch_bounds__(NT, (d_waves<T, RH, NT>())) void k_smooth_tile64_d(const SmoothArgs<T> a) {
  int bx, by;
  xcd_tile(bx, by);
  smooth_tile64d_body<T, RH, NT>(a, bx, by, by * gridDim.x + bx);
}

// Red-black smoothing with the region in registers.  A workgroup of NW
// wavefronts stages a 64 x (NW*BR) region; lane x owns region column x, wave
// w owns rows [w*BR, (w+1)*BR).  Rows are held in pairs (2m, 2m+1) as
// A[m] = the colour-0 cell and B[m] = the colour-1 cell of the pair, so a
// colour pass updates exactly one register per pair and lane:
//   horizontal neighbours: the other colour of lanes x-1 / x+1 (DPP)
//   vertical neighbours:   the other colour in this lane's own pairs
//                          (pair m-1 / m+1), across wave bands via a 4 KB
//                          LDS exchange of the band-edge values.
// LDS holds only those edges, so occupancy is set by registers.  With R even
// the 2x2 restriction blocks coincide with (lane pair, row pair).

// Strip exchange for a wavefront split into 64 / COLS strips of COLS lanes:
// lane l receives v from lane l - COLS if `down`, else from lane l + COLS
// (lanes whose source lies outside the wave receive an unspecified value).
// gfx950's v_permlane16_swap / v_permlane32_swap: for (old, src),
// permlane16 returns rows [old0, src0, old2, src2] and [old1, src1, old3,
// src3] (rows of 16 lanes), permlane32 returns [old_lo, src_lo] and
// [old_hi, src_hi] (halves of 32 lanes).
template <int COLS>
__device__ __forceinline__ uint32_t strip_exchange_u32(uint32_t v, bool down) {
  const int lane = threadIdx.x & 63;
  if constexpr (COLS == 32) {
    const auto h = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return lane >= 32 ? h[0] : h[1];  // the other half holds the other strip
  } else {
    static_assert(COLS == 16, "strips of 16 or 32 lanes");
    // a = rows [v0 v0 v2 v2], b = [v1 v1 v3 v3]
    const auto ab = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    // c = [v0 v0 v1 v1], d = [v2 v2 v3 v3]
    const auto cd = __builtin_amdgcn_permlane32_swap(ab[0], ab[1], false, false);
    const int row = lane >> 4;
    // from below: row 1 <- v0 (a), row 2 <- v1 (c), row 3 <- v2 (a)
    // from above: row 0 <- v1 (b), row 1 <- v2 (d), row 2 <- v3 (b)
    return down ? (row == 2 ? cd[0] : ab[0]) : (row == 1 ? cd[1] : ab[1]);
  }
}
template <int COLS, typename T>
__device__ __forceinline__ T strip_exchange(T v, bool down) {
  uint32_t w[sizeof(T) / 4];
  __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) w[i] = strip_exchange_u32<COLS>(w[i], down);
  T r;
  __builtin_memcpy(&r, w, sizeof(T));
  return r;
}

template <typename T>
__device__ __forceinline__ T rb_relax(T ce, T l, T r, T d, T u, T b, uint32_t f, double alpha,
                                      double omega) {
  const T lv = (f & F_LEFT) ? l : ce;
  const T rv = (f & F_RIGHT) ? r : ce;
  const T dv = (f & F_DOWN) ? d : ce;
  const T uv = (f & F_UP) ? u : ce;
  const T ns = lv + rv + dv + uv;
  return (T)relax_value(0, ce, ns, b, alpha, omega);
}

template <typename T>
__device__ __forceinline__ double rb_resid(T ce, T l, T r, T d, T u, T b, uint32_t f,
                                           double alpha) {
  const T lv = (f & F_LEFT) ? l : ce;
  const T rv = (f & F_RIGHT) ? r : ce;
  const T dv = (f & F_DOWN) ? d : ce;
  const T uv = (f & F_UP) ? u : ce;
  const T ns = lv + rv + dv + uv;
  return residual_value(ce, ns, b, alpha);
}

template <typename T, int NW, int BR>
__global__ __launch_bounds__(NW * 64) void k_rb_reg(const SmoothArgs<T> a) {
  constexpr int NT = NW * 64;
  constexpr int RW = 64, RH = NW * BR, P = BR / 2;
  static_assert(BR % 2 == 0, "bands hold whole row pairs");
  __shared__ T edge[2][NW][2][64];
  __shared__ double red[NW];
  if (a.gate != nullptr && *a.gate == 0) return;

  const int R = a.R;
  const int TX = RW - 2 * R, TY = RH - 2 * R;
  const int H = a.H, W = a.W;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int px = lane & 1;
  const int gx0 = blockIdx.x * TX - R, gy0 = blockIdx.y * TY - R;
  const int gx = gx0 + lane;
  const bool colin = gx >= 0 && gx < W;
  const int gxc = gx < 0 ? 0 : (gx >= W ? W - 1 : gx);
  const int r0 = wave * BR;

  T A[P], B[P], RA[P], RB[P];
  uint32_t FA[P], FB[P];
  // ---- stage: unconditional clamped loads, in two halves of the band to
  // bound the registers live at once
#pragma unroll
  for (int hb = 0; hb < 2; ++hb) {
    constexpr int HR = BR / 2;
    T pv[HR], rv[HR];
    uint32_t fv[HR];
#pragma unroll
    for (int k = 0; k < HR; ++k) {
      const int gy = gy0 + r0 + hb * HR + k;
      const int gyc = gy < 0 ? 0 : (gy >= H ? H - 1 : gy);
      const size_t g = (size_t)gyc * W + gxc;
      fv[k] = a.flags[g];
      rv[k] = a.rhs[g];
      pv[k] = a.zero_init ? (T)0 : a.p_in[g];
    }
    if (a.coarse_e != nullptr) {
      T corr[HR];
#pragma unroll
      for (int k = 0; k < HR; ++k) {
        const int gy = gy0 + r0 + hb * HR + k;
        corr[k] = (T)sample_field(a.coarse_e, a.Hc, a.Wc, (gx + 0.5) * 0.5 - 0.5,
                                  (gy + 0.5) * 0.5 - 0.5);
      }
#pragma unroll
      for (int k = 0; k < HR; ++k) pv[k] = pv[k] + ((fv[k] & F_LFLUID) ? corr[k] : (T)0);
    }
#pragma unroll
    for (int k = 0; k < HR; ++k) {
      const int gy = gy0 + r0 + hb * HR + k;
      const bool in = colin && gy >= 0 && gy < H;
      if (!in) {
        pv[k] = (T)0;
        rv[k] = (T)0;
        fv[k] = 0;
      }
    }
#pragma unroll
    for (int q = 0; q < HR / 2; ++q) {
      const int m = hb * (HR / 2) + q;
      A[m] = px ? pv[2 * q + 1] : pv[2 * q];
      B[m] = px ? pv[2 * q] : pv[2 * q + 1];
      RA[m] = px ? rv[2 * q + 1] : rv[2 * q];
      RB[m] = px ? rv[2 * q] : rv[2 * q + 1];
      FA[m] = px ? fv[2 * q + 1] : fv[2 * q];
      FB[m] = px ? fv[2 * q] : fv[2 * q + 1];
    }
  }

  // ---- colour passes
  for (int pass = 0; pass < a.npass; ++pass) {
    const int colour = (a.first_colour + pass) & 1;
    const int lo = pass + 1;
    const bool xin = lane >= lo && lane < RW - lo;
    const int slot = pass & 1;
    // publish this band's edge values of the colour read in this pass
    edge[slot][wave][0][lane] = colour == 0 ? B[0] : A[0];
    edge[slot][wave][1][lane] = colour == 0 ? B[P - 1] : A[P - 1];
    __syncthreads();
    const T below = wave > 0 ? edge[slot][wave - 1][1][lane] : (T)0;
    const T above = wave < NW - 1 ? edge[slot][wave + 1][0][lane] : (T)0;
    if (colour == 0) {
#pragma unroll
      for (int m = 0; m < P; ++m) {
        const T bd = m > 0 ? B[m - 1] : below;
        const T bu = m < P - 1 ? B[m + 1] : above;
        const T l = from_left(B[m]), r = from_right(B[m]);
        const T d = px ? B[m] : bd;
        const T u = px ? bu : B[m];
        const int t = r0 + 2 * m + px;
        const bool ok = xin && t >= lo && t < RH - lo && (FA[m] & F_FLUID);
        const T nv = rb_relax(A[m], l, r, d, u, RA[m], FA[m], a.alpha, a.omega);
        A[m] = ok ? nv : A[m];
      }
    } else {
#pragma unroll
      for (int m = 0; m < P; ++m) {
        const T ad = m > 0 ? A[m - 1] : below;
        const T au = m < P - 1 ? A[m + 1] : above;
        const T l = from_left(A[m]), r = from_right(A[m]);
        const T d = px ? ad : A[m];
        const T u = px ? A[m] : au;
        const int t = r0 + 2 * m + 1 - px;
        const bool ok = xin && t >= lo && t < RH - lo && (FB[m] & F_FLUID);
        const T nv = rb_relax(B[m], l, r, d, u, RB[m], FB[m], a.alpha, a.omega);
        B[m] = ok ? nv : B[m];
      }
    }
  }

  // ---- write the tile interior
  const bool xout = lane >= R && lane < R + TX && colin;
#pragma unroll
  for (int m = 0; m < P; ++m) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int ly = r0 + 2 * m + k;
      const int gy = gy0 + ly;
      const T v = (k == px) ? A[m] : B[m];
      if (xout && ly >= R && ly < R + TY && gy < H) a.p_out[(size_t)gy * W + gx] = v;
    }
  }

  if (a.coarse_rhs == nullptr && a.norm_partial == nullptr) return;

  // ---- residuals of the final iterate (both colours)
  const int slot = a.npass & 1;
  edge[slot][wave][0][lane] = A[0];
  edge[slot][wave][1][lane] = A[P - 1];
  __syncthreads();
  const T a_below = wave > 0 ? edge[slot][wave - 1][1][lane] : (T)0;
  const T a_above = wave < NW - 1 ? edge[slot][wave + 1][0][lane] : (T)0;
  __syncthreads();
  edge[slot][wave][0][lane] = B[0];
  edge[slot][wave][1][lane] = B[P - 1];
  __syncthreads();
  const T b_below = wave > 0 ? edge[slot][wave - 1][1][lane] : (T)0;
  const T b_above = wave < NW - 1 ? edge[slot][wave + 1][0][lane] : (T)0;
  const int Wc = W / 2, Hc = H / 2;
  double acc = 0.0;
#pragma unroll
  for (int m = 0; m < P; ++m) {
    const T bd = m > 0 ? B[m - 1] : b_below;
    const T bu = m < P - 1 ? B[m + 1] : b_above;
    const T ad = m > 0 ? A[m - 1] : a_below;
    const T au = m < P - 1 ? A[m + 1] : a_above;
    // residuals in float64; the restriction takes them in the level's dtype
    // (as stored by the reference), the norm in float64
    const double dA = rb_resid(A[m], from_left(B[m]), from_right(B[m]), px ? B[m] : bd,
                               px ? bu : B[m], RA[m], FA[m], a.alpha);
    const double dB = rb_resid(B[m], from_left(A[m]), from_right(A[m]), px ? ad : A[m],
                               px ? A[m] : au, RB[m], FB[m], a.alpha);
    const double dlo = px ? dB : dA;  // row 2m
    const double dhi = px ? dA : dB;  // row 2m+1
    const T rlo = (T)dlo, rhi = (T)dhi;
    const uint32_t flo = px ? FB[m] : FA[m];
    const uint32_t fhi = px ? FA[m] : FB[m];
    const int ly = r0 + 2 * m;
    const int gy = gy0 + ly;
    const bool yin = ly >= R && ly < R + TY;
    if (a.coarse_rhs != nullptr) {
      // 2x2 block = (lanes x, x+1) x (rows 2m, 2m+1), x even (R even); the
      // children are summed in the reference's (dj, di) order
      const T r01 = from_right(rlo), r11 = from_right(rhi);
      const uint32_t f01 = (uint32_t)from_right((int)flo), f11 = (uint32_t)from_right((int)fhi);
      double total = 0.0;
      int count = 0;
      total += (flo & F_LFLUID) ? (double)rlo : 0.0;
      count += (flo & F_LFLUID) ? 1 : 0;
      total += (f01 & F_LFLUID) ? (double)r01 : 0.0;
      count += (f01 & F_LFLUID) ? 1 : 0;
      total += (fhi & F_LFLUID) ? (double)rhi : 0.0;
      count += (fhi & F_LFLUID) ? 1 : 0;
      total += (f11 & F_LFLUID) ? (double)r11 : 0.0;
      count += (f11 & F_LFLUID) ? 1 : 0;
      const int Jc = gy >> 1, Ic = gx >> 1;
      if (px == 0 && xout && yin && Jc < Hc && Ic < Wc) {
        const T mean = count > 0 ? (T)(total / (double)count) : (T)0;
        a.coarse_rhs[(size_t)Jc * Wc + Ic] = (T)((double)mean / a.coarse_div);
      }
    }
    if (a.norm_partial != nullptr) {
      // each row of the pair against the output window on its own: with an
      // odd halo a pair straddles the window's edge
      const bool yin1 = ly + 1 >= R && ly + 1 < R + TY;
      const bool in0 = xout && yin && gy < H && (flo & F_LFLUID);
      const bool in1 = xout && yin1 && gy + 1 < H && (fhi & F_LFLUID);
      acc += in0 ? dlo * dlo : 0.0;
      acc += in1 ? dhi * dhi : 0.0;
    }
  }
  if (a.norm_partial != nullptr) {
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
  const float* cop;    // red-black coarsest level by its precomputed operator (or null)
  const int* gate;     // nothing is done while *gate == 0
};

// One red-black colour pass or Jacobi iteration over a whole level in LDS.
// Levels up to 64 columns map lane -> column (row pairs for red-black, as in
// k_smooth_tile64); new values are computed in batches of four before they
// are stored.  Wider levels use a flat index.
template <typename T, int MODE, int NT>
__device__ __forceinline__ void small_pass(const T* cur, T* dst, const T* rhs,
                           const uint8_t* fl, int H, int W, double alpha, double omega,
                           int colour) {
  constexpr int NW = NT / 64;
  constexpr int B = 4;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int N = H * W;
  // loads at a clamped index (always inside the level), selects by flags;
  // returns whether the cell is written in this pass
  auto update = [&](int o, T& out) -> bool {
    const uint8_t f = fl[o];
    const T ce = cur[o];
    const T lv = cur[o > 0 ? o - 1 : 0];
    const T rv = cur[o + 1 < N ? o + 1 : N - 1];
    const T dv = cur[o >= W ? o - W : 0];
    const T uv = cur[o + W < N ? o + W : N - 1];
    const T b = rhs[o];
    const T l = (f & F_LEFT) ? lv : ce;
    const T r = (f & F_RIGHT) ? rv : ce;
    const T d = (f & F_DOWN) ? dv : ce;
    const T u = (f & F_UP) ? uv : ce;
    const T ns = l + r + d + u;
    const bool fluid = f & F_FLUID;
    out = fluid ? (T)relax_value(MODE == MODE_RB ? 0 : 1, ce, ns, b, alpha, omega) : ce;
    return fluid || MODE == MODE_JACOBI;  // Jacobi copies non-fluid cells
  };
  if (W <= 64) {
    const bool lin = lane < W;
    if (MODE == MODE_RB) {
      const int par = (lane + colour) & 1;
      const int npair = (H + 1) >> 1;
      for (int m0 = wave; m0 < npair; m0 += B * NW) {
        T nv[B];
        int oo[B];
        bool w_[B];
#pragma unroll
        for (int b = 0; b < B; ++b) {
          const int j = 2 * (m0 + b * NW) + par;
          const bool ok = lin && (m0 + b * NW) < npair && j < H;
          oo[b] = ok ? j * W + lane : 0;
          w_[b] = update(oo[b], nv[b]) && ok;
        }
#pragma unroll
        for (int b = 0; b < B; ++b)
          if (w_[b]) dst[oo[b]] = nv[b];
      }
    } else {
      for (int j0 = wave; j0 < H; j0 += B * NW) {
        T nv[B];
        int oo[B];
        bool w_[B];
#pragma unroll
        for (int b = 0; b < B; ++b) {
          const int j = j0 + b * NW;
          const bool ok = lin && j < H;
          oo[b] = ok ? j * W + lane : 0;
          w_[b] = update(oo[b], nv[b]) && ok;
        }
#pragma unroll
        for (int b = 0; b < B; ++b)
          if (w_[b]) dst[oo[b]] = nv[b];
      }
    }
    return;
  }
  if (MODE == MODE_RB) {
    const int nh = (W + 1) >> 1;
    for (int k = threadIdx.x; k < H * nh; k += NT) {
      const int j = k / nh, c = k - j * nh;
      const int i = ((colour + j) & 1) + 2 * c;
      if (i >= W) continue;
      T v;
      if (update(j * W + i, v)) dst[j * W + i] = v;
    }
  } else {
    for (int o = threadIdx.x; o < H * W; o += NT) {
      T v;
      if (update(o, v)) dst[o] = v;
    }
  }
}

// Red-black passes over a level of at most 64 x 64 cells with the cells in
// registers (the layout of k_rb_reg): lane x owns column x, wave w owns the
// row pairs [w*P, (w+1)*P); A[m] / B[m] are the colour-0 / colour-1 cell of
// pair m.  Horizontal neighbours come from lanes x -/+ 1 (__shfl), vertical
// ones from the lane's own pairs, and across band edges from the level's LDS
// array: a pass first stores the cells the previous pass updated (the colour
// this pass reads), so after one barrier the band-edge rows are current.  A
// pass reads one colour's cells and the next writes the other colour, so a
// single barrier per pass separates every store from the reads it could
// disturb.  Cells outside the level are never updated and never used: the
// flags of the level's border cells have those neighbour bits cleared.
// The per-cell arithmetic is small_pass's, in the same order.
template <typename T, int NT, int PMAX>
__device__ __forceinline__ void small_smooth_rb_reg(T* p, const T* rhs, const uint8_t* fl, int H,
                                                    int W, double alpha, double omega,
                                                    int npass) {
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int px = lane & 1;
  constexpr int P = PMAX;        // ceil(ceil(H / 2) / NW): the caller dispatches on it
  const int r0 = 2 * wave * P;  // first row of this wave's band
  const bool lin = lane < W;
  T A[PMAX], B[PMAX], RA[PMAX], RB[PMAX];
  uint32_t FA[PMAX], FB[PMAX];
#pragma unroll
  for (int m = 0; m < PMAX; ++m) {
    const int ya = r0 + 2 * m + px, yb = r0 + 2 * m + 1 - px;
    const bool oka = lin && m < P && ya < H, okb = lin && m < P && yb < H;
    const int oa = oka ? ya * W + lane : 0, ob = okb ? yb * W + lane : 0;
    A[m] = oka ? p[oa] : (T)0;
    B[m] = okb ? p[ob] : (T)0;
    RA[m] = oka ? rhs[oa] : (T)0;
    RB[m] = okb ? rhs[ob] : (T)0;
    FA[m] = oka ? fl[oa] : 0u;
    FB[m] = okb ? fl[ob] : 0u;
  }
  const int yb_lo = r0 - 1, yb_hi = r0 + 2 * P;
  const bool has_lo = lin && yb_lo >= 0 && yb_lo < H, has_hi = lin && yb_hi < H;
  for (int pass = 0; pass < npass; ++pass) {
    const int colour = pass & 1;
    if (pass > 0) {  // publish the cells the previous pass updated
#pragma unroll
      for (int m = 0; m < PMAX; ++m) {
        const int yo = colour == 0 ? r0 + 2 * m + 1 - px : r0 + 2 * m + px;
        if (lin && m < P && yo < H) p[yo * W + lane] = colour == 0 ? B[m] : A[m];
      }
    }
    __syncthreads();
    const T below = has_lo ? p[yb_lo * W + lane] : (T)0;
    const T above = has_hi ? p[yb_hi * W + lane] : (T)0;
    if (colour == 0) {
#pragma unroll
      for (int m = 0; m < PMAX; ++m) {
        const T bd = m > 0 ? B[m - 1] : below;
        const T bu = m + 1 < P ? B[m + 1 < PMAX ? m + 1 : m] : above;
        const T l = from_left(B[m]), r = from_right(B[m]);
        const T d = px ? B[m] : bd;
        const T u = px ? bu : B[m];
        const bool ok = (FA[m] & F_FLUID) != 0u;
        const T nv = rb_relax(A[m], l, r, d, u, RA[m], FA[m], alpha, omega);
        A[m] = ok ? nv : A[m];
      }
    } else {
#pragma unroll
      for (int m = 0; m < PMAX; ++m) {
        const T ad = m > 0 ? A[m - 1] : below;
        const T au = m + 1 < P ? A[m + 1 < PMAX ? m + 1 : m] : above;
        const T l = from_left(A[m]), r = from_right(A[m]);
        const T d = px ? ad : A[m];
        const T u = px ? A[m] : au;
        const bool ok = (FB[m] & F_FLUID) != 0u;
        const T nv = rb_relax(B[m], l, r, d, u, RB[m], FB[m], alpha, omega);
        B[m] = ok ? nv : B[m];
      }
    }
  }
  // store the colour the last pass updated; the other one is already there
  if (npass > 0) {
    const int last = (npass - 1) & 1;
#pragma unroll
    for (int m = 0; m < PMAX; ++m) {
      const int yo = last == 0 ? r0 + 2 * m + px : r0 + 2 * m + 1 - px;
      if (lin && m < P && yo < H) p[yo * W + lane] = last == 0 ? A[m] : B[m];
    }
  }
  __syncthreads();
}

constexpr int WAVE_COARSEST = 1, WAVE_LEVELS = 2;

// Red-black passes over a level of at most COLS x (128 / COLS * P) cells on
// one wavefront with the cells in registers.  Lane = x + COLS * s owns
// column x of strip s, rows [2Ps, 2P(s+1)), held as pairs m (rows 2Ps + 2m,
// 2Ps + 2m + 1): A[m] / B[m] is the pair's colour-0 / colour-1 cell, at row
// 2Ps + 2m + px / 2Ps + 2m + 1 - px with px = x & 1.  A colour pass updates
// one register per pair.  Horizontal neighbours are the other colour of
// lanes l -/+ 1 (DPP); vertical ones the other colour in the lane's own
// pairs and, across a strip edge, of the lane COLS below or above
// (strip_exchange), so a pass touches neither LDS nor a barrier.  Neighbour
// values from outside the level are never used: the flags of the level's
// border cells have those neighbour bits cleared.  The per-cell arithmetic is
// rb_relax's.  All 64 lanes of the wave must call it.
template <typename T, int COLS, int P>
__device__ __forceinline__ void wave_smooth_rb(T* p, const T* rhs, const uint8_t* fl, int H, int W,
                                               double alpha, double omega, int npass) {
  const int lane = threadIdx.x & 63;
  const int x = lane % COLS, s = lane / COLS;
  const int px = x & 1;
  const int r0 = 2 * P * s;
  const bool lin = x < W;
  T A[P], B[P], RA[P], RB[P];
  uint32_t FA[P], FB[P];
#pragma unroll
  for (int m = 0; m < P; ++m) {
    const int ya = r0 + 2 * m + px, yb = r0 + 2 * m + 1 - px;
    const bool oka = lin && ya < H, okb = lin && yb < H;
    const int oa = oka ? ya * W + x : 0, ob = okb ? yb * W + x : 0;
    A[m] = oka ? p[oa] : (T)0;
    B[m] = okb ? p[ob] : (T)0;
    RA[m] = oka ? rhs[oa] : (T)0;
    RB[m] = okb ? rhs[ob] : (T)0;
    FA[m] = oka ? fl[oa] : 0u;
    FB[m] = okb ? fl[ob] : 0u;
  }
  for (int pass = 0; pass < npass; ++pass) {
    if ((pass & 1) == 0) {
      // px = 0: A[0]'s lower neighbour is B[P-1] of the strip below;
      // px = 1: A[P-1]'s upper neighbour is B[0] of the strip above
      const T e = strip_exchange<COLS>(px ? B[0] : B[P - 1], px == 0);
#pragma unroll
      for (int m = 0; m < P; ++m) {
        const T d = px ? B[m] : (m > 0 ? B[m > 0 ? m - 1 : 0] : e);
        const T u = px ? (m + 1 < P ? B[m + 1 < P ? m + 1 : m] : e) : B[m];
        const T nv = rb_relax(A[m], from_left(B[m]), from_right(B[m]), d, u, RA[m], FA[m], alpha,
                              omega);
        A[m] = (FA[m] & F_FLUID) ? nv : A[m];
      }
    } else {
      // px = 1: B[0]'s lower neighbour is A[P-1] of the strip below;
      // px = 0: B[P-1]'s upper neighbour is A[0] of the strip above
      const T e = strip_exchange<COLS>(px ? A[P - 1] : A[0], px == 1);
#pragma unroll
      for (int m = 0; m < P; ++m) {
        const T d = px ? (m > 0 ? A[m > 0 ? m - 1 : 0] : e) : A[m];
        const T u = px ? A[m] : (m + 1 < P ? A[m + 1 < P ? m + 1 : m] : e);
        const T nv = rb_relax(B[m], from_left(A[m]), from_right(A[m]), d, u, RB[m], FB[m], alpha,
                              omega);
        B[m] = (FB[m] & F_FLUID) ? nv : B[m];
      }
    }
  }
#pragma unroll
  for (int m = 0; m < P; ++m) {
    const int ya = r0 + 2 * m + px, yb = r0 + 2 * m + 1 - px;
    if (lin && ya < H) p[ya * W + x] = A[m];
    if (lin && yb < H) p[yb * W + x] = B[m];
  }
}

// Red-black passes over a level of at most 32 x 32 cells on the calling
// wavefront (all 64 lanes): strips of 16 lanes for levels up to 16 x 16,
// of 32 lanes up to 32 x 32, with the pairs per lane as a compile-time count.
template <typename T>
__device__ __forceinline__ void wave_smooth_level(T* p, const T* rhs, const uint8_t* fl, int H,
                                                  int W, double alpha, double omega, int npass) {
  if (W <= 16 && H <= 16) {
    if (H <= 8)
      wave_smooth_rb<T, 16, 1>(p, rhs, fl, H, W, alpha, omega, npass);
    else
      wave_smooth_rb<T, 16, 2>(p, rhs, fl, H, W, alpha, omega, npass);
    return;
  }
  switch ((H + 3) / 4) {  // 2 strips of 2P rows
    case 1: wave_smooth_rb<T, 32, 1>(p, rhs, fl, H, W, alpha, omega, npass); return;
    case 2: wave_smooth_rb<T, 32, 2>(p, rhs, fl, H, W, alpha, omega, npass); return;
    case 3: wave_smooth_rb<T, 32, 3>(p, rhs, fl, H, W, alpha, omega, npass); return;
    case 4: wave_smooth_rb<T, 32, 4>(p, rhs, fl, H, W, alpha, omega, npass); return;
    case 5: wave_smooth_rb<T, 32, 5>(p, rhs, fl, H, W, alpha, omega, npass); return;
    case 6: wave_smooth_rb<T, 32, 6>(p, rhs, fl, H, W, alpha, omega, npass); return;
    case 7: wave_smooth_rb<T, 32, 7>(p, rhs, fl, H, W, alpha, omega, npass); return;
    default: wave_smooth_rb<T, 32, 8>(p, rhs, fl, H, W, alpha, omega, npass); return;
  }
}

template <typename T, int MODE, int NT, int WAVE>
__device__ __forceinline__ void small_smooth(T* p, T* tmp, const T* rhs, const uint8_t* fl, int H, int W,
                             double alpha, double omega, int npass) {
  if (MODE == MODE_RB && (WAVE & WAVE_LEVELS) && W <= 32 && H <= 32) {
    if (threadIdx.x < 64) wave_smooth_level<T>(p, rhs, fl, H, W, alpha, omega, npass);
    __syncthreads();
    return;
  }
  if (MODE == MODE_RB && W <= 64 && H <= 64) {
    // pairs per wave as a compile-time count, so no pass runs predicated-off
    // work for pairs the level does not have
    constexpr int NW = NT / 64;
    const int P = (((H + 1) >> 1) + NW - 1) / NW;
    switch (P) {
      case 1: small_smooth_rb_reg<T, NT, 1>(p, rhs, fl, H, W, alpha, omega, npass); return;
      case 2: small_smooth_rb_reg<T, NT, 2>(p, rhs, fl, H, W, alpha, omega, npass); return;
      case 3: small_smooth_rb_reg<T, NT, 3>(p, rhs, fl, H, W, alpha, omega, npass); return;
      case 4: small_smooth_rb_reg<T, NT, 4>(p, rhs, fl, H, W, alpha, omega, npass); return;
      case 5: small_smooth_rb_reg<T, NT, 5>(p, rhs, fl, H, W, alpha, omega, npass); return;
      case 6: small_smooth_rb_reg<T, NT, 6>(p, rhs, fl, H, W, alpha, omega, npass); return;
      case 7: small_smooth_rb_reg<T, NT, 7>(p, rhs, fl, H, W, alpha, omega, npass); return;
      case 8: small_smooth_rb_reg<T, NT, 8>(p, rhs, fl, H, W, alpha, omega, npass); return;
      default: break;  // more than 8 pairs per wave: the LDS passes below
    }
  }
  const int N = H * W;
  T* cur = p;
  for (int pass = 0; pass < npass; ++pass) {
    if (MODE == MODE_RB) {
      small_pass<T, MODE, NT>(cur, cur, rhs, fl, H, W, alpha, omega, pass & 1);
    } else {
      T* nxt = (cur == p) ? tmp : p;
      small_pass<T, MODE, NT>(cur, nxt, rhs, fl, H, W, alpha, omega, 0);
      cur = nxt;
    }
    __syncthreads();
  }
  if (cur != p) {  // Jacobi with an odd count: copy back (jacobi.py:51-52)
    for (int o = threadIdx.x; o < N; o += NT) p[o] = cur[o];
    __syncthreads();
  }
}

// Coarsest level of at most 16 x 16 cells on one wavefront with the cells in
// registers: lane (bx, by) = (lane & 7, lane >> 3) owns cells (2bx+a, 2by+b).
// A pass reads only values not written in that pass (the other colour, or the
// previous Jacobi iterate), taken from the lane's own registers or from the
// neighbouring lane (DPP across columns, __shfl across rows), so no pass
// touches LDS memory or a barrier.  A pass is one wave's chain of dependent
// instructions, so everything constant across passes is taken out of it.
// Neighbour values from lanes outside the level are never used: the cell
// flags of border cells have those neighbour bits cleared.
template <typename T, int MODE>
__device__ __forceinline__ void coarsest_wave(T* p, const T* rhs, const uint8_t* fl, int H, int W, double alpha,
                              double omega, int npass) {
  const int lane = threadIdx.x & 63;
  const int bx = lane & 7, by = lane >> 3;
  T c[2][2];
  double ab[2][2];  // alpha * rhs, the same product every pass
  // the cell's flags as lane masks, tested once instead of every pass
  bool fluid[2][2], oL[2][2], oR[2][2], oD[2][2], oU[2][2];
#pragma unroll
  for (int b = 0; b < 2; ++b)
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const int i = 2 * bx + a, j = 2 * by + b;
      const bool in = i < W && j < H;
      const uint8_t f = in ? fl[j * W + i] : 0;
      c[b][a] = in ? p[j * W + i] : (T)0;
      ab[b][a] = alpha * (double)(in ? rhs[j * W + i] : (T)0);
      fluid[b][a] = (f & F_FLUID) != 0;
      oL[b][a] = (f & F_LEFT) != 0;
      oR[b][a] = (f & F_RIGHT) != 0;
      oD[b][a] = (f & F_DOWN) != 0;
      oU[b][a] = (f & F_UP) != 0;
    }
  // vertical neighbours are 8 lanes away (a row of 2x2 blocks is 8 lanes):
  // through the LDS crossbar; horizontal ones are the adjacent lanes: DPP
  // (lane 0 / 63 keep their own value where __shfl would wrap; both are at
  // the level's edge, whose faces are closed, so the value is never used)
  const int dm = (lane - 8) & 63, dp = (lane + 8) & 63;
  // relax_value / poisson_new of the cell, in the same order: the neighbour
  // sum in T, (sum + alpha * rhs) * 0.25 and the SOR blend in float64
  auto relax = [&](int a, int b, T left, T right, T down, T up) -> T {
    const T ce = c[b][a];
    const T l = oL[b][a] ? left : ce;
    const T rr = oR[b][a] ? right : ce;
    const T d = oD[b][a] ? down : ce;
    const T u = oU[b][a] ? up : ce;
    const T sum = l + rr + d + u;
    const double nw = ((double)sum + ab[b][a]) * 0.25;
    const T v = MODE == MODE_RB ? (T)((1.0 - omega) * (double)ce + omega * nw) : (T)nw;
    return fluid[b][a] ? v : ce;
  };
  // Each pass issues its LDS-crossbar reads before anything waits on one, so
  // a pass costs one crossbar round trip.
  auto colour0 = [&]() {  // cells (0,0) and (1,1)
    const T x10 = __shfl(c[1][0], dm), y01 = __shfl(c[0][1], dp);
    __builtin_amdgcn_sched_barrier(0);
    const T x01 = from_left(c[0][1]), y10 = from_right(c[1][0]);
    const T n00 = relax(0, 0, x01, c[0][1], x10, c[1][0]);
    const T n11 = relax(1, 1, c[1][0], y10, c[0][1], y01);
    c[0][0] = n00;
    c[1][1] = n11;
  };
  auto colour1 = [&]() {  // cells (1,0) and (0,1)
    const T x11 = __shfl(c[1][1], dm), y00 = __shfl(c[0][0], dp);
    __builtin_amdgcn_sched_barrier(0);
    const T x00 = from_right(c[0][0]), y11 = from_left(c[1][1]);
    const T n10 = relax(1, 0, c[0][0], x00, x11, c[1][1]);
    const T n01 = relax(0, 1, y11, c[1][1], c[0][0], y00);
    c[0][1] = n10;
    c[1][0] = n01;
  };
  if (MODE == MODE_RB) {
    int pass = 0;
    for (; pass + 2 <= npass; pass += 2) {
      colour0();
      colour1();
    }
    if (pass < npass) colour0();
  } else {
    for (int pass = 0; pass < npass; ++pass) {
      const T d0 = __shfl(c[1][0], dm), d1 = __shfl(c[1][1], dm);
      const T u0 = __shfl(c[0][0], dp), u1 = __shfl(c[0][1], dp);
      __builtin_amdgcn_sched_barrier(0);
      const T l0 = from_left(c[0][1]), l1 = from_left(c[1][1]);
      const T r0 = from_right(c[0][0]), r1 = from_right(c[1][0]);
      const T n00 = relax(0, 0, l0, c[0][1], d0, c[1][0]);
      const T n10 = relax(1, 0, c[0][0], r0, d1, c[1][1]);
      const T n01 = relax(0, 1, l1, c[1][1], c[0][0], u0);
      const T n11 = relax(1, 1, c[1][0], r1, c[0][1], u1);
      c[0][0] = n00;
      c[0][1] = n10;
      c[1][0] = n01;
      c[1][1] = n11;
    }
  }
#pragma unroll
  for (int b = 0; b < 2; ++b)
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const int i = 2 * bx + a, j = 2 * by + b;
      if (i < W && j < H) p[j * W + i] = c[b][a];
    }
}

// x = op b for a coarsest level of n <= 256 cells: op[j * n + i] is the
// result in cell i of the coarsest red-black passes from a zero guess with
// rhs = unit vector j (k_coarse_op).  The passes are linear in the rhs, so
// this is their result for b up to rounding (k_coarse_op; or the exact
// solution operator, coarse_exact_operator).  Thread t covers rows
// 4 (t % 64) .. + 3 of a column set: every load of a set's matrix columns is
// a float4 issued before any is used (one memory latency per set, where a
// column-by-column loop waits once per column); the sets' partial sums meet
// in LDS.
template <typename T, int NT>
__device__ __forceinline__ void coarsest_apply_op(const float* __restrict__ op, const T* b, T* x,
                                                  int n) {
  // 16 column sets j = g (mod 16), one float partial per (set, row) summed
  // over the set's columns in order, then the 16 partials of a row in double
  // in set order: the order of 1024 threads, one set per wavefront.  With
  // fewer threads a wavefront takes several sets in turn, so the result does
  // not depend on NT.
  constexpr int G = 16;
  constexpr int JMAX = 256 / G;  // columns per set at most
  static_assert(NT % 64 == 0 && NT <= 64 * G, "whole wavefronts, at most one set each");
  __shared__ float part[G][256];
  const int q = threadIdx.x % 64;
  const int i0 = 4 * q;
  for (int g = threadIdx.x / 64; g < G; g += NT / 64) {
    if ((n & 3) == 0) {
      float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
      float4 m[JMAX];
      float bj[JMAX];
#pragma unroll
      for (int k = 0; k < JMAX; ++k) {
        const int j = g + k * G;
        const bool ok = j < n && i0 < n;
        m[k] = ok ? *reinterpret_cast<const float4*>(op + (size_t)j * n + i0)
                  : make_float4(0.f, 0.f, 0.f, 0.f);
        bj[k] = j < n ? (float)b[j] : 0.f;
      }
#pragma unroll
      for (int k = 0; k < JMAX; ++k) {
        acc.x += m[k].x * bj[k];
        acc.y += m[k].y * bj[k];
        acc.z += m[k].z * bj[k];
        acc.w += m[k].w * bj[k];
      }
      part[g][i0] = acc.x;
      part[g][i0 + 1] = acc.y;
      part[g][i0 + 2] = acc.z;
      part[g][i0 + 3] = acc.w;
    } else {  // rows one at a time
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + r;
        float acc = 0.f;
        for (int j = g; j < n; j += G)
          if (i < n) acc += op[(size_t)j * n + i] * (float)b[j];
        part[g][i] = acc;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < n) {
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < G; ++k) acc += (double)part[k][threadIdx.x];
    x[threadIdx.x] = (T)acc;
  }
  __syncthreads();
}

// x = op (alpha b) for a level of n <= COARSE_EXACT_MAX cells (op[j * n + i],
// the level's exact float32 operator K^+ without its alpha, so that one build
// serves every dt; alpha b is formed in float32) over many workgroups:
// workgroup g owns rows
// EX_ROWS g .. + EX_ROWS - 1, and its thread t covers 4 of them
// (4 (t % 4) .. + 3, one float4 per column) over the column set
// s = t / 4 (columns j = s, s + EX_SETS, ...).  Every load of a thread is
// issued before any is used; each (set, row) partial is summed in float over
// the set's columns in order, then a row's EX_SETS partials in double in set
// order -- fixed by the code, whatever the launch.
constexpr int EX_SETS = 64, EX_ROWS = 16, EX_NT = 256;
constexpr int EX_JMAX = COARSE_EXACT_MAX / EX_SETS;  // columns per set at most
static_assert(EX_NT == EX_SETS * EX_ROWS / 4, "one thread per (set, 4 rows)");

__global__ __launch_bounds__(EX_NT) void k_apply_exact(const float* __restrict__ op,
                                                       const float* __restrict__ b,
                                                       float* __restrict__ x, int n,
                                                       float alpha,
                                                       const int* __restrict__ gate) {
  if (gate != nullptr && *gate == 0) return;
  __shared__ float part[EX_SETS][EX_ROWS];
  const int t = threadIdx.x;
  const int r4 = 4 * (t % 4), set = t / 4;
  const int row0 = blockIdx.x * EX_ROWS;
  const int i0 = row0 + r4;
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  if ((n & 3) == 0) {  // rows i0 .. i0 + 3 all inside or all outside; 16-byte aligned
    // every load at an index clamped into the arrays, all issued before any
    // is used; columns past n add nothing, rows past n are not stored
    const int ic = i0 < n ? i0 : 0;
    float4 m[EX_JMAX];
    float bv[EX_JMAX];
#pragma unroll
    for (int k = 0; k < EX_JMAX; ++k) {
      const int j = set + k * EX_SETS;
      const int jc = j < n ? j : n - 1;
      m[k] = *reinterpret_cast<const float4*>(op + (size_t)jc * n + ic);
      bv[k] = b[jc];
    }
#pragma unroll
    for (int k = 0; k < EX_JMAX; ++k) {
      if (set + k * EX_SETS >= n) continue;
      const float bj = alpha * bv[k];
      acc[0] += m[k].x * bj;
      acc[1] += m[k].y * bj;
      acc[2] += m[k].z * bj;
      acc[3] += m[k].w * bj;
    }
  } else {  // row by row
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = i0 + r;
      for (int j = set; j < n; j += EX_SETS)
        if (i < n) acc[r] += op[(size_t)j * n + i] * (alpha * b[j]);
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) part[set][r4 + r] = acc[r];
  __syncthreads();
  if (t < EX_ROWS && row0 + t < n) {
    double sum = 0.0;
#pragma unroll 8
    for (int k = 0; k < EX_SETS; ++k) sum += (double)part[k][t];
    x[row0 + t] = (float)sum;
  }
}

// WAVE (bits): WAVE_COARSEST runs a red-black coarsest level of at most
// 16 x 16 cells with wave_smooth_level instead of coarsest_wave;
// WAVE_LEVELS also runs every red-black level of at most 32 x 32 cells, the
// coarsest included, on wavefront 0 while the others wait at the barrier.
template <typename T, int MODE, int NT, int WAVE>
__device__ __forceinline__ void vcycle_small_body(const SmallArgs<T>& a) {
  extern __shared__ __align__(16) unsigned char smem[];
  if (a.gate != nullptr && *a.gate == 0) return;
  // Every level array is addressed as (LDS base + per-level offset): pointers
  // picked from an array at run time would lose the LDS address space and
  // turn every access into a FLAT one.  Layout: [p_l, b_l]* tmp | flags_l*.
  T* const sT = reinterpret_cast<T*>(smem);
  int offP[SMALL_MAX], offB[SMALL_MAX], offF[SMALL_MAX];
  int toff = 0, foff = 0, nmax = 0;
  constexpr int PADT = 16 / (int)sizeof(T);
  for (int l = 0; l < a.nl; ++l) {
    const int n = a.H[l] * a.W[l];
    const int np = (n + PADT - 1) / PADT * PADT;
    nmax = n > nmax ? n : nmax;
    offP[l] = toff;
    toff += np;
    offB[l] = toff;
    toff += np;
    offF[l] = foff;
    foff += (n + 15) & ~15;
  }
  const int offTmp = toff;
  if (MODE == MODE_JACOBI) toff += (nmax + PADT - 1) / PADT * PADT;
  uint8_t* const fb = smem + (size_t)toff * sizeof(T);
#define SP(l) (sT + offP[(l)])
#define SB(l) (sT + offB[(l)])
#define SF(l) (fb + offF[(l)])
  T* const tmp = sT + offTmp;  // Jacobi ping-pong partner

  // stage the top level; batches of 8 independent loads per thread
  {
    constexpr int BL = 8;
    const int n = a.H[0] * a.W[0];
    for (int o0 = threadIdx.x; o0 < n; o0 += BL * NT) {
      T pv[BL], rv[BL];
      uint8_t fv[BL];
#pragma unroll
      for (int b = 0; b < BL; ++b) {
        const int o = o0 + b * NT;
        const int oc = o < n ? o : n - 1;
        pv[b] = a.zero_top ? (T)0 : a.p_top_in[oc];
        rv[b] = a.rhs_top[oc];
        fv[b] = a.flags[0][oc];
      }
#pragma unroll
      for (int b = 0; b < BL; ++b) {
        const int o = o0 + b * NT;
        if (o < n) {
          SP(0)[o] = pv[b];
          SB(0)[o] = rv[b];
          SF(0)[o] = fv[b];
        }
      }
    }
    for (int l = 1; l < a.nl; ++l) {
      const int nl_ = a.H[l] * a.W[l];
      for (int o0 = threadIdx.x; o0 < nl_; o0 += BL * NT) {
        uint8_t fv[BL];
#pragma unroll
        for (int b = 0; b < BL; ++b) {
          const int o = o0 + b * NT;
          fv[b] = a.flags[l][o < nl_ ? o : nl_ - 1];
        }
#pragma unroll
        for (int b = 0; b < BL; ++b) {
          const int o = o0 + b * NT;
          if (o < nl_) {
            SF(l)[o] = fv[b];
            SP(l)[o] = (T)0;  // solver.py:127
          }
        }
      }
    }
  }
  __syncthreads();

  const int last = a.nl - 1;
  // down
  for (int l = 0; l < last; ++l) {
    const int H = a.H[l], W = a.W[l];
    small_smooth<T, MODE, NT, WAVE>(SP(l), tmp, SB(l), SF(l), H, W, a.alpha[l], a.omega[l],
                                    a.npass_smooth);
    const int Hc = a.H[l + 1], Wc = a.W[l + 1];
    const int N = H * W;
    for (int k = threadIdx.x; k < Hc * Wc; k += NT) {
      const int J = k / Wc, I = k - J * Wc;
      double total = 0.0;
      int count = 0;
#pragma unroll
      for (int dj = 0; dj < 2; ++dj)
#pragma unroll
        for (int di = 0; di < 2; ++di) {
          const int o = (2 * J + dj) * W + 2 * I + di;
          const uint8_t f = SF(l)[o];
          const T* p = SP(l);
          const T ce = p[o];
          const T lf = (f & F_LEFT) ? p[o > 0 ? o - 1 : 0] : ce;
          const T rt = (f & F_RIGHT) ? p[o + 1 < N ? o + 1 : N - 1] : ce;
          const T dn = (f & F_DOWN) ? p[o >= W ? o - W : 0] : ce;
          const T up = (f & F_UP) ? p[o + W < N ? o + W : N - 1] : ce;
          const T rv = (T)residual_value(ce, (T)(lf + rt + dn + up), SB(l)[o], a.alpha[l]);
          const bool fl = f & F_LFLUID;
          total += fl ? (double)rv : 0.0;  // +0.0 for a non-fluid child is exact
          count += fl ? 1 : 0;
        }
      const T mean = count > 0 ? (T)(total / (double)count) : (T)0;
      SB(l + 1)[k] = (T)((double)mean / a.coarse_div[l]);
    }
    __syncthreads();
  }
  // coarsest
  if (MODE == MODE_RB && a.cop != nullptr) {
    coarsest_apply_op<T, NT>(a.cop, SB(last), SP(last), a.H[last] * a.W[last]);
  } else if (MODE == MODE_RB &&
      (((WAVE & WAVE_LEVELS) && a.H[last] <= 32 && a.W[last] <= 32) ||
       ((WAVE & WAVE_COARSEST) && a.H[last] <= 16 && a.W[last] <= 16))) {
    if (threadIdx.x < 64)
      wave_smooth_level<T>(SP(last), SB(last), SF(last), a.H[last], a.W[last], a.alpha[last],
                           a.omega[last], a.npass_coarsest);
    __syncthreads();
  } else if (a.H[last] <= 16 && a.W[last] <= 16) {
    if (threadIdx.x < 64)
      coarsest_wave<T, MODE>(SP(last), SB(last), SF(last), a.H[last], a.W[last], a.alpha[last],
                             a.omega[last], a.npass_coarsest);
    __syncthreads();
  } else {
    small_smooth<T, MODE, NT, false>(SP(last), tmp, SB(last), SF(last), a.H[last], a.W[last],
                                     a.alpha[last], a.omega[last], a.npass_coarsest);
  }
  // up
  for (int l = last - 1; l >= 0; --l) {
    const int H = a.H[l], W = a.W[l];
    const int Hc = a.H[l + 1], Wc = a.W[l + 1];
    for (int o = threadIdx.x; o < H * W; o += NT) {
      const int j = o / W, i = o - j * W;
      T corr = (T)0;
      if (SF(l)[o] & F_LFLUID)
        corr = (T)sample_field(SP(l + 1), Hc, Wc, (i + 0.5) * 0.5 - 0.5, (j + 0.5) * 0.5 - 0.5);
      SP(l)[o] = SP(l)[o] + corr;
    }
    __syncthreads();
    small_smooth<T, MODE, NT, WAVE>(SP(l), tmp, SB(l), SF(l), H, W, a.alpha[l], a.omega[l],
                                    a.npass_smooth);
  }
  const int n0 = a.H[0] * a.W[0];
  for (int o = threadIdx.x; o < n0; o += NT) a.p_top_out[o] = SP(0)[o];
#undef SP
#undef SB
#undef SF
}

template <typename T, int MODE, int NT, int WAVE = 0>
__global__ __launch_bounds__(NT) void k_vcycle_small(const SmallArgs<T> a) {
  vcycle_small_body<T, MODE, NT, WAVE>(a);
}

// Column j of the float32 coarsest-level operator (coarsest_apply_op): the
// red-black passes of coarsest_wave, from p = 0 with rhs = unit vector j, on
// one wavefront per column.
__global__ __launch_bounds__(64) void k_coarse_op(const uint8_t* __restrict__ flags, int H, int W,
                                                  double alpha, double omega, int npass,
                                                  float* __restrict__ op) {
  __shared__ float p[256], rhs[256];
  __shared__ uint8_t fl[256];
  const int n = H * W, j = blockIdx.x;
  for (int i = threadIdx.x; i < 256; i += 64) {
    p[i] = 0.0f;
    rhs[i] = i == j ? 1.0f : 0.0f;
    fl[i] = i < n ? flags[i] : (uint8_t)0;
  }
  __syncthreads();
  coarsest_wave<float, MODE_RB>(p, rhs, fl, H, W, alpha, omega, npass);
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += 64) op[(size_t)j * n + i] = p[i];
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
                                                        float* __restrict__ defect,
                                                        const int* __restrict__ gate) {
  __shared__ double red[NB_T / 64];
  if (gate != nullptr && *gate == 0) return;
  int bx, by;
  xcd_tile(bx, by);
  const int i = bx * NB_X + (threadIdx.x & (NB_X - 1));
  const int j = by * NB_Y + (threadIdx.x / NB_X);
  double acc = 0.0;
  if (i < W && j < H) {
    const size_t o = (size_t)j * W + i;
    // as in k_correct_defect: every operand loaded (clamped indices) before
    // the flags are tested
    const size_t n = (size_t)H * W;
    const uint8_t f = flags[o];
    const T ce = p[o], pl = p[o > 0 ? o - 1 : 0], pr = p[o + 1 < n ? o + 1 : o];
    const T pd = p[o >= (size_t)W ? o - W : o], pu = p[o + W < n ? o + W : o];
    const T b = rhs[o];
    double rv = 0.0;
    if (f & F_LFLUID) {
      const T l = (f & F_LEFT) ? pl : ce;
      const T rr = (f & F_RIGHT) ? pr : ce;
      const T d = (f & F_DOWN) ? pd : ce;
      const T u = (f & F_UP) ? pu : ce;
      rv = poisson_residual(l, rr, d, u, ce, b, alpha);
      acc = rv * rv;
    }
    if (defect != nullptr) defect[o] = (f & F_LFLUID) ? (float)(rv / alpha) : 0.0f;
  }
  const double tot = block_sum<NB_T>(acc, red);
  if (threadIdx.x == 0) partial[by * gridDim.x + bx] = tot;
}

// The FP64 refinement step of the mixed solve in one pass: p_out = p_in + e
// on every cell (k_add_correction), then the defect of p_out exactly as
// k_residual_norm computes it, with every corrected value formed as
// p_in + e in the same rounding.  p_out must not alias p_in (neighbours read
// p_in while other threads write p_out).  p_in == nullptr: p_in is all zero.
template <typename T>
__global__ __launch_bounds__(NB_T) void k_correct_defect(const T* __restrict__ p_in,
                                                         const float* __restrict__ e,
                                                         T* __restrict__ p_out,
                                                         const T* __restrict__ rhs,
                                                         const uint8_t* __restrict__ flags,
                                                         int H, int W, double alpha,
                                                         double* __restrict__ partial,
                                                         float* __restrict__ defect,
                                                         const int* __restrict__ gate) {
  __shared__ double red[NB_T / 64];
  if (gate != nullptr && *gate == 0) return;
  int bx, by;
  xcd_tile(bx, by);
  const int i = bx * NB_X + (threadIdx.x & (NB_X - 1));
  const int j = by * NB_Y + (threadIdx.x / NB_X);
  double acc = 0.0;
  if (i < W && j < H) {
    const size_t o = (size_t)j * W + i;
    // the five corrected values are loaded at clamped indices before the
    // flags are tested (a closed face's value is never used), so a lane's
    // loads start together instead of waiting for its flags
    const size_t n = (size_t)H * W;
    const size_t ol = o > 0 ? o - 1 : 0, orr = o + 1 < n ? o + 1 : o;
    const size_t od = o >= (size_t)W ? o - W : o, ou = o + W < n ? o + W : o;
    auto q = [&](size_t k) -> T { return (p_in != nullptr ? p_in[k] : (T)0) + (T)e[k]; };
    const uint8_t f = flags[o];
    const T ce = q(o), ql = q(ol), qr = q(orr), qd = q(od), qu = q(ou);
    const T b = rhs[o];
    p_out[o] = ce;
    double rv = 0.0;
    if (f & F_LFLUID) {
      const T l = (f & F_LEFT) ? ql : ce;
      const T rr = (f & F_RIGHT) ? qr : ce;
      const T d = (f & F_DOWN) ? qd : ce;
      const T u = (f & F_UP) ? qu : ce;
      rv = poisson_residual(l, rr, d, u, ce, b, alpha);
      acc = rv * rv;
    }
    defect[o] = (f & F_LFLUID) ? (float)(rv / alpha) : 0.0f;
  }
  const double tot = block_sum<NB_T>(acc, red);
  if (threadIdx.x == 0) partial[by * gridDim.x + bx] = tot;
}

// ||alpha*rhs||^2 partials over fluid cells (the p = 0 residual), laid out
// as k_residual_norm's: for p = 0 both give the same partials, bit for bit.
// defect (optional): the FP32 defect of p = 0 exactly as k_residual_norm
// writes it.
template <typename T>
__global__ __launch_bounds__(NB_T) void k_rhs_norm(const T* __restrict__ rhs,
                                                   const uint8_t* __restrict__ flags, int H,
                                                   int W, double alpha,
                                                   double* __restrict__ partial,
                                                   float* __restrict__ defect) {
  __shared__ double red[NB_T / 64];
  const int i = blockIdx.x * NB_X + (threadIdx.x & (NB_X - 1));
  const int j = blockIdx.y * NB_Y + (threadIdx.x / NB_X);
  double acc = 0.0;
  if (i < W && j < H) {
    const size_t o = (size_t)j * W + i;
    const bool fl = (flags[o] & F_LFLUID) != 0;
    const T r = rhs[o];  // loaded with the flags, not after their test
    if (fl) {
      const double b = alpha * (double)r;
      acc = b * b;
    }
    if (defect != nullptr) {
      const double rv = fl ? poisson_residual((T)0, (T)0, (T)0, (T)0, (T)0, r, alpha) : 0.0;
      defect[o] = fl ? (float)(rv / alpha) : 0.0f;
    }
  }
  const double tot = block_sum<NB_T>(acc, red);
  if (threadIdx.x == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = tot;
}

// Sum n partials in a fixed order into out[slot].
// Convergence state of a tolerance-driven solve, kept on the device so that
// the host can queue the next iteration before it knows the outcome of the
// previous one: every launch of an iteration is gated on `active`.
struct SolveState {
  int active;   // 1 while the solve continues
  int cycles;   // refinements (mixed) or V-cycles (native) queued to run
  double rel;   // relative residual of the last check
};

// With `init`, also starts a gated solve's state (active, no cycle yet).
__global__ __launch_bounds__(1024) void k_sum_partials(const double* __restrict__ partial, int n,
                                                       double* __restrict__ out, int slot,
                                                       SolveState* __restrict__ init) {
  __shared__ double red[1024 / 64];
  // the partials loaded in batches before any is added, added in the order
  // of one after another
  constexpr int B = 16;
  double acc = 0.0;
  for (int k0 = threadIdx.x; k0 < n; k0 += B * 1024) {
    double v[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int k = k0 + b * 1024;
      v[b] = partial[k < n ? k : n - 1];
    }
#pragma unroll
    for (int b = 0; b < B; ++b)
      if (k0 + b * 1024 < n) acc += v[b];
  }
  const double tot = block_sum<1024>(acc, red);
  if (threadIdx.x == 0) {
    out[slot] = tot;
    if (init != nullptr) *init = SolveState{1, 0, -1.0};
  }
}

template <typename T>
__global__ void k_add_correction(T* __restrict__ p, const float* __restrict__ e, long n,
                                 const int* __restrict__ gate) {
  if (gate != nullptr && *gate == 0) return;
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = p[k] + (T)e[k];
}

// A slot of the pinned host ring: the state after one check, then the
// check's sequence number, written last (after a system-scope fence), so
// the host can poll the slot instead of waiting on an event.
struct RingSlot {
  int active;
  int cycles;
  double rel;
  unsigned seq;
  unsigned pad;
};

// A slot of the batched scenes' pinned ring: the active scene count after
// one iteration, then its sequence number.
struct CountSlot {
  int count;
  unsigned seq;
};

__global__ void k_state_init(SolveState* __restrict__ st) {
  st->active = 1;
  st->cycles = 0;
  st->rel = -1.0;
}

// One convergence check: the k_sum_partials reduction of ||r||^2, the
// relative residual against ||alpha div||^2 = *b2, and the decision the host
// loops of the solvers take (stop at rel <= tol, or after `cap` cycles --
// `inner` V-cycles per check).
//
// The resulting state is also written to *host_out (a slot of a pinned host
// ring, may be null) with sequence number seq, so the host learns the
// outcome without a copy or an event of its own on the stream.  A check
// queued after the solve has ended (state inactive) publishes nothing: the
// host never waits for it, and it may run after the ring has been handed to
// a later solve.
__device__ __forceinline__ void publish(RingSlot* slot, const SolveState& v, unsigned seq) {
  slot->active = v.active;
  slot->cycles = v.cycles;
  slot->rel = v.rel;
  __threadfence_system();
  slot->seq = seq;
}

__global__ __launch_bounds__(1024) void k_check(const double* __restrict__ partial, int n,
                                                const double* __restrict__ b2,
                                                SolveState* __restrict__ st, double tol,
                                                int cap, int inner,
                                                RingSlot* __restrict__ host_out, unsigned seq) {
  __shared__ double red[1024 / 64];
  // every load is issued before any result is used (the state, the rhs
  // norm, the partials in batches), so the check waits on memory once; a
  // thread adds its partials in the same order as one after another
  constexpr int B = 16;
  const int active = st->active, c = st->cycles;
  const double bb = *b2;
  double acc = 0.0;
  for (int k0 = threadIdx.x; k0 < n; k0 += B * 1024) {
    double v[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int k = k0 + b * 1024;
      v[b] = partial[k < n ? k : n - 1];
    }
#pragma unroll
    for (int b = 0; b < B; ++b)
      if (k0 + b * 1024 < n) acc += v[b];
  }
  if (active == 0) return;
  const double tot = block_sum<1024>(acc, red);
  if (threadIdx.x != 0) return;
  const double rel = bb > 0.0 ? sqrt(tot / bb) : 0.0;
  st->rel = rel;
  const bool done = (tol > 0.0 ? rel <= tol : c >= cap) || c * inner >= cap;
  SolveState out{done ? 0 : 1, done ? c : c + 1, rel};
  *st = out;
  if (host_out != nullptr) publish(host_out, out, seq);
}

__global__ void k_solid_from_flags(const uint8_t* __restrict__ flags, long n,
                                   uint8_t* __restrict__ solid) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) solid[k] = (flags[k] & F_LFLUID) ? 0 : 1;
}

// ------------------------------------------------------------------ host side
constexpr int TILE_NT = 256;
constexpr int TILE_RH = 48;
constexpr int MAX_R = 8;
constexpr int MAX_R_DEEP = 24;  // red-black single-grid launches on small grids
constexpr int SMALL_NT = 256;

inline int tile_tx(int R) { return 64 - 2 * R; }
inline int tile_ty(int R) { return TILE_RH - 2 * R; }

// ------------------------------------------------------------ profiling
// Optional per-launch timing with HIP events recorded on the launch stream,
// tagged with the launch's algorithmic bytes (one compulsory read or write per
// field per logical pass: a red-black sweep reads p and rhs and writes p; a
// residual-restriction reads p, rhs and writes a quarter-size coarse rhs; a
// prolong-correction reads a quarter-size coarse e and reads/writes p; a norm
// reads p and rhs).  Off by default; the benchmark turns it on for a few
// untimed steps.
struct ProfRec {
  const char* name;
  hipEvent_t start, stop;
  double bytes;
  int launches;  // kernel launches inside the event pair
};
static bool g_prof_on = false;
// Finest-level smoothing launches are idempotent (p_in -> p_out, coarse rhs
// or norm partials from p_in): while profiling, each one is repeated this
// many times inside its event pair, so the pair's own overhead (a few us,
// against a ~30 us kernel) is spread over the repeats.
static int g_prof_repeat = 1;
static std::vector<ProfRec> g_prof;
// Events are recycled through a pool: creating two per launch on the host
// would slow submission enough to leave gaps inside the timed intervals of
// short kernels.
static std::vector<hipEvent_t> g_prof_pool;

static hipEvent_t prof_event() {
  if (!g_prof_pool.empty()) {
    hipEvent_t e = g_prof_pool.back();
    g_prof_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

static hipEvent_t prof_begin(hipStream_t s) {
  if (!g_prof_on) return nullptr;
  hipEvent_t e = prof_event();
  if (e) (void)hipEventRecord(e, s);
  return e;
}

static void prof_end(hipEvent_t start, const char* name, double bytes, hipStream_t s,
                     int launches = 1) {
  if (start == nullptr) return;
  hipEvent_t e = prof_event();
  if (e == nullptr) return;
  (void)hipEventRecord(e, s);
  g_prof.push_back({name, start, e, bytes * launches, launches});
}

template <typename T>
double smooth_alg_bytes(int mode, const SmoothArgs<T>& a) {
  const double es = (double)sizeof(T), n = (double)a.H * a.W;
  double b = (mode == MODE_RB ? 0.5 * a.npass : (double)a.npass) * 3.0 * es * n;
  if (a.coarse_rhs != nullptr) b += 2.25 * es * n;
  if (a.coarse_e != nullptr) b += 2.25 * es * n;
  if (a.norm_partial != nullptr) b += 2.0 * es * n;
  return b;
}

constexpr int RB_NW = 4, RB_BR = 16, RB_RH = RB_NW * RB_BR;  // register kernel region 64 x 64
constexpr int JAC_RH = 32;

// Tile variants of the smoothing launch.  The fine level (HBM-bound) and the
// coarse levels (latency-bound: few tiles) are tuned separately, and the fine
// level gets its own kernel symbol so profiles separate the two.
enum : int {
  VAR_LDS48 = 0,   // 64 x 48 region, p / rhs / flags in LDS
  VAR_LDS32 = 1,   // 64 x 32 region, p / rhs / flags in LDS
  VAR_REG64 = 2,   // 64 x 64 region in registers (k_rb_reg)
  VAR_LDS48R = 3,  // 64 x 48 region, p in LDS, rhs / flags in registers, 256 threads
  VAR_LDS32R = 4,  // 64 x 32 region, as VAR_LDS48R
  VAR_LDS64R = 5,  // 64 x 64 region, as VAR_LDS48R with 512 threads
  VAR_LDS16 = 6,   // 64 x 16 region, p / rhs / flags in LDS
  VAR_DMA32 = 7,   // 64 x 32 region, p to LDS by LDS-DMA, rhs / flags in registers
  VAR_DMA64 = 8,   // 64 x 64 region, as VAR_DMA32 with 512 threads
  VAR_DMA16 = 9,   // 64 x 16 region, as VAR_DMA32
  NUM_VARIANTS = 10
};
// Defaults per dtype [f32?][finest, coarse, small coarse], from timing every
// variant on the launches of the 2048^2 V(2,2) workload (4 passes and the
// restriction, or the prolongation and 4 passes): finest level 64 x 64 with p
// in LDS and the rhs and flags in registers (float64 31 / 30 us against 34 /
// 33 us with p staged by LDS-DMA; float32 27 / 28 us), the 1024^2 level
// 64 x 32 likewise (float64 26 us for both launches against 28 us all in
// LDS).  Coarse levels of at most SMALL_TILED_CELLS cells occupy a few CUs,
// so a launch lasts as long as one workgroup: 64 x 16 regions all in LDS halve
// the work on that critical path (two cells per lane and pass instead of
// four).
static int g_variant[2][3] = {{VAR_LDS64R, VAR_LDS32R, VAR_LDS16},
                              {VAR_LDS64R, VAR_LDS32R, VAR_LDS16}};
constexpr long SMALL_TILED_CELLS = 512L * 512L;

inline int variant_rh(int v);

// The variant of a launch; a region too short for the launch's halo (output
// rows = region rows - 2R < 2) falls back to the 64 x 32 LDS tile.
inline int pick_variant(bool f32, bool fine, int H, int W, int R) {
  const int v = fine ? g_variant[f32][0]
                     : ((long)H * W <= SMALL_TILED_CELLS ? g_variant[f32][2] : g_variant[f32][1]);
  return variant_rh(v) - 2 * R < 2 ? VAR_LDS32 : v;
}

// Threads of the single-workgroup small-level V-cycle, per dtype [f32?].
static int g_small_nt[2] = {1024, 1024};
// red-black levels of at most 32 x 32 cells (and the coarsest) on one
// wavefront with DPP / permlane exchange, else the multi-wave LDS passes
static int g_small_wave[2] = {0, 0};
constexpr int SMALL_NT_CHOICES[8] = {256, 512, 1024, 256, 512, 256, 512, 1024};
constexpr int SMALL_WAVE_CHOICES[8] = {0, 0, 0, 3, 3, 1, 1, 1};

inline int variant_rh(int v) {
  switch (v) {
    case VAR_LDS48: case VAR_LDS48R: return 48;
    case VAR_LDS32: case VAR_LDS32R: case VAR_DMA32: return 32;
    case VAR_LDS16: case VAR_DMA16: return 16;
    default: return RB_RH;  // VAR_REG64, VAR_LDS64R, VAR_DMA64
  }
}

template <typename T>
int launch_tiled_once(int mode, const SmoothArgs<T>& a, hipStream_t s, bool fine);

template <typename T>
int launch_tiled(int mode, const SmoothArgs<T>& a, hipStream_t s, bool fine = false) {
  const int rlim = mode == MODE_RB ? MAX_R_DEEP : (JAC_RH - 8) / 2;
  if (a.R < 1 || a.R > rlim) return fail(SMK_EINVAL, "halo %d outside [1, %d]", a.R, rlim);
  hipEvent_t ev = prof_begin(s);
  const bool f32 = sizeof(T) == 4;
  const char* tag = mode == MODE_RB
                        ? (fine ? (f32 ? "rb_smooth_fine_f32" : "rb_smooth_fine_f64")
                                : (f32 ? "rb_smooth_coarse_f32" : "rb_smooth_coarse_f64"))
                        : (fine ? (f32 ? "jacobi_fine_f32" : "jacobi_fine_f64")
                                : (f32 ? "jacobi_coarse_f32" : "jacobi_coarse_f64"));
  const int reps = (ev != nullptr && fine) ? g_prof_repeat : 1;
  struct Done {
    hipEvent_t ev;
    const char* tag;
    double bytes;
    hipStream_t s;
    int reps;
    ~Done() { prof_end(ev, tag, bytes, s, reps); }
  } done{ev, tag, ev ? smooth_alg_bytes<T>(mode, a) : 0.0, s, reps};
  for (int rep = 1; rep < reps; ++rep) {
    int rc = launch_tiled_once<T>(mode, a, s, fine);
    if (rc) return rc;
  }
  return launch_tiled_once<T>(mode, a, s, fine);
}

template <typename T>
int launch_tiled_once(int mode, const SmoothArgs<T>& a, hipStream_t s, bool fine) {
  if (mode == MODE_RB) {
    const int v = pick_variant(sizeof(T) == 4, fine, a.H, a.W, a.R);
    const int rh = variant_rh(v);
    if (v == VAR_REG64 && a.coarse_rhs != nullptr && (a.R & 1))
      return fail(SMK_EINVAL, "restriction needs an even halo (got %d)", a.R);
    dim3 grid(cdiv(a.W, tile_tx(a.R)), cdiv(a.H, rh - 2 * a.R));
    if (v == VAR_REG64)
      hipLaunchKernelGGL((k_rb_reg<T, RB_NW, RB_BR>), grid, dim3(RB_NW * 64), 0, s, a);
#define SMK_R_LAUNCH(RHV, NTV)                                                             \
  do {                                                                                     \
    if (fine)                                                                              \
      hipLaunchKernelGGL((k_smooth_fine_r<T, RHV, NTV>), grid, dim3(NTV), 0, s, a);        \
    else                                                                                   \
      hipLaunchKernelGGL((k_smooth_tile64_r<T, RHV, NTV>), grid, dim3(NTV), 0, s, a);      \
  } while (0)
    else if (v == VAR_LDS48R)
      SMK_R_LAUNCH(48, 256);
    else if (v == VAR_LDS32R)
      SMK_R_LAUNCH(32, 256);
    else if (v == VAR_LDS64R)
      SMK_R_LAUNCH(64, 512);
#undef SMK_R_LAUNCH
#define SMK_D_LAUNCH(RHV, NTV)                                                             \
  do {                                                                                     \
    if (fine)                                                                              \
      hipLaunchKernelGGL((k_smooth_fine_d<T, RHV, NTV>), grid, dim3(NTV), 0, s, a);        \
    else                                                                                   \
      hipLaunchKernelGGL((k_smooth_tile64_d<T, RHV, NTV>), grid, dim3(NTV), 0, s, a);      \
  } while (0)
    else if (v == VAR_DMA32)
      SMK_D_LAUNCH(32, 256);
    else if (v == VAR_DMA64)
      SMK_D_LAUNCH(64, 512);
    else if (v == VAR_DMA16)
      SMK_D_LAUNCH(16, 256);
#undef SMK_D_LAUNCH
    else if (v == VAR_LDS32)
      hipLaunchKernelGGL((k_smooth_tile64<T, MODE_RB, 32, TILE_NT>), grid, dim3(TILE_NT), 0, s, a);
    else if (v == VAR_LDS16)
      hipLaunchKernelGGL((k_smooth_tile64<T, MODE_RB, 16, TILE_NT>), grid, dim3(TILE_NT), 0, s, a);
    else if (fine)
      hipLaunchKernelGGL((k_smooth_fine<T, 48, TILE_NT>), grid, dim3(TILE_NT), 0, s, a);
    else
      hipLaunchKernelGGL((k_smooth_tile64<T, MODE_RB, 48, TILE_NT>), grid, dim3(TILE_NT), 0, s, a);
    return check_launch("red-black tile kernel");
  }
  hipLaunchKernelGGL((k_smooth_tile64<T, MODE_JACOBI, JAC_RH, TILE_NT>),
                     dim3(cdiv(a.W, tile_tx(a.R)), cdiv(a.H, JAC_RH - 2 * a.R)), dim3(TILE_NT), 0,
                     s, a);
  return check_launch("k_smooth_tile64");
}

// partial-sum slots one tiled launch can produce
inline int tiled_blocks(int mode, int H, int W, int R, bool fine = false, bool f32 = false) {
  const int rh = mode == MODE_RB ? variant_rh(pick_variant(f32, fine, H, W, R)) : JAC_RH;
  return (int)(cdiv(W, tile_tx(R)) * cdiv(H, rh - 2 * R));
}

inline size_t small_lds_bytes(const int* Hs, const int* Ws, int nl, size_t es, int mode) {
  // must match the layout computed in k_vcycle_small
  const size_t padt = 16 / es;
  size_t t = 0, f = 0, nmax = 0;
  for (int l = 0; l < nl; ++l) {
    const size_t n = (size_t)Hs[l] * Ws[l];
    const size_t np = (n + padt - 1) / padt * padt;
    nmax = n > nmax ? n : nmax;
    t += 2 * np;
    f += (n + 15) & ~(size_t)15;
  }
  if (mode == MODE_JACOBI) t += (nmax + padt - 1) / padt * padt;
  return t * es + f;
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
    const void* fns[] = {
        reinterpret_cast<const void*>(&k_vcycle_small<T, MODE_RB, 256>),
        reinterpret_cast<const void*>(&k_vcycle_small<T, MODE_JACOBI, 256>),
        reinterpret_cast<const void*>(&k_vcycle_small<T, MODE_RB, 512>),
        reinterpret_cast<const void*>(&k_vcycle_small<T, MODE_JACOBI, 512>),
        reinterpret_cast<const void*>(&k_vcycle_small<T, MODE_RB, 1024>),
        reinterpret_cast<const void*>(&k_vcycle_small<T, MODE_JACOBI, 1024>),
        reinterpret_cast<const void*>(&k_vcycle_small<T, MODE_RB, 256, 3>),
        reinterpret_cast<const void*>(&k_vcycle_small<T, MODE_RB, 512, 3>),
        reinterpret_cast<const void*>(&k_vcycle_small<T, MODE_RB, 256, 1>),
        reinterpret_cast<const void*>(&k_vcycle_small<T, MODE_RB, 512, 1>),
        reinterpret_cast<const void*>(&k_vcycle_small<T, MODE_RB, 1024, 1>)};
    bool ok = true;
    for (const void* f : fns)
      ok = ok && hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)want) ==
                     hipSuccess;
    if (!ok) {
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
  hipEvent_t ev = prof_begin(s);
  double bytes = 0.0;
  if (ev) {
    const double es = (double)sizeof(T);
    const double sweeps = mode == MODE_RB ? 0.5 : 1.0;
    for (int l = 0; l < a.nl; ++l) {
      const double n = (double)a.H[l] * a.W[l];
      const bool last = l == a.nl - 1;
      bytes += (last ? a.npass_coarsest : 2 * a.npass_smooth) * sweeps * 3.0 * es * n;
      if (!last) bytes += 4.5 * es * n;
    }
  }
  struct Done {
    hipEvent_t ev;
    double bytes;
    hipStream_t s;
    ~Done() { prof_end(ev, sizeof(T) == 4 ? "vcycle_small_f32" : "vcycle_small_f64", bytes, s); }
  } done{ev, bytes, s};
  const int nt = g_small_nt[sizeof(T) == 4];
  const int wave = mode == MODE_RB ? g_small_wave[sizeof(T) == 4] : 0;
#define SMK_SMALL_LAUNCH(NTV)                                                                   \
  do {                                                                                         \
    if (mode == MODE_RB)                                                                       \
      hipLaunchKernelGGL((k_vcycle_small<T, MODE_RB, NTV>), dim3(1), dim3(NTV), lds, s, a);    \
    else                                                                                       \
      hipLaunchKernelGGL((k_vcycle_small<T, MODE_JACOBI, NTV>), dim3(1), dim3(NTV), lds, s, a); \
  } while (0)
#define SMK_SMALL_WAVE(NTV, WV) \
  hipLaunchKernelGGL((k_vcycle_small<T, MODE_RB, NTV, WV>), dim3(1), dim3(NTV), lds, s, a)
  if (wave == 3 && nt == 512)
    SMK_SMALL_WAVE(512, 3);
  else if (wave == 3)
    SMK_SMALL_WAVE(256, 3);
  else if (wave == 1 && nt == 1024)
    SMK_SMALL_WAVE(1024, 1);
  else if (wave == 1 && nt == 512)
    SMK_SMALL_WAVE(512, 1);
  else if (wave == 1)
    SMK_SMALL_WAVE(256, 1);
#undef SMK_SMALL_WAVE
  else if (nt == 1024)
    SMK_SMALL_LAUNCH(1024);
  else if (nt == 512)
    SMK_SMALL_LAUNCH(512);
  else
    SMK_SMALL_LAUNCH(256);
#undef SMK_SMALL_LAUNCH
  return check_launch("k_vcycle_small");
}

constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

inline int norm_blocks(int H, int W) { return (int)(cdiv(W, NB_X) * cdiv(H, NB_Y)); }

// Workspace of the single-grid solvers: flags | p scratch | partials | results
struct SolverWs {
  uint8_t* flags;
  void* ptmp;
  double* partial;
  double* result;  // [0] = ||r||^2, [1] = ||alpha b||^2
  SolveState* state;
  float* e32[2];   // FP32 correction ping-pong (mixed precision, FP64 only)
  float* defect32;
  size_t bytes;
};

inline size_t solver_ws_bytes(int H, int W, int dtype) {
  const size_t n = (size_t)H * W;
  const size_t es = dtype == SMK_F64 ? 8 : 4;
  const size_t mixed = dtype == SMK_F64 ? 3 * align256(n * 4) : 0;
  return align256(n) + align256(n * es) + align256((size_t)norm_blocks(H, W) * 8) + 256 + 256 +
         mixed;
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
  b += 256;
  w.state = reinterpret_cast<SolveState*>(b);
  b += 256;
  w.e32[0] = w.e32[1] = w.defect32 = nullptr;
  if (dtype == SMK_F64) {
    w.e32[0] = reinterpret_cast<float*>(b);
    b += align256(n * 4);
    w.e32[1] = reinterpret_cast<float*>(b);
    b += align256(n * 4);
    w.defect32 = reinterpret_cast<float*>(b);
  }
  w.bytes = solver_ws_bytes(H, W, dtype);
  return w;
}

// Passes per launch of the single-grid solvers (red-black and Jacobi) on
// grids of at most RB_DEEP_CELLS cells, where a launch is latency-bound: a deeper halo
// costs redundant passes on chip but saves launches (512^2 FP64 RB-SOR with
// a check every 20 sweeps: 90.6 ms per step at 8 passes per launch, 74.8 at
// 16, 68.5 at 20, with the 40 passes between checks in two launches).
// Capped so that the finest-level variant keeps at least 16 exact rows per
// region; only the register-rhs bodies take halos beyond MAX_R + 1.
constexpr long RB_DEEP_CELLS = 1024L * 1024L;

template <typename T>
int max_passes_per_launch(int mode, int H, int W) {
  // Jacobi: the 32-row regions keep at least 8 exact rows (12 passes; the
  // 128^2 Jacobi-200 workload 0.318 -> 0.273 ms per step against 8)
  if (mode == MODE_JACOBI && (long)H * W <= RB_DEEP_CELLS) return (JAC_RH - 8) / 2;
  if (mode != MODE_RB || (long)H * W > RB_DEEP_CELLS) return MAX_R;
  const int v = g_variant[sizeof(T) == 4][0];
  if (v != VAR_LDS64R && v != VAR_LDS48R && v != VAR_LDS32R) return MAX_R;
  const int rh = variant_rh(v);
  int c = MAX_R_DEEP;
  while (c > MAX_R && rh - 2 * c < 16) --c;
  return c;
}

// The pass counts of a run of launches without storing them: nl launches
// sharing npass passes evenly (the first npass % nl one more), launch
// `split` (if >= 0) divided into two.
struct PassChunks {
  int nl, npass, split;
  int count() const { return nl + (split >= 0 ? 1 : 0); }
  int base(int q) const { return npass / nl + (q < npass % nl ? 1 : 0); }
  int at(int q) const {
    if (split < 0 || q < split) return base(q);
    if (q == split) return base(split) / 2;
    if (q == split + 1) return base(split) - base(split) / 2;
    return base(q - 1);
  }
};

// Runs `npass_total` passes from src into the buffer `dst_final`, using
// `other` as the ping-pong partner; chunked into launches of at most
// max_passes_per_launch passes.  The chunking keeps the launch count's
// parity such that the last launch writes dst_final (src may equal
// dst_final).
template <typename T>
int run_passes(int mode, T* dst_final, T* other, const T* src, const T* rhs,
               const uint8_t* flags, int H, int W, double alpha, double omega,
               int npass_total, int first_colour, hipStream_t s, const int* gate = nullptr,
               bool zero_first = false) {
  if (npass_total <= 0) {
    if (src != dst_final)
      SMK_CHECK_HIP(hipMemcpyAsync(dst_final, src, (size_t)H * W * sizeof(T),
                                   hipMemcpyDeviceToDevice, s));
    return SMK_OK;
  }
  // as few launches as the cap allows, the passes spread evenly over them
  const int cmax = max_passes_per_launch<T>(mode, H, W);
  PassChunks chunks{(npass_total + cmax - 1) / cmax, npass_total, -1};
  // launches alternate buffers: src -> X -> Y -> ... ; the last must be dst_final.
  // With src == dst_final an even launch count ends in dst_final; with src ==
  // other an odd count does.  Split a chunk to fix the parity, or copy.
  const bool src_is_final = (src == dst_final);
  const bool need_even = src_is_final || (src != other);
  // Generic source (neither buffer): first launch writes `other` if the count
  // is even, dst_final if odd -- so any count works when src is foreign.
  const bool foreign = (src != dst_final && src != other);
  if (!foreign) {
    const bool even = (chunks.count() % 2) == 0;
    if (need_even != even) {
      if (chunks.base(0) >= 2) {  // the largest chunk comes first
        chunks.split = 0;
      } else {
        // single 1-pass chunk with the wrong parity: copy first
        T* to = src_is_final ? other : dst_final;
        SMK_CHECK_HIP(hipMemcpyAsync(to, src, (size_t)H * W * sizeof(T),
                                     hipMemcpyDeviceToDevice, s));
        src = to;
      }
    }
  }
  const int L = chunks.count();
  // destination of launch q: counted back from the last, which is dst_final
  const T* in = src;
  int colour = first_colour;
  for (int q = 0; q < L; ++q) {
    const int npass = chunks.at(q);
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
    a.npass = npass;
    a.first_colour = colour;
    a.R = npass;
    a.zero_init = (zero_first && q == 0) ? 1 : 0;
    a.coarse_div = 1.0;
    a.gate = gate;
    int rc = launch_tiled<T>(mode, a, s, true);
    if (rc) return rc;
    colour = (colour + npass) & 1;
    in = out;
  }
  return SMK_OK;
}

template <typename T>
int residual_norm2_dev(const T* p, const T* rhs, const uint8_t* flags, int H, int W,
                       double alpha, double* partial, double* result, int slot, float* defect,
                       hipStream_t s, const int* gate = nullptr) {
  dim3 g(cdiv(W, NB_X), cdiv(H, NB_Y));
  hipEvent_t ev = prof_begin(s);
  hipLaunchKernelGGL(k_residual_norm<T>, g, dim3(NB_T), 0, s, p, rhs, flags, H, W, alpha,
                     partial, defect, gate);
  prof_end(ev, defect ? "defect_fp64" : "residual_norm",
           (2.0 * sizeof(T) + (defect ? 4.0 : 0.0)) * (double)H * W, s);
  int rc = check_launch("k_residual_norm");
  if (rc) return rc;
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, partial, (int)(g.x * g.y),
                     result, slot, nullptr);
  return check_launch("k_sum_partials");
}

// k_residual_norm only: the partials for a k_check of the same launch shape.
template <typename T>
int residual_partials_dev(const T* p, const T* rhs, const uint8_t* flags, int H, int W,
                          double alpha, double* partial, float* defect, const int* gate,
                          hipStream_t s) {
  dim3 g(cdiv(W, NB_X), cdiv(H, NB_Y));
  hipEvent_t ev = prof_begin(s);
  hipLaunchKernelGGL(k_residual_norm<T>, g, dim3(NB_T), 0, s, p, rhs, flags, H, W, alpha,
                     partial, defect, gate);
  prof_end(ev, defect ? "defect_fp64" : "residual_norm",
           (2.0 * sizeof(T) + (defect ? 4.0 : 0.0)) * (double)H * W, s);
  return check_launch("k_residual_norm");
}

// ||alpha rhs||^2 over the fluid cells into result[slot] (with `defect`, also
// the defect of p = 0; with `init`, the reduction also starts a gated solve's
// state, in place of a k_state_init launch).
template <typename T>
int rhs_norm2_dev(const T* rhs, const uint8_t* flags, int H, int W, double alpha,
                  double* partial, double* result, int slot, hipStream_t s,
                  float* defect = nullptr, SolveState* init = nullptr) {
  dim3 g(cdiv(W, NB_X), cdiv(H, NB_Y));
  hipLaunchKernelGGL(k_rhs_norm<T>, g, dim3(NB_T), 0, s, rhs, flags, H, W, alpha, partial,
                     defect);
  int rc = check_launch("k_rhs_norm");
  if (rc) return rc;
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, partial, (int)(g.x * g.y),
                     result, slot, init);
  return check_launch("k_sum_partials");
}

inline double calc_alpha(double h, double rho, double dt) { return (-(h * h) * rho) / dt; }

inline double optimal_omega(int size) {
  const double w = 2.0 / (1.0 + sin(M_PI / (double)size));
  return w < 1.95 ? w : 1.95;
}

// Queue gated iterations -- iter(k, slot, seq) enqueues iteration k, ending
// with a k_check on d_state that also publishes the state to `slot`, the
// device address of pinned ring slot k % 4, under sequence number `seq` --
// until the device reports the solve done.  The outcome of iteration k-1 is
// polled from its slot while iteration k is already queued, so the GPU never
// waits for the host; the one iteration queued after the last useful one is
// gated off and does nothing.  Polling replaces an event per iteration (an
// event record costs the stream a system-scope release, ~5 us of idle GPU).
//
// When a slot reports the solve done, that slot holds the final state (the
// queued no-op iteration leaves it unchanged), so the function returns at
// once, without waiting for the stream: work the caller queues next follows
// the no-op iteration on the stream and the GPU does not idle while the host
// catches up.  While polling, the stream is queried now and then, so an
// error on the stream ends the wait; a loop that runs out of iterations
// reads the state after a stream synchronisation.
//
// The no-op iteration is not free (each of its launches still costs ~2 us of
// dispatch), so with tol > 0 the loop predicts: when the last two outcomes
// read, r2 then r1, extrapolate geometrically to r1 * (r1 / r2) <= tol for
// the iteration in flight, it reads that iteration's outcome before queueing
// the next one.  A right guess saves the no-op iteration for the host's
// return latency; a wrong one costs the GPU that latency once.  The
// iterations and their results are the same either way.
// Sequence numbers of the pinned read-back rings (single-grid solves,
// multigrid handles, batches), unique in the process.  A ring's memory can
// be that of a freed ring, whose slots may still show the old bytes (a
// publication that lands after the new ring was cleared); with every
// publication numbered once per process, no such slot can pass for the
// answer to a later check.
inline unsigned next_ring_sequence(unsigned count) {
  static std::atomic<unsigned> next{1u};
  return next.fetch_add(count);
}

template <typename F>
int run_gated_ring(SolveState* d_state, RingSlot* h_ring, RingSlot* d_ring, int max_iter,
                   double tol, F&& iter, SolveState* fin, hipStream_t s) {
  volatile RingSlot* ring = h_ring;
  const unsigned base = next_ring_sequence((unsigned)max_iter + 2u);
  // while profiling, wait for each iteration before queueing the next, so
  // that no gated-off launch enters the per-kernel averages
  const int lag = g_prof_on ? 0 : 1;
  int known = -1;               // the last iteration whose outcome was read
  double r1 = -1.0, r2 = -1.0;  // the last two relative residuals read
  // read iteration it's outcome; *done = 1 when it reports the solve done
  auto read_outcome = [&](int it, int* done) -> int {
    const int slot = it & 3;
    const unsigned want = base + (unsigned)it;
    for (unsigned spin = 1; ring[slot].seq != want; ++spin) {
      if ((spin & 1023u) == 0u) {
        const hipError_t e = hipStreamQuery(s);
        if (e == hipSuccess && ring[slot].seq != want)
          return fail((int)hipErrorUnknown, "gated solve: check %u never published", want);
        if (e != hipSuccess && e != hipErrorNotReady)
          return fail((int)e, "gated solve: %s", hipGetErrorString(e));
      }
    }
    __atomic_thread_fence(__ATOMIC_ACQUIRE);
    known = it;
    r2 = r1;
    r1 = ring[slot].rel;
    *done = ring[slot].active == 0;
    if (*done) {
      fin->active = ring[slot].active;
      fin->cycles = ring[slot].cycles;
      fin->rel = ring[slot].rel;
    }
    return SMK_OK;
  };
  for (int k = 0; k <= max_iter; ++k) {
    int done = 0, rc;
    if (lag > 0 && k - 1 > known && tol > 0.0 && r1 > 0.0 && r2 > 0.0 &&
        r1 * (r1 / r2) <= tol) {
      rc = read_outcome(k - 1, &done);
      if (rc || done) return rc;
    }
    rc = iter(k, d_ring + (k & 3), base + (unsigned)k);
    if (rc) return rc;
    if (k - lag < 0 || k - lag <= known) continue;
    rc = read_outcome(k - lag, &done);
    if (rc || done) return rc;
  }
  SMK_CHECK_HIP(hipStreamSynchronize(s));
  SMK_CHECK_HIP(hipMemcpy(fin, d_state, sizeof(SolveState), hipMemcpyDeviceToHost));
  return SMK_OK;
}

// Pinned read-back ring and events of the single-grid solvers (which take a
// caller workspace, not a handle): one set per host thread, kept until exit.
struct GateRing {
  RingSlot* h_ring = nullptr;
  RingSlot* d_ring = nullptr;  // device address of h_ring
};

static int thread_gate_ring(GateRing** out) {
  thread_local GateRing g;
  if (g.h_ring == nullptr) {
    SMK_CHECK_HIP(hipHostMalloc((void**)&g.h_ring, 4 * sizeof(RingSlot), hipHostMallocDefault));
    memset(g.h_ring, 0, 4 * sizeof(RingSlot));
    SMK_CHECK_HIP(hipHostGetDevicePointer((void**)&g.d_ring, g.h_ring, 0));
  }
  *out = &g;
  return SMK_OK;
}

// Tolerance-driven sweeps, gated: iteration 0 checks the initial guess,
// iteration k >= 1 runs chunk k of `check_every` sweeps (the last chunk may
// be shorter) then checks.  k_check's cap is the number of chunks.
template <typename F>
int sweeps_gated(SolverWs& w, int iterations, int check_every, double tol, F&& chunk,
                 const double* b2, int nparts, int* done_out, double* rel_out, hipStream_t s) {
  GateRing* g = nullptr;
  int rc = thread_gate_ring(&g);
  if (rc) return rc;
  const int nchunks = (iterations + check_every - 1) / check_every;
  hipLaunchKernelGGL(k_state_init, dim3(1), dim3(1), 0, s, w.state);
  rc = check_launch("k_state_init");
  if (rc) return rc;
  SolveState fin{};
  rc = run_gated_ring(w.state, g->h_ring, g->d_ring, nchunks + 1, tol,
                      [&](int k, RingSlot* slot, unsigned sq) -> int {
    int r = SMK_OK;
    if (k > 0) {
      const int start = (k - 1) * check_every;
      const int n = iterations - start < check_every ? iterations - start : check_every;
      if (n > 0) r = chunk(n, &w.state->active);
    }
    if (r) return r;
    r = chunk(0, &w.state->active);  // the residual partials
    if (r) return r;
    hipLaunchKernelGGL(k_check, dim3(1), dim3(1024), 0, s, w.partial, nparts, b2, w.state, tol,
                       nchunks, 1, slot, sq);
    return check_launch("k_check");
  }, &fin, s);
  if (rc) return rc;
  const long done = (long)fin.cycles * check_every;
  *done_out = (int)(done < iterations ? done : iterations);
  *rel_out = fin.rel;
  return SMK_OK;
}

template <typename T>
int rbsor_solve_t(T* p, const T* div, const uint8_t* fmu, const uint8_t* fmv, int H, int W,
                  double rho, double h, double dt, int iterations, double omega, double tol,
                  int check_every, void* ws, smk_solve_info* info, hipStream_t s,
                  bool gated = true) {
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
  if (gated) {
    int done = 0;
    double rel = -1.0;
    rc = sweeps_gated(w, iterations, check_every, tol, [&](int n, const int* gate) -> int {
      if (n == 0)
        return residual_partials_dev<T>(p, div, w.flags, H, W, alpha, w.partial, nullptr, gate, s);
      return run_passes<T>(MODE_RB, p, tmp, p, div, w.flags, H, W, alpha, omega, 2 * n, 0, s,
                           gate);
    }, w.result + 1, norm_blocks(H, W), &done, &rel, s);
    if (rc) return rc;
    if (info) {
      info->iterations = done;
      info->converged = rel <= tol;
      info->rel_residual = rel;
    }
    return SMK_OK;
  }
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

// gauss_seidel.py:8-27 -- lexicographic Gauss-Seidel, as a wavefront.
// In sweep s, cell (j, i) reads the new values of (j, i-1) and (j-1, i) and
// the previous sweep's (j, i+1) and (j+1, i).  With t(s, j, i) = i + j + 2s
// every one of those is produced at time t-1, and no two cells of one time
// step are neighbours, so all cells of a time step update in place at once
// and each reads exactly the operands the serial loop gives it: the result
// is the reference's sequential sweep, bit for bit.  One workgroup walks the
// (H-1)+(W-1)+2(S-1)+1 time steps with a barrier between them.
constexpr int GS_NT = 1024;

template <typename T>
__global__ __launch_bounds__(GS_NT) void k_gs_wavefront(T* __restrict__ p,
                                                        const T* __restrict__ div,
                                                        const uint8_t* __restrict__ flags, int H,
                                                        int W, double alpha, int sweeps,
                                                        const int* __restrict__ gate) {
  if (gate != nullptr && *gate == 0) return;
  const int tmax = (H - 1) + (W - 1) + 2 * (sweeps - 1);
  for (int t = 0; t <= tmax; ++t) {
    const int over = t - (H - 1) - (W - 1);
    const int s_lo = over <= 0 ? 0 : (over + 1) / 2;
    const int s_hi = (t / 2) < (sweeps - 1) ? (t / 2) : (sweeps - 1);
    for (int sw = s_lo; sw <= s_hi; ++sw) {
      const int d = t - 2 * sw;
      const int j_lo = d - (W - 1) > 0 ? d - (W - 1) : 0;
      const int j_hi = d < H - 1 ? d : H - 1;
      for (int j = j_lo + (int)threadIdx.x; j <= j_hi; j += GS_NT) {
        const int i = d - j;
        const size_t o = (size_t)j * W + i;
        const uint8_t f = flags[o];
        if (!(f & F_FLUID)) continue;
        const T ce = p[o];
        const T l = (f & F_LEFT) ? p[o - 1] : ce;
        const T r = (f & F_RIGHT) ? p[o + 1] : ce;
        const T dn = (f & F_DOWN) ? p[o - W] : ce;
        const T up = (f & F_UP) ? p[o + W] : ce;
        p[o] = (T)poisson_new(l, r, dn, up, div[o], alpha);
      }
    }
    __syncthreads();
  }
}

int gs_solve(void* p, const void* div, const uint8_t* fmu, const uint8_t* fmv, int H, int W,
             double rho, double h, double dt, int iterations, void* ws, int dtype,
             hipStream_t s) {
  SolverWs w = carve_ws(ws, H, W, dtype);
  int rc = smk_cell_flags(fmu, fmv, H, W, w.flags, s);
  if (rc || iterations <= 0) return rc;
  const double alpha = calc_alpha(h, rho, dt);
  hipEvent_t ev = prof_begin(s);
  if (dtype == SMK_F64)
    hipLaunchKernelGGL(k_gs_wavefront<double>, dim3(1), dim3(GS_NT), 0, s, (double*)p,
                       (const double*)div, w.flags, H, W, alpha, iterations, nullptr);
  else
    hipLaunchKernelGGL(k_gs_wavefront<float>, dim3(1), dim3(GS_NT), 0, s, (float*)p,
                       (const float*)div, w.flags, H, W, alpha, iterations, nullptr);
  prof_end(ev, "gauss_seidel", 3.0 * (dtype == SMK_F64 ? 8.0 : 4.0) * H * W * iterations, s);
  return check_launch("k_gs_wavefront");
}

// FP32 red-black SOR sweeps as the inner solver of FP64 defect correction:
// per chunk of `check_every` sweeps, r = alpha*div - A p in FP64 (with its
// norm), defect = r/alpha in FP32, e = the chunk's sweeps from zero on
// A e = alpha*defect in FP32, p += e in FP64.  A stationary iteration from
// zero on the residual equation is the same linear map as continuing it from
// p, so the sweep count to a tolerance matches the FP64 solver's.
int rbsor_solve_mixed(double* p, const double* div, const uint8_t* fmu, const uint8_t* fmv,
                      int H, int W, double rho, double h, double dt, int iterations,
                      double omega, double tol, int check_every, void* ws, smk_solve_info* info,
                      hipStream_t s) {
  SolverWs w = carve_ws(ws, H, W, SMK_F64);
  int rc = smk_cell_flags(fmu, fmv, H, W, w.flags, s);
  if (rc) return rc;
  const double alpha = calc_alpha(h, rho, dt);
  if (omega <= 0.0 || omega > 2.0) omega = optimal_omega(H > W ? H : W);
  if (check_every < 1) check_every = 1;
  const long n = (long)H * W;
  rc = rhs_norm2_dev<double>(div, w.flags, H, W, alpha, w.partial, w.result, 1, s);
  if (rc) return rc;
  int done = 0;
  double rel = -1.0;
  rc = sweeps_gated(w, iterations, check_every, tol, [&](int k, const int* gate) -> int {
    if (k == 0)
      return residual_partials_dev<double>(p, div, w.flags, H, W, alpha, w.partial, w.defect32,
                                           gate, s);
    int r = run_passes<float>(MODE_RB, w.e32[0], w.e32[1], w.e32[0], w.defect32, w.flags, H, W,
                              alpha, omega, 2 * k, 0, s, gate, true);
    if (r) return r;
    hipEvent_t ev = prof_begin(s);
    hipLaunchKernelGGL(k_add_correction<double>, dim3(cdiv(n, 256)), dim3(256), 0, s, p,
                       w.e32[0], n, gate);
    prof_end(ev, "correct_fp64", 20.0 * (double)n, s);
    return check_launch("k_add_correction");
  }, w.result + 1, norm_blocks(H, W), &done, &rel, s);
  if (rc) return rc;
  if (info) {
    info->iterations = done;
    info->converged = tol > 0.0 && rel <= tol;
    info->rel_residual = rel;
    info->inner_cycles = 0;
    info->refinements = (done + check_every - 1) / check_every;
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
  // device-side convergence (gated iterations); each check also writes the
  // state to a pinned ring, one slot per queued iteration
  smk::SolveState* d_state;
  smk::RingSlot* h_ring;
  smk::RingSlot* d_ring;  // device address of h_ring
  int small_from[2];  // first level handled by k_vcycle_small: [native dtype, float32]
  int small_mode[2];  // smoother mode small_from was computed for
  // float32 coarsest-level operator (SMK_COARSEST_OPERATOR) and what it was
  // built for
  float* cop = nullptr;
  double cop_alpha = 0.0, cop_omega = 0.0;
  int cop_npass = -1;
  // float32 exact operator (without alpha) of the first level of at most 1024
  // cells (SMK_COARSEST_EXACT_1024): its level (valid = false after a mask
  // change)
  float* copw = nullptr;
  int copw_level = -1;
  bool copw_valid = false;
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
  double coarsest_omega;
  const int* gate;  // device flag gating every launch (nullptr: ungated)
  int si, ci;
  int rediscretize;
  double rho, dt;
  const float* cop;  // float32 coarsest-level operator for the small kernel (or null)
  bool fast;         // float32 arithmetic in the float32 levels' tile passes
  // float32 V-cycles: the level whose exact operator copw (without alpha)
  // they apply in place of it and the levels below (-1: none), and that
  // level's alpha
  int exact_level;
  const float* copw;
  double exact_alpha;
};

inline double level_omega(const MgRun& run, const Level& L, bool coarsest = false) {
  if (run.mode != MODE_RB) return 1.0;
  const int n = L.H > L.W ? L.H : L.W;
  if (coarsest && run.coarsest_omega != 0.0)
    return run.coarsest_omega > 0.0 ? run.coarsest_omega : optimal_omega(n);
  return run.omega_override > 0.0 ? run.omega_override : optimal_omega(n);
}

// First level from which the small-level kernel takes the rest of the
// hierarchy: the first that fits one workgroup's LDS.  When tiled levels
// precede it anyway, the kernel starts no higher than SMALL_TOP_CELLS cells:
// it runs on one CU, where a 64x64 level costs more (12.8 us of smoothing,
// restriction and prolongation) than two tiled launches (~10 us); measured
// 1 % per 2048^2 step.  A hierarchy that fits entirely stays one launch.
constexpr long SMALL_TOP_CELLS = 32L * 32L;

template <typename T>
int mg_small_from(smk_mg* mg, int mode) {
  if (mode == MODE_GS) return -1;  // the lexicographic sweep has its own kernel
  const size_t lim = small_lds_limit<T>();
  int Hs[64], Ws[64];
  for (int l = 0; l < mg->nl; ++l) {
    Hs[l] = mg->lv[l].H;
    Ws[l] = mg->lv[l].W;
  }
  int from0 = -1;
  for (int from = 0; from < mg->nl && from0 < 0; ++from) {
    if (mg->nl - from > SMALL_MAX) continue;
    if (small_lds_bytes(Hs + from, Ws + from, mg->nl - from, sizeof(T), mode) <= lim) from0 = from;
  }
  if (from0 <= 0) return from0;  // -1: the coarsest level alone does not fit
  for (int from = from0; from < mg->nl; ++from)
    if ((long)Hs[from] * Ws[from] <= SMALL_TOP_CELLS) return from;
  return from0;
}

// Smooth level l with npass passes (tiled kernels), optionally prolonging a
// coarse correction in the first launch and restricting the residual or
// reducing the residual norm in the last.  Flips lv.cur per launch.
template <typename T>
int mg_smooth(std::vector<Level>& lv, int l, bool use32, int npass, bool zero_init,
              const T* coarse_e, int Hc, int Wc, T* coarse_rhs, double coarse_div,
              double* norm_partial, int* nparts, double alpha, double omega, int mode,
              hipStream_t s, const int* gate = nullptr, bool fast_inner = false) {
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
    // the register red-black variant restricts (lane pair, row pair) blocks
    if (mode == MODE_RB && last && coarse_rhs != nullptr &&
        pick_variant(sizeof(T) == 4, l == 0, L.H, L.W, a.R) == VAR_REG64)
      a.R += a.R & 1;
    a.zero_init = (first && zero_init) ? 1 : 0;
    a.fast_math = (use32 && fast_inner) ? 1 : 0;
    a.gate = gate;
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
      *nparts = tiled_blocks(mode, L.H, L.W, a.R, l == 0, sizeof(T) == 4);
    }
    int rc = launch_tiled<T>(mode, a, s, l == 0);
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
    a.omega[k] = level_omega(run, L, k == a.nl - 1);
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
  a.cop = (sizeof(T) == 4 && use32) ? run.cop : nullptr;
  a.gate = run.gate;
  const size_t lds = small_lds_bytes(a.H, a.W, a.nl, sizeof(T), run.mode);
  return launch_small<T>(run.mode, a, lds, s);
}

// solver.py:94-139.  norm_partial (level 0 only): reduce ||r||^2 of the
// result in the last post-smooth launch; *nparts receives the slot count.
template <typename T>
int mg_vcycle(smk_mg* mg, int l, bool use32, const MgRun& run, double h, bool zero_init,
              double* norm_partial, int* nparts, hipStream_t s);

// The float32 V-cycle's level l and everything below it in one launch:
// p32 of level l = the level's exact operator applied to its rhs32
// (k_apply_exact; the operator is the V-cycle's limit, so no smoothing of
// these levels is needed).
inline int mg_apply_exact(smk_mg* mg, int l, const MgRun& run, hipStream_t s) {
  Level& L = mg->lv[l];
  const int n = L.H * L.W;
  if (run.copw == nullptr || n > COARSE_EXACT_MAX)
    return fail(SMK_EINVAL, "exact level %d (%dx%d): no operator", l, L.H, L.W);
  float* x = reinterpret_cast<float*>(L.p32[L.cur32]);
  const float* b = reinterpret_cast<const float*>(L.rhs32);
  hipEvent_t ev = prof_begin(s);
  hipLaunchKernelGGL(k_apply_exact, dim3(cdiv(n, EX_ROWS)), dim3(EX_NT), 0, s, run.copw, b, x,
                     n, (float)run.exact_alpha, run.gate);
  prof_end(ev, "exact_apply_f32", 4.0 * (double)n * (double)n, s);
  return check_launch("k_apply_exact");
}

template <typename T>
__global__ void k_fill_gated(T* __restrict__ x, long n, T v, const int* __restrict__ gate) {
  if (gate != nullptr && *gate == 0) return;
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) x[k] = v;
}

// A V-cycle with the lexicographic Gauss-Seidel smoother (solver.py:107,137
// with smoother_type "gauss_seidel"): the sweeps run in place on the level's
// current p with the wavefront kernel, the residual-restriction and the
// prolong-correction are tile launches without smoothing passes (each moves
// p to the other buffer, two flips per level and cycle).  The level-0
// residual norm is computed by the caller (*nparts = -1).
template <typename T>
int mg_vcycle_gs(smk_mg* mg, int l, bool use32, const MgRun& run, double h, bool zero_init,
                 int* nparts, hipStream_t s) {
  std::vector<Level>& lv = mg->lv;
  Level& L = lv[l];
  const bool coarsest = (l == mg->nl - 1);
  const double alpha = calc_alpha(h, run.rho, run.dt);
  const long n = (long)L.H * L.W;
  auto sweep = [&](int iters) -> int {
    void** pb = use32 ? L.p32 : L.p;
    T* p = reinterpret_cast<T*>(pb[use32 ? L.cur32 : L.cur]);
    const T* rhs = reinterpret_cast<const T*>(use32 ? L.rhs32 : L.rhs);
    if (iters <= 0) return SMK_OK;
    hipEvent_t ev = prof_begin(s);
    hipLaunchKernelGGL(k_gs_wavefront<T>, dim3(1), dim3(GS_NT), 0, s, p, rhs, L.flags, L.H, L.W,
                       alpha, iters, run.gate);
    prof_end(ev, "gauss_seidel", 3.0 * sizeof(T) * (double)n * iters, s);
    return check_launch("k_gs_wavefront");
  };
  if (nparts != nullptr) *nparts = -1;
  if (zero_init) {
    void** pb = use32 ? L.p32 : L.p;
    hipLaunchKernelGGL(k_fill_gated<T>, dim3(cdiv(n, 256)), dim3(256), 0, s,
                       reinterpret_cast<T*>(pb[use32 ? L.cur32 : L.cur]), n, (T)0, run.gate);
    int rc = check_launch("k_fill_gated");
    if (rc) return rc;
  }
  int rc = sweep(coarsest ? run.ci : run.si);
  if (rc || coarsest) return rc;
  Level& C = lv[l + 1];
  T* crhs = reinterpret_cast<T*>(use32 ? C.rhs32 : C.rhs);
  const double coarse_div = run.rediscretize ? alpha : calc_alpha(h * 2.0, run.rho, run.dt);
  rc = mg_smooth<T>(lv, l, use32, 0, false, nullptr, 0, 0, crhs, coarse_div, nullptr, nullptr,
                    alpha, 1.0, MODE_RB, s, run.gate);
  if (rc) return rc;
  rc = mg_vcycle<T>(mg, l + 1, use32, run, h * 2.0, true, nullptr, nullptr, s);
  if (rc) return rc;
  const T* ce = reinterpret_cast<const T*>(use32 ? C.p32[C.cur32] : C.p[C.cur]);
  rc = mg_smooth<T>(lv, l, use32, 0, false, ce, C.H, C.W, nullptr, 1.0, nullptr, nullptr, alpha,
                    1.0, MODE_RB, s, run.gate);
  if (rc) return rc;
  return sweep(run.si);
}

template <typename T>
int mg_vcycle(smk_mg* mg, int l, bool use32, const MgRun& run, double h, bool zero_init,
              double* norm_partial, int* nparts, hipStream_t s) {
  std::vector<Level>& lv = mg->lv;
  if (run.mode == MODE_GS) return mg_vcycle_gs<T>(mg, l, use32, run, h, zero_init, nparts, s);
  if (use32 && l == run.exact_level) {  // this level and those below: one product
    if (norm_partial != nullptr) *nparts = -1;
    return mg_apply_exact(mg, l, run, s);
  }
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
  const double omega = level_omega(run, L, coarsest);
  if (coarsest)
    return mg_smooth<T>(lv, l, use32, npass, zero_init, nullptr, 0, 0, nullptr, 1.0,
                        norm_partial, nparts, alpha, omega, run.mode, s, run.gate, run.fast);
  Level& C = lv[l + 1];
  T* crhs = reinterpret_cast<T*>(use32 ? C.rhs32 : C.rhs);
  const double coarse_div = run.rediscretize ? alpha : calc_alpha(h * 2.0, run.rho, run.dt);
  int rc = mg_smooth<T>(lv, l, use32, npass, zero_init, nullptr, 0, 0, crhs, coarse_div, nullptr,
                        nullptr, alpha, omega, run.mode, s, run.gate, run.fast);
  if (rc) return rc;
  rc = mg_vcycle<T>(mg, l + 1, use32, run, h * 2.0, true, nullptr, nullptr, s);
  if (rc) return rc;
  const T* ce = reinterpret_cast<const T*>(use32 ? C.p32[C.cur32] : C.p[C.cur]);
  return mg_smooth<T>(lv, l, use32, npass, false, ce, C.H, C.W, nullptr, 1.0, norm_partial,
                      nparts, alpha, omega, run.mode, s, run.gate, run.fast);
}

inline MgRun make_run(const smk_mg_params* prm, double rho, double dt) {
  MgRun run{};
  run.mode = prm->smoother == SMK_SMOOTH_JACOBI
                 ? MODE_JACOBI
                 : (prm->smoother == SMK_SMOOTH_GS ? MODE_GS : MODE_RB);
  run.omega_override = prm->smoother_omega;
  run.coarsest_omega = prm->coarsest_omega;
  run.si = prm->smoother_iterations;
  run.ci = prm->coarsest_iterations;
  run.rediscretize = prm->coarse_scale == SMK_COARSE_REDISCRETIZED;
  run.rho = rho;
  run.dt = dt;
  run.exact_level = -1;
  run.copw = nullptr;
  run.exact_alpha = 0.0;
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
                       mg->result, 0, nullptr);
    rc = check_launch("k_sum_partials");
  }
  if (rc) return rc;
  SMK_CHECK_HIP(hipMemcpyAsync(host, mg->result, count * sizeof(double), hipMemcpyDeviceToHost, s));
  SMK_CHECK_HIP(hipStreamSynchronize(s));
  return SMK_OK;
}

// Queue gated iterations -- iter(k) enqueues iteration k, ending with a
// k_check on mg->d_state -- until the device reports the solve done.  The
// outcome of iteration k-1 is read (from a pinned ring slot) while iteration
// k is already queued, so the GPU never waits for the host; the one
// iteration queued after the last useful one is gated off and does nothing.
template <typename F>
int run_gated(smk_mg* mg, int max_iter, double tol, F&& iter, SolveState* fin, hipStream_t s) {
  return run_gated_ring(mg->d_state, mg->h_ring, mg->d_ring, max_iter, tol, iter,
                        fin, s);
}

// Gated iterations need the finest level's ping-pong flips per V-cycle to be
// even (one launch per pre- and post-smoothing, or none when the small-level
// kernel owns level 0), so that skipped V-cycles leave the host's record of
// which buffer holds the result where executed ones would.  Coarser levels
// restart from zero every cycle, so their buffer parity does not matter.
inline bool gateable(const smk_mg* mg, const MgRun& run, int slot) {
  if (mg->small_from[slot] == 0 || run.mode == MODE_GS) return true;
  const int npass = run.mode == MODE_RB ? 2 * run.si : run.si;
  return mg->nl > 1 && npass + 1 <= MAX_R;
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
  const long n = (long)mg->H * mg->W;
  // zero guess: the gated loop below never reads p before writing it (the
  // first V-cycle starts from zero, the initial check reuses the rhs norm's
  // partials, which equal the p = 0 residual's); the other loops start from
  // a zeroed p
  const bool zg = prm->zero_guess != 0;
  const bool gated = prm->tolerance > 0.0 && gateable(mg, run, 0);
  if (zg && !gated) {
    hipLaunchKernelGGL(k_fill_gated<T>, dim3(cdiv(n, 256)), dim3(256), 0, s, p, n, (T)0,
                       nullptr);
    rc = check_launch("k_fill_gated");
    if (rc) return rc;
  }
  if (gated) {
    // norm of the initial guess, then V-cycle + norm per gated iteration
    rc = rhs_norm2_dev<T>(div, L0.flags, mg->H, mg->W, alpha, mg->partial, mg->result, 1, s,
                          nullptr, mg->d_state);
    if (rc) return rc;
    MgRun grun = run;
    grun.gate = &mg->d_state->active;
    SolveState fin{};
    rc = run_gated(mg, prm->v_cycles + 1, prm->tolerance,
                   [&](int k, RingSlot* slot, unsigned sq) -> int {
      int nparts = -1;
      int r = SMK_OK;
      if (k > 0) r = mg_vcycle<T>(mg, 0, false, grun, h, zg && k == 1, mg->partial, &nparts, s);
      else if (zg) nparts = norm_blocks(mg->H, mg->W);
      if (!r && nparts < 0) {
        nparts = mg->npartial;
        r = residual_partials_dev<T>(reinterpret_cast<const T*>(L0.p[L0.cur]), div, L0.flags,
                                     mg->H, mg->W, alpha, mg->partial, nullptr, grun.gate, s);
      }
      if (r) return r;
      hipLaunchKernelGGL(k_check, dim3(1), dim3(1024), 0, s, mg->partial, nparts,
                         mg->result + 1, mg->d_state, prm->tolerance, prm->v_cycles, 1, slot,
                         sq);
      return check_launch("k_check");
    }, &fin, s);
    if (rc) return rc;
    cycles = fin.cycles;
    rel = fin.rel;
    converged = rel <= prm->tolerance;
    if (zg && cycles == 0) {  // no V-cycle ran: the result is the zero guess
      hipLaunchKernelGGL(k_fill_gated<T>, dim3(cdiv(n, 256)), dim3(256), 0, s, p, n, (T)0,
                         nullptr);
      rc = check_launch("k_fill_gated");
      if (rc) return rc;
    }
  } else if (prm->tolerance > 0.0) {
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

// The float32 operator of the coarsest level (built once per handle, mask
// state and alpha / omega / pass count): the red-black passes' matrix, or
// with `exact` the level's exact least-squares solution operator
// (coarse_exact_operator, built on the host from the level's flags; the
// pass count and omega do not enter).  Null when the coarsest level has
// more than 256 cells (or the smoother is not red-black).
inline int ensure_coarse_op(smk_mg* mg, const MgRun& run, double h, bool exact,
                            const float** out, hipStream_t s) {
  *out = nullptr;
  const Level& C = mg->lv[mg->nl - 1];
  if (run.mode != MODE_RB || C.H > 16 || C.W > 16) return SMK_OK;
  const double alpha = calc_alpha(h * (double)(1L << (mg->nl - 1)), run.rho, run.dt);
  const double omega = exact ? 0.0 : level_omega(run, C, true);
  const int npass = exact ? -2 : 2 * run.ci;
  if (mg->cop == nullptr) {
    int rc = mg_alloc(reinterpret_cast<void**>(&mg->cop), (size_t)256 * 256 * sizeof(float));
    if (rc) return rc;
    mg->cop_npass = -1;
  }
  if (mg->cop_npass != npass || mg->cop_alpha != alpha || mg->cop_omega != omega) {
    const int n = C.H * C.W;
    if (exact) {
      uint8_t fl[256];
      SMK_CHECK_HIP(hipMemcpyAsync(fl, C.flags, (size_t)n, hipMemcpyDeviceToHost, s));
      SMK_CHECK_HIP(hipStreamSynchronize(s));
      std::vector<float> op((size_t)n * n);
      SMK_REQUIRE(coarse_exact_operator(fl, C.H, C.W, alpha, op.data()) == 0,
                  "coarsest level %dx%d: no exact operator (inconsistent face flags)", C.H,
                  C.W);
      SMK_CHECK_HIP(hipMemcpyAsync(mg->cop, op.data(), op.size() * sizeof(float),
                                   hipMemcpyHostToDevice, s));
      SMK_CHECK_HIP(hipStreamSynchronize(s));
    } else {
      hipLaunchKernelGGL(k_coarse_op, dim3(n), dim3(64), 0, s, C.flags, C.H, C.W, alpha, omega,
                         npass, mg->cop);
      int rc = check_launch("k_coarse_op");
      if (rc) return rc;
    }
    mg->cop_alpha = alpha;
    mg->cop_omega = omega;
    mg->cop_npass = npass;
  }
  *out = mg->cop;
  return SMK_OK;
}

// The float32 exact operator K^+ (alpha applied by the product) of the first
// level of at most 1024 cells (SMK_COARSEST_EXACT_1024), built on the host
// once per handle and mask state; *level = -1 and *out = null when no level
// qualifies or the smoother is not red-black (the V-cycles then keep their
// coarsest sweeps).  *alpha = that level's alpha for this solve.
inline int ensure_wide_op(smk_mg* mg, const MgRun& run, double h, int* level, const float** out,
                          double* alpha_out, hipStream_t s) {
  *level = -1;
  *out = nullptr;
  if (run.mode != MODE_RB) return SMK_OK;
  int l = 0;
  while (l < mg->nl && (long)mg->lv[l].H * mg->lv[l].W > COARSE_EXACT_MAX) ++l;
  if (l == mg->nl) return SMK_OK;
  const Level& C = mg->lv[l];
  *alpha_out = calc_alpha(h * (double)(1L << l), run.rho, run.dt);
  if (mg->copw == nullptr) {
    int rc = mg_alloc(reinterpret_cast<void**>(&mg->copw),
                      (size_t)COARSE_EXACT_MAX * COARSE_EXACT_MAX * sizeof(float));
    if (rc) return rc;
    mg->copw_valid = false;
  }
  if (!mg->copw_valid || mg->copw_level != l) {
    const int n = C.H * C.W;
    std::vector<uint8_t> fl((size_t)n);
    SMK_CHECK_HIP(hipMemcpyAsync(fl.data(), C.flags, (size_t)n, hipMemcpyDeviceToHost, s));
    SMK_CHECK_HIP(hipStreamSynchronize(s));
    std::vector<float> op((size_t)n * n);
    SMK_REQUIRE(coarse_exact_operator(fl.data(), C.H, C.W, 1.0, op.data(), COARSE_EXACT_MAX) == 0,
                "level %dx%d: no exact operator (inconsistent face flags)", C.H, C.W);
    SMK_CHECK_HIP(hipMemcpyAsync(mg->copw, op.data(), op.size() * sizeof(float),
                                 hipMemcpyHostToDevice, s));
    SMK_CHECK_HIP(hipStreamSynchronize(s));
    mg->copw_level = l;
    mg->copw_valid = true;
  }
  *level = l;
  *out = mg->copw;
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
  MgRun run = make_run(prm, rho, dt);
  mg_prepare_small<float>(mg, 1, run.mode);
  if (prm->coarsest_solve == SMK_COARSEST_OPERATOR || prm->coarsest_solve == SMK_COARSEST_EXACT) {
    rc = ensure_coarse_op(mg, run, h, prm->coarsest_solve == SMK_COARSEST_EXACT, &run.cop, s);
    if (rc) return rc;
  } else if (prm->coarsest_solve == SMK_COARSEST_EXACT_1024) {
    rc = ensure_wide_op(mg, run, h, &run.exact_level, &run.copw, &run.exact_alpha, s);
    if (rc) return rc;
  }
  run.fast = prm->inner_arithmetic == SMK_INNER_F32;
  const double alpha = calc_alpha(h, rho, dt);
  const int H = mg->H, W = mg->W;
  const long n = (long)H * W;
  const int inner = prm->inner_cycles > 0 ? prm->inner_cycles : 1;
  const double tol = prm->tolerance;
  // zero guess: in the gated loop the rhs norm's kernel also writes the
  // defect of p = 0 and the first correction reads no p; the other loop
  // starts from a zeroed p
  const bool zg = prm->zero_guess != 0;
  const bool gated = gateable(mg, run, 1);
  float* const defect0 = reinterpret_cast<float*>(L0.rhs32);
  rc = rhs_norm2_dev<double>(div, L0.flags, H, W, alpha, mg->partial, mg->result, 1, s,
                             zg && gated ? defect0 : nullptr, gated ? mg->d_state : nullptr);
  if (rc) return rc;
  if (zg && !gated) {
    hipLaunchKernelGGL(k_fill_gated<double>, dim3(cdiv(n, 256)), dim3(256), 0, s, p, n, 0.0,
                       nullptr);
    rc = check_launch("k_fill_gated");
    if (rc) return rc;
  }
  int refinements = 0, inner_total = 0;
  double rel = -1.0, b2 = 0.0;
  double host[2];
  if (gated) {
    // iteration 0: defect of the initial p + check; iteration k > 0: inner
    // FP32 V-cycles, p += e, defect + check -- all gated on the device state
    MgRun grun = run;
    grun.gate = &mg->d_state->active;
    const int nparts = norm_blocks(H, W);
    // refinement k reads p from buf[(k-1) & 1] and writes buf[k & 1] (the
    // correction and the defect in one pass); refinements gated off write
    // nothing, so the result is in buf[executed refinements & 1]
    double* buf[2] = {p, reinterpret_cast<double*>(L0.p[1])};
    SolveState fin{};
    rc = run_gated(mg, prm->v_cycles + 1, prm->tolerance,
                   [&](int k, RingSlot* slot, unsigned sq) -> int {
      int r = SMK_OK;
      float* defect = reinterpret_cast<float*>(L0.rhs32);
      if (k > 0) {
        for (int c = 0; c < inner && !r; ++c)
          r = mg_vcycle<float>(mg, 0, true, grun, h, c == 0, nullptr, nullptr, s);
        if (r) return r;
        const float* e = reinterpret_cast<const float*>(L0.p32[L0.cur32]);
        hipEvent_t ev = prof_begin(s);
        const double* pin = (zg && k == 1) ? nullptr : buf[(k - 1) & 1];
        hipLaunchKernelGGL(k_correct_defect<double>, dim3(cdiv(W, NB_X), cdiv(H, NB_Y)),
                           dim3(NB_T), 0, s, pin, e, buf[k & 1], div, L0.flags, H, W, alpha,
                           mg->partial, defect, grun.gate);
        prof_end(ev, "correct_defect_fp64", 32.0 * (double)n, s);
        r = check_launch("k_correct_defect");
      } else if (!zg) {
        r = residual_partials_dev<double>(p, div, L0.flags, H, W, alpha, mg->partial, defect,
                                          grun.gate, s);
      }  // zero guess: the rhs norm's partials and defect stand for p = 0
      if (r) return r;
      hipLaunchKernelGGL(k_check, dim3(1), dim3(1024), 0, s, mg->partial, nparts,
                         mg->result + 1, mg->d_state, tol, prm->v_cycles, inner, slot, sq);
      return check_launch("k_check");
    }, &fin, s);
    if (rc) return rc;
    refinements = fin.cycles;
    inner_total = fin.cycles * inner;
    rel = fin.rel;
    if (refinements & 1)
      SMK_CHECK_HIP(hipMemcpyAsync(p, buf[1], (size_t)n * sizeof(double),
                                   hipMemcpyDeviceToDevice, s));
    if (zg && refinements == 0) {  // no refinement ran: the result is the zero guess
      hipLaunchKernelGGL(k_fill_gated<double>, dim3(cdiv(n, 256)), dim3(256), 0, s, p, n, 0.0,
                         nullptr);
      rc = check_launch("k_fill_gated");
      if (rc) return rc;
    }
  }
  while (!gateable(mg, run, 1)) {
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
    hipEvent_t ev = prof_begin(s);
    hipLaunchKernelGGL(k_add_correction<double>, dim3(cdiv(n, 256)), dim3(256), 0, s, p, e, n,
                       nullptr);
    prof_end(ev, "correct_fp64", 20.0 * (double)n, s);
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

// ------------------------------------------------- no-allocation Jacobi solve
// Scratch of smk_jacobi_solve_noalloc: the cell flags (1 B per cell) and the
// ping-pong partner of p (8 B per cell in float64) of a grid of at most
// SMK_NOALLOC_MAX_CELLS cells, 9 MiB in all, resident on every GPU the
// library runs on until the process ends.  Host code never passes the name
// `scratch` to a kernel (on the host it is a stand-in, not a device
// address): it uses the address hipGetSymbolAddress gives for the current
// device, looked up once per device.
constexpr size_t NOALLOC_FLAGS_BYTES = align256((size_t)SMK_NOALLOC_MAX_CELLS);
constexpr size_t NOALLOC_BYTES = NOALLOC_FLAGS_BYTES + align256((size_t)SMK_NOALLOC_MAX_CELLS * 8);
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::atomic<unsigned char*> g_scratch[NOALLOC_MAX_DEVICES];
std::mutex g_scratch_mu;

// The current device's copy of `scratch` (looked up on the device's first
// call, then reused).
int noalloc_scratch(unsigned char** out) {
  int dev = -1;
  SMK_CHECK_HIP(hipGetDevice(&dev));
  SMK_REQUIRE(dev >= 0 && dev < NOALLOC_MAX_DEVICES, "no-allocation scratch: device %d", dev);
  unsigned char* p = g_scratch[dev].load(std::memory_order_acquire);
  if (p == nullptr) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    p = g_scratch[dev].load(std::memory_order_relaxed);
    if (p == nullptr) {
      void* a = nullptr;
      const hipError_t e = hipGetSymbolAddress(&a, HIP_SYMBOL(scratch));
      if (e != hipSuccess || a == nullptr)
        return fail(e != hipSuccess ? (int)e : (int)hipErrorUnknown,
                    "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
      p = static_cast<unsigned char*>(a);
      g_scratch[dev].store(p, std::memory_order_release);
    }
  }
  *out = p;
  return SMK_OK;
}
}  // namespace

extern "C" {

int smk_jacobi_solve_noalloc(void* p, const void* div, const uint8_t* fmu, const uint8_t* fmv,
                             int H, int W, double rho, double h, double dt, int iterations,
                             int dtype, void* stream) {
  SMK_REQUIRE(p && div && fmu && fmv, "smk_jacobi_solve_noalloc: null pointer");
  SMK_REQUIRE(H > 0 && W > 0 && iterations >= 0, "smk_jacobi_solve_noalloc: bad arguments");
  SMK_REQUIRE(dtype == SMK_F64 || dtype == SMK_F32, "smk_jacobi_solve_noalloc: bad dtype %d",
              dtype);
  SMK_REQUIRE((long)H * W <= SMK_NOALLOC_MAX_CELLS,
              "smk_jacobi_solve_noalloc: %d x %d = %ld cells exceeds the static scratch of %ld "
              "cells; use smk_jacobi_solve with a workspace",
              H, W, (long)H * W, (long)SMK_NOALLOC_MAX_CELLS);
  unsigned char* base = nullptr;
  int rc = noalloc_scratch(&base);
  if (rc) return rc;
  hipStream_t s = as_stream(stream);
  uint8_t* flags = base;
  void* ptmp = base + NOALLOC_FLAGS_BYTES;
  rc = smk_cell_flags(fmu, fmv, H, W, flags, stream);
  if (rc) return rc;
  const double alpha = calc_alpha(h, rho, dt);
  if (dtype == SMK_F64)
    return run_passes<double>(MODE_JACOBI, (double*)p, (double*)ptmp, (const double*)p,
                              (const double*)div, flags, H, W, alpha, 1.0, iterations, 0, s);
  return run_passes<float>(MODE_JACOBI, (float*)p, (float*)ptmp, (const float*)p,
                           (const float*)div, flags, H, W, alpha, 1.0, iterations, 0, s);
}

int smk_profile_enable(int on) {
  g_prof_on = on != 0;
  g_prof_repeat = on > 1 ? (on < 64 ? on : 64) : 1;
  // warm the event pool so that profiled launches create none
  while (g_prof_on && g_prof_pool.size() < 4096) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) break;
    g_prof_pool.push_back(e);
  }
  return SMK_OK;
}

int smk_profile_read(smk_profile_record* out, int n, int reset) {
  // aggregate by tag in first-seen order
  std::vector<smk_profile_record> agg;
  for (ProfRec& r : g_prof) {
    hipError_t e = hipEventSynchronize(r.stop);
    if (e != hipSuccess) return fail((int)e, "smk_profile_read: %s", hipGetErrorString(e));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, r.start, r.stop);
    size_t k = 0;
    while (k < agg.size() && strncmp(agg[k].name, r.name, sizeof(agg[k].name)) != 0) ++k;
    if (k == agg.size()) {
      smk_profile_record z{};
      strncpy(z.name, r.name, sizeof(z.name) - 1);
      agg.push_back(z);
    }
    agg[k].launches += r.launches;
    agg[k].total_ms += ms;
    agg[k].alg_bytes += r.bytes;
  }
  if (reset) {
    for (ProfRec& r : g_prof) {
      g_prof_pool.push_back(r.start);
      g_prof_pool.push_back(r.stop);
    }
    g_prof.clear();
  }
  const int m = (int)agg.size() < n ? (int)agg.size() : n;
  for (int k = 0; k < m && out != nullptr; ++k) out[k] = agg[k];
  return (int)agg.size();
}

int smk_set_smoother_variant(int dtype, int which, int variant) {
  SMK_REQUIRE(dtype == SMK_F64 || dtype == SMK_F32, "smk_set_smoother_variant: bad dtype");
  SMK_REQUIRE(which >= 0 && which <= 3,
              "smk_set_smoother_variant: which is 0 (fine), 1 (coarse), 2 (small levels) or "
              "3 (coarse levels of at most 512x512 cells)");
  if (which == 2) {
    SMK_REQUIRE(variant >= 0 && variant < 8, "smk_set_smoother_variant: bad variant %d", variant);
    g_small_nt[dtype == SMK_F32] = SMALL_NT_CHOICES[variant];
    g_small_wave[dtype == SMK_F32] = SMALL_WAVE_CHOICES[variant];
    return SMK_OK;
  }
  SMK_REQUIRE(variant >= 0 && variant < NUM_VARIANTS, "smk_set_smoother_variant: bad variant %d",
              variant);
  g_variant[dtype == SMK_F32][which == 3 ? 2 : which] = variant;
  return SMK_OK;
}

int smk_get_smoother_variant(int dtype, int which) {
  if ((dtype != SMK_F64 && dtype != SMK_F32) || which < 0 || which > 3) return -1;
  if (which == 2) {
    const int nt = g_small_nt[dtype == SMK_F32];
    const int wave = g_small_wave[dtype == SMK_F32];
    for (int v = 0; v < 8; ++v)
      if (SMALL_NT_CHOICES[v] == nt && SMALL_WAVE_CHOICES[v] == wave) return v;
    return -1;
  }
  return g_variant[dtype == SMK_F32][which == 3 ? 2 : which];
}

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

int smk_gs_solve(void* p, const void* div, const uint8_t* fmu, const uint8_t* fmv, int H, int W,
                 double rho, double h, double dt, int iterations, void* ws, size_t ws_bytes,
                 int dtype, void* stream) {
  SMK_REQUIRE(p && div && fmu && fmv && ws, "smk_gs_solve: null pointer");
  SMK_REQUIRE(H > 0 && W > 0 && iterations >= 0, "smk_gs_solve: bad arguments");
  SMK_REQUIRE(dtype == SMK_F64 || dtype == SMK_F32, "smk_gs_solve: bad dtype %d", dtype);
  SMK_REQUIRE(ws_bytes >= solver_ws_bytes(H, W, dtype),
              "smk_gs_solve: workspace too small (%zu < %zu)", ws_bytes,
              solver_ws_bytes(H, W, dtype));
  return gs_solve(p, div, fmu, fmv, H, W, rho, h, dt, iterations, ws, dtype, as_stream(stream));
}

int smk_rbsor_solve_mixed(void* p, const void* div, const uint8_t* fmu, const uint8_t* fmv,
                          int H, int W, double rho, double h, double dt, int iterations,
                          double omega, double tol, int check_every, void* ws, size_t ws_bytes,
                          smk_solve_info* info, void* stream) {
  SMK_REQUIRE(p && div && fmu && fmv && ws, "smk_rbsor_solve_mixed: null pointer");
  SMK_REQUIRE(H > 0 && W > 0 && iterations >= 0, "smk_rbsor_solve_mixed: bad arguments");
  SMK_REQUIRE(ws_bytes >= solver_ws_bytes(H, W, SMK_F64),
              "smk_rbsor_solve_mixed: workspace too small (%zu < %zu)", ws_bytes,
              solver_ws_bytes(H, W, SMK_F64));
  return rbsor_solve_mixed((double*)p, (const double*)div, fmu, fmv, H, W, rho, h, dt,
                           iterations, omega, tol, check_every, ws, info, as_stream(stream));
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

// The masks of every level from the level-0 face masks (solver.py:34-73):
// level-0 flags, solid = ~is_fluid (solver.py:38-40), then per coarse level
// the coarsened solid, its face masks and flags.  In place, stream-ordered.
static int mg_build_masks(smk_mg* mg, const uint8_t* fmu, const uint8_t* fmv, hipStream_t s) {
  int rc = smk_cell_flags(fmu, fmv, mg->H, mg->W, mg->lv[0].flags, s);
  if (!rc) {
    const long n = (long)mg->H * mg->W;
    hipLaunchKernelGGL(k_solid_from_flags, dim3(cdiv(n, 256)), dim3(256), 0, s,
                       mg->lv[0].flags, n, mg->lv[0].solid);
    rc = check_launch("k_solid_from_flags");
  }
  for (int l = 1; l < mg->nl && !rc; ++l) {
    Level& F = mg->lv[l - 1];
    Level& C = mg->lv[l];
    rc = smk_coarsen_solid(F.solid, F.H, F.W, C.solid, s);
    if (!rc) rc = smk_face_masks_from_solid(C.solid, C.H, C.W, C.fmu, C.fmv, s);
    if (!rc) rc = smk_cell_flags(C.fmu, C.fmv, C.H, C.W, C.flags, s);
  }
  mg->cop_npass = -1;  // the coarsest operator follows the coarsest flags
  mg->copw_valid = false;
  return rc;
}

int smk_mg_refresh_masks(smk_mg* mg, const uint8_t* fmu, const uint8_t* fmv, int all_levels,
                         void* stream) {
  SMK_REQUIRE(mg && fmu && fmv, "smk_mg_refresh_masks: null pointer");
  hipStream_t s = as_stream(stream);
  if (all_levels) return mg_build_masks(mg, fmu, fmv, s);
  if (mg->nl == 1) mg->cop_npass = -1;  // the coarsest level is level 0
  if (mg->copw_level == 0) mg->copw_valid = false;
  return refresh_cell_flags(fmu, fmv, mg->H, mg->W, mg->lv[0].flags, s);
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
  (void)hipFree(mg->d_state);
  (void)hipFree(mg->cop);
  (void)hipFree(mg->copw);
  if (mg->h_ring) (void)hipHostFree(mg->h_ring);
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
  if (!rc) rc = mg_alloc((void**)&mg->d_state, sizeof(SolveState));
  if (!rc) {
    hipError_t e = hipHostMalloc((void**)&mg->h_ring, 4 * sizeof(RingSlot), hipHostMallocDefault);
    if (e == hipSuccess) {
      memset(mg->h_ring, 0, 4 * sizeof(RingSlot));
      e = hipHostGetDevicePointer((void**)&mg->d_ring, mg->h_ring, 0);
    }
    if (e != hipSuccess) rc = fail((int)e, "smk_mg_create: %s", hipGetErrorString(e));
  }
  if (!rc) rc = mg_build_masks(mg, fmu, fmv, s);
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
  SMK_REQUIRE(prm->smoother == SMK_SMOOTH_JACOBI || prm->smoother == SMK_SMOOTH_RBSOR ||
                  prm->smoother == SMK_SMOOTH_GS,
              "smk_mg_solve: smoother must be SMK_SMOOTH_JACOBI, SMK_SMOOTH_GS or "
              "SMK_SMOOTH_RBSOR");
  SMK_REQUIRE(prm->v_cycles >= 0 && prm->smoother_iterations >= 0 &&
                  prm->coarsest_iterations >= 0,
              "smk_mg_solve: negative iteration count");
  SMK_REQUIRE(prm->coarse_scale == SMK_COARSE_REFERENCE ||
                  prm->coarse_scale == SMK_COARSE_REDISCRETIZED,
              "smk_mg_solve: bad coarse_scale");
  SMK_REQUIRE(prm->coarsest_solve == SMK_COARSEST_SWEEPS ||
                  prm->coarsest_solve == SMK_COARSEST_OPERATOR ||
                  prm->coarsest_solve == SMK_COARSEST_EXACT ||
                  prm->coarsest_solve == SMK_COARSEST_EXACT_1024,
              "smk_mg_solve: bad coarsest_solve %d", prm->coarsest_solve);
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

// batched scenes (multigrid side and step driver)
#include "smk_batch.inl"
