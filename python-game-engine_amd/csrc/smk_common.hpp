// smk_common.hpp -- shared device helpers and host error plumbing for
// libsmoke_hip.so (gfx950).
//
// Layout conventions (same as the reference, src/cfd/common/simulation.py:69-99):
//   cell fields p, div, density, temperature : H x W
//   u faces : H x (W+1)      v faces : (H+1) x W
//   row-major, index [j*W + i]; j is y (row), i is x (column)
//   masks : uint8 0/1 (torch.bool storage)
//
// All device arithmetic is compiled with -ffp-contract=off so that every
// expression rounds exactly where the reference's expression does.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>

#include "smk_flags.hpp"
#include "smoke_hip.h"

namespace smk {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
int fail(int code, const char* fmt, ...);
int check_launch(const char* what);

#define SMK_CHECK_HIP(expr)                                                   \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess)                                                     \
      return ::smk::fail((int)e_, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

#define SMK_REQUIRE(cond, ...)                                                \
  do {                                                                        \
    if (!(cond)) return ::smk::fail(SMK_EINVAL, __VA_ARGS__);                 \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }

// Re-derive F_FLUID and the neighbour bits of flags from the face masks,
// keeping each cell's F_LFLUID (stream-ordered, no allocation).
int refresh_cell_flags(const uint8_t* fmu, const uint8_t* fmv, int H, int W, uint8_t* flags,
                       hipStream_t s);

// --------------------------------------------------------- sampling
// advection.py:6-49: floor/ceil, clamp to [0, n-1], t = x - x0 (clamped x0),
// lerp(a, b, t) = a + (b - a) * t.  Coordinates are float64 whatever the
// field type (the reference's coordinates are Python floats).  The first two
// lerps take two field values: (b - a) is formed in the field's type and
// only then promoted by the float64 weight (Numba's typing of
// lerp(float32, float32, float64)); the last lerp is float64 throughout.
__device__ __forceinline__ int clamp_index(double f, int hi) {
  if (!(f > 0.0)) return 0;  // also maps NaN to 0
  if (f > (double)hi) return hi;
  return (int)f;
}

// clamp_index(ceil(x), hi) from x0 = clamp_index(floor(x), hi).  The floor
// itself is not needed: x <= 0 and NaN clamp to 0 either way, and for x > 0
// the truncation in clamp_index is the floor.  Above x0 the ceiling is
// x0 + 1 unless x is a whole number; both forms agree for every double
// (including the infinities and NaN) and clamp bound.
__device__ __forceinline__ int clamp_ceil(double x, int x0, int hi) {
  const int c = x0 + ((double)x0 < x ? 1 : 0);
  return c < hi ? c : hi;
}

// clamp_index of a whole number held as an int
__device__ __forceinline__ int clamp_whole(int v, int hi) { return v > 0 ? (v < hi ? v : hi) : 0; }

// The lerps of sample_field from the four corner values (cYX: row Y,
// column X) and the weights tx = x - x0, ty = y - y0.
template <typename T>
__device__ __forceinline__ double lerp4(T c00, T c10, T c01, T c11, double tx, double ty) {
  const double c0 = (double)c00 + (double)(T)(c10 - c00) * tx;
  const double c1 = (double)c01 + (double)(T)(c11 - c01) * tx;
  return c0 + (c1 - c0) * ty;
}

// The lerps of sample_field once the four clamped indices are known.
template <typename T>
__device__ __forceinline__ double bilerp(const T* __restrict__ f, int fw, int x0, int x1, int y0,
                                         int y1, double x, double y) {
  const T c00 = f[(size_t)y0 * fw + x0];
  const T c10 = f[(size_t)y0 * fw + x1];
  const T c01 = f[(size_t)y1 * fw + x0];
  const T c11 = f[(size_t)y1 * fw + x1];
  return lerp4(c00, c10, c01, c11, x - (double)x0, y - (double)y0);
}

template <typename T>
__device__ __forceinline__ double sample_field(const T* __restrict__ f, int fh,
                                               int fw, double x, double y) {
  const int x0 = clamp_index(x, fw - 1);
  const int y0 = clamp_index(y, fh - 1);
  return bilerp(f, fw, x0, clamp_ceil(x, x0, fw - 1), y0, clamp_ceil(y, y0, fh - 1), x, y);
}

// Two consecutive elements of a row, read as one load (the hardware takes
// 8- and 16-byte global loads at the element's own alignment).
template <typename T>
struct __attribute__((aligned(sizeof(T)))) Pair {
  T a, b;
};
template <typename T>
__device__ __forceinline__ Pair<T> load_pair(const T* __restrict__ p) {
  Pair<T> r;
  __builtin_memcpy(&r, p, sizeof(r));
  return r;
}

// sample_field for a field in global memory, each row's two corners read by
// one load.  x1 is x0 or x0 + 1, so the pair starting at
// xs = min(x0, fw - 2) holds both; the corner values, weights and lerps are
// sample_field's.  A one-column field reads its corners singly.
template <typename T>
__device__ __forceinline__ double sample_field_rows(const T* __restrict__ f, int fh, int fw,
                                                    double x, double y) {
  const int x0 = clamp_index(x, fw - 1);
  const int y0 = clamp_index(y, fh - 1);
  const int x1 = clamp_ceil(x, x0, fw - 1), y1 = clamp_ceil(y, y0, fh - 1);
  if (fw < 2) return bilerp(f, fw, x0, x1, y0, y1, x, y);
  const int xs = x0 < fw - 2 ? x0 : fw - 2;
  const Pair<T> r0 = load_pair(f + (size_t)y0 * fw + xs);
  const Pair<T> r1 = load_pair(f + (size_t)y1 * fw + xs);
  const bool a0 = x0 == xs, a1 = x1 == xs;
  return lerp4(a0 ? r0.a : r0.b, a1 ? r0.a : r0.b, a0 ? r1.a : r1.b, a1 ? r1.a : r1.b,
               x - (double)x0, y - (double)y0);
}

// sample_field at a point whose floor and ceiling are known whole numbers
// (fx, cx, fy, cy): the staggered-grid averages of the advection step, whose
// positions are a cell centre or face plus a multiple of one half.  Same
// indices, weights and lerps as sample_field, without the float64 rounding
// and comparisons that find them.
template <typename T>
__device__ __forceinline__ double sample_known(const T* __restrict__ f, int fh, int fw, double x,
                                               double y, int fx, int cx, int fy, int cy) {
  return bilerp(f, fw, clamp_whole(fx, fw - 1), clamp_whole(cx, fw - 1), clamp_whole(fy, fh - 1),
                clamp_whole(cy, fh - 1), x, y);
}

// --------------------------------------------------------- lane moves
// Apply a 32-bit lane operation to every dword of a value.
template <typename T, typename F>
__device__ __forceinline__ T per_dword(T v, F f) {
  static_assert(sizeof(T) % 4 == 0, "dword-sized values only");
  uint32_t w[sizeof(T) / 4];
  __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) w[i] = f(w[i]);
  T r;
  __builtin_memcpy(&r, w, sizeof(T));
  return r;
}

// Lane l receives lane l-1's value, lane 0 keeps its own (DPP wave_shr:1;
// the result of __shfl_up(v, 1, 64) without a trip through the LDS
// crossbar).  All 64 lanes must be active.
template <typename T>
__device__ __forceinline__ T from_left(T v) {
  return per_dword(v, [](uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0x138, 0xf, 0xf, false);
  });
}
// Lane l receives lane l+1's value, lane 63 keeps its own (DPP wave_shl:1).
template <typename T>
__device__ __forceinline__ T from_right(T v) {
  return per_dword(v, [](uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0x130, 0xf, 0xf, false);
  });
}

// --------------------------------------------------------- XCD-aware order
// Workgroups are handed to the chip's 8 XCDs round-robin by launch order, so
// blocks id and id+1 run on different XCDs with separate L2s.  remap_block
// turns the launch id into a tile id such that each XCD walks a contiguous
// run of tiles (rows of tiles for a row-major grid): vertically adjacent
// tiles, which share halo rows, then meet in the same L2.  Bijective on
// [0, nb); the nb % 8 tail keeps its ids.
constexpr int NUM_XCD = 8;

__device__ __forceinline__ int remap_block(int id, int nb) {
  const int full = nb - nb % NUM_XCD;
  if (id >= full) return id;
  return (id % NUM_XCD) * (full / NUM_XCD) + id / NUM_XCD;
}

// (bx, by) of the tile a 2-D launch's workgroup handles, in XCD order.
__device__ __forceinline__ void xcd_tile(int& bx, int& by) {
  const int gx = (int)gridDim.x;
  const int v = remap_block((int)(blockIdx.y * gridDim.x + blockIdx.x), gx * (int)gridDim.y);
  by = v / gx;
  bx = v - by * gx;
}

// --------------------------------------------------------- reductions
// Wave64 sum with DPP-backed cross-lane moves (__shfl_xor lowers to
// ds_swizzle / DPP on gfx950).  Fixed order => deterministic.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block sum over NT threads (NT multiple of 64); result valid in thread 0.
template <int NT, typename T>
__device__ __forceinline__ T block_sum(T v, T* scratch /* >= NT/64 */) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) scratch[wid] = v;
  __syncthreads();
  T r = 0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) r += scratch[w];
  }
  __syncthreads();
  return r;
}

}  // namespace smk
