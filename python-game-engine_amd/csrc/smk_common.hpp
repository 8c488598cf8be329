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

template <typename T>
__device__ __forceinline__ double sample_field(const T* __restrict__ f, int fh,
                                               int fw, double x, double y) {
  const int x0 = clamp_index(floor(x), fw - 1);
  const int x1 = clamp_index(ceil(x), fw - 1);
  const int y0 = clamp_index(floor(y), fh - 1);
  const int y1 = clamp_index(ceil(y), fh - 1);
  const double tx = x - (double)x0;
  const double ty = y - (double)y0;
  const T c00 = f[(size_t)y0 * fw + x0];
  const T c10 = f[(size_t)y0 * fw + x1];
  const T c01 = f[(size_t)y1 * fw + x0];
  const T c11 = f[(size_t)y1 * fw + x1];
  const double c0 = (double)c00 + (double)(T)(c10 - c00) * tx;
  const double c1 = (double)c01 + (double)(T)(c11 - c01) * tx;
  return c0 + (c1 - c0) * ty;
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
