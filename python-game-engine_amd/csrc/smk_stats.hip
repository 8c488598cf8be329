// smk_stats.hip -- field reductions of the profiling metrics on the device.
//
// profiling/config.py:268-301 defines six reducers (rms, mean, mean_abs,
// max, max_abs, std -- numpy semantics, std with ddof 0) applied over three
// scopes of a 2-D field (profiling/metrics.py:7-28): "all", "interior"
// ([1:-1, 1:-1]) and "boundary" (the rest).  smk_field_stats computes all 18
// numbers without copying the field to the host:
//   pass 1  per-block partial n, sum, sum of squares, sum of |x|, max, max|x|
//           for the interior and the boundary (grid-stride, fixed order)
//   pass 2  one workgroup combines them (all = interior + boundary), forms the
//           means and writes them for
//   pass 3  per-block partial sums of (x - mean_scope)^2 (two-pass std, as
//           numpy computes it)
//   pass 4  one workgroup writes the 18 results.
// Deterministic: fixed per-thread order, fixed reduction tree.
#include <cfloat>
#include <cmath>

#include "smk_common.hpp"

namespace smk {
namespace {

constexpr int ST_NT = 256;
constexpr int ST_MAXB = 512;  // partial slots (grid size cap)
constexpr int NQ = 6;         // n, sum, sumsq, sumabs, max, maxabs

template <typename T>
__global__ __launch_bounds__(ST_NT) void k_stats_pass1(const T* __restrict__ f, int rows, int cols,
                                                       double* __restrict__ part) {
  __shared__ double red[ST_NT / 64];
  double q[2][NQ];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    q[s][0] = q[s][1] = q[s][2] = q[s][3] = 0.0;
    q[s][4] = -INFINITY;
    q[s][5] = -INFINITY;
  }
  const long n = (long)rows * cols;
  for (long k = (long)blockIdx.x * ST_NT + threadIdx.x; k < n; k += (long)gridDim.x * ST_NT) {
    const int j = (int)(k / cols), i = (int)(k - (long)j * cols);
    const int s = (j >= 1 && j <= rows - 2 && i >= 1 && i <= cols - 2) ? 0 : 1;
    const double x = (double)f[k];
    const double a = fabs(x);
    q[s][0] += 1.0;
    q[s][1] += x;
    q[s][2] += x * x;
    q[s][3] += a;
    q[s][4] = fmax(q[s][4], x);
    q[s][5] = fmax(q[s][5], a);
  }
  for (int s = 0; s < 2; ++s)
    for (int t = 0; t < NQ; ++t) {
      double v = q[s][t];
      if (t < 4) {
        v = block_sum<ST_NT>(v, red);
      } else {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
        __syncthreads();
        if (threadIdx.x == 0)
          for (int w = 1; w < ST_NT / 64; ++w) v = fmax(v, red[w]);
        __syncthreads();
      }
      if (threadIdx.x == 0) part[((size_t)blockIdx.x * 2 + s) * NQ + t] = v;
    }
}

// combine the partials; means[3] = scope means (all, interior, boundary),
// tot[3][NQ] kept for the final pass
__global__ __launch_bounds__(ST_NT) void k_stats_pass2(const double* __restrict__ part, int nb,
                                                       double* __restrict__ tot,
                                                       double* __restrict__ means) {
  if (threadIdx.x != 0) return;
  double q[2][NQ];
  for (int s = 0; s < 2; ++s) {
    q[s][0] = q[s][1] = q[s][2] = q[s][3] = 0.0;
    q[s][4] = q[s][5] = -INFINITY;
    for (int b = 0; b < nb; ++b) {
      const double* p = part + ((size_t)b * 2 + s) * NQ;
      q[s][0] += p[0];
      q[s][1] += p[1];
      q[s][2] += p[2];
      q[s][3] += p[3];
      q[s][4] = fmax(q[s][4], p[4]);
      q[s][5] = fmax(q[s][5], p[5]);
    }
  }
  // scope order: all, interior, boundary
  for (int t = 0; t < NQ; ++t) {
    tot[1 * NQ + t] = q[0][t];
    tot[2 * NQ + t] = q[1][t];
    tot[0 * NQ + t] = t < 4 ? q[0][t] + q[1][t] : fmax(q[0][t], q[1][t]);
  }
  for (int s = 0; s < 3; ++s) {
    const double cnt = tot[s * NQ + 0];
    means[s] = cnt > 0.0 ? tot[s * NQ + 1] / cnt : NAN;
  }
}

template <typename T>
__global__ __launch_bounds__(ST_NT) void k_stats_pass3(const T* __restrict__ f, int rows, int cols,
                                                       const double* __restrict__ means,
                                                       double* __restrict__ part) {
  __shared__ double red[ST_NT / 64];
  double d[2] = {0.0, 0.0};
  const double m_in = means[1], m_bd = means[2];
  const long n = (long)rows * cols;
  for (long k = (long)blockIdx.x * ST_NT + threadIdx.x; k < n; k += (long)gridDim.x * ST_NT) {
    const int j = (int)(k / cols), i = (int)(k - (long)j * cols);
    const bool in = j >= 1 && j <= rows - 2 && i >= 1 && i <= cols - 2;
    const double x = (double)f[k];
    // deviations from the interior / boundary means, and from the overall
    // mean (kept in slot 2 of the partial)
    const double e = x - (in ? m_in : m_bd);
    d[in ? 0 : 1] += e * e;
  }
  for (int s = 0; s < 2; ++s) {
    const double v = block_sum<ST_NT>(d[s], red);
    if (threadIdx.x == 0) part[(size_t)blockIdx.x * 4 + s] = v;
  }
  // overall-mean deviations in a separate sweep of the same elements
  double da = 0.0;
  const double m_all = means[0];
  for (long k = (long)blockIdx.x * ST_NT + threadIdx.x; k < n; k += (long)gridDim.x * ST_NT) {
    const double e = (double)f[k] - m_all;
    da += e * e;
  }
  const double v = block_sum<ST_NT>(da, red);
  if (threadIdx.x == 0) part[(size_t)blockIdx.x * 4 + 2] = v;
}

// out[scope * 6 + reducer], scopes (all, interior, boundary), reducers
// (rms, mean, mean_abs, max, max_abs, std)
__global__ void k_stats_pass4(const double* __restrict__ tot, const double* __restrict__ part,
                              int nb, double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double dev[3] = {0.0, 0.0, 0.0};
  for (int b = 0; b < nb; ++b) {
    dev[1] += part[(size_t)b * 4 + 0];
    dev[2] += part[(size_t)b * 4 + 1];
    dev[0] += part[(size_t)b * 4 + 2];
  }
  for (int s = 0; s < 3; ++s) {
    const double* q = tot + s * NQ;
    const double cnt = q[0];
    double* o = out + s * 6;
    if (cnt <= 0.0) {
      for (int r = 0; r < 6; ++r) o[r] = NAN;
      continue;
    }
    o[0] = sqrt(q[2] / cnt);
    o[1] = q[1] / cnt;
    o[2] = q[3] / cnt;
    o[3] = q[4];
    o[4] = q[5];
    o[5] = sqrt(dev[s] / cnt);
  }
}

template <typename T>
int field_stats(const T* f, int rows, int cols, double* out, double* ws, hipStream_t s) {
  const long n = (long)rows * cols;
  int nb = (int)((n + ST_NT - 1) / ST_NT);
  if (nb > ST_MAXB) nb = ST_MAXB;
  if (nb < 1) nb = 1;
  double* part = ws;                        // ST_MAXB * 2 * NQ
  double* tot = ws + ST_MAXB * 2 * NQ;      // 3 * NQ
  double* means = tot + 3 * NQ;             // 3
  double* part3 = means + 4;                // ST_MAXB * 4
  hipLaunchKernelGGL(k_stats_pass1<T>, dim3(nb), dim3(ST_NT), 0, s, f, rows, cols, part);
  int rc = check_launch("k_stats_pass1");
  if (rc) return rc;
  hipLaunchKernelGGL(k_stats_pass2, dim3(1), dim3(64), 0, s, part, nb, tot, means);
  rc = check_launch("k_stats_pass2");
  if (rc) return rc;
  hipLaunchKernelGGL(k_stats_pass3<T>, dim3(nb), dim3(ST_NT), 0, s, f, rows, cols, means, part3);
  rc = check_launch("k_stats_pass3");
  if (rc) return rc;
  hipLaunchKernelGGL(k_stats_pass4, dim3(1), dim3(64), 0, s, tot, part3, nb, out);
  return check_launch("k_stats_pass4");
}

}  // namespace
}  // namespace smk

using namespace smk;

extern "C" {

size_t smk_field_stats_workspace_bytes(void) {
  return sizeof(double) * ((size_t)ST_MAXB * 2 * NQ + 3 * NQ + 4 + (size_t)ST_MAXB * 4);
}

int smk_field_stats(const void* field, int rows, int cols, void* out18, void* workspace,
                    size_t workspace_bytes, int dtype, void* stream) {
  SMK_REQUIRE(field && out18 && workspace, "smk_field_stats: null pointer");
  SMK_REQUIRE(rows > 0 && cols > 0, "smk_field_stats: bad shape %dx%d", rows, cols);
  SMK_REQUIRE(workspace_bytes >= smk_field_stats_workspace_bytes(),
              "smk_field_stats: workspace too small");
  hipStream_t s = as_stream(stream);
  double* ws = static_cast<double*>(workspace);
  double* out = static_cast<double*>(out18);
  if (dtype == SMK_F64) return field_stats<double>((const double*)field, rows, cols, out, ws, s);
  if (dtype == SMK_F32) return field_stats<float>((const float*)field, rows, cols, out, ws, s);
  return fail(SMK_EINVAL, "smk_field_stats: dtype must be SMK_F64 or SMK_F32, got %d", dtype);
}

}  // extern "C"
