// smk_coarse_exact.hpp -- host code: the exact solution operator of a
// multigrid level (at most 1024 cells), applied by the FP32 V-cycles of a
// mixed-precision solve as one matrix-vector product in place of that level
// and the levels below it.
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

#include "smk_flags.hpp"

namespace smk {

// The coarsest level's red-black passes relax each fluid cell i toward
//   p_i = (v_left + v_right + v_down + v_up + alpha b_i) / 4,
// v = the neighbour's p across an open face (flag bits F_LEFT .. F_UP) and
// p_i across a closed one, and leave non-fluid cells at p = 0.  Their fixed
// points solve K p = alpha b over the fluid cells, with
//   K_ii = the number of open faces of cell i,
//   K_ij = -1 for an open face between fluid cells i and j
// (an open face to a non-fluid cell couples to its fixed p = 0).  K is
// symmetric positive semi-definite; its null space is spanned by the
// indicator vectors of the connected fluid components that have no open face
// to a non-fluid cell (the pure-Neumann parts of the level), along which the
// passes drift rather than converge when b has a component there.
//
// coarse_exact_operator writes x = alpha K^+ b, the least-squares solution of
// least norm (the passes' limit with its null-space component removed), as
// out[j * n + i] = the value in cell i for b = unit vector j (n = H * W, rows
// and columns of non-fluid cells zero).  K^+ = (K + P)^-1 - P, P the
// orthogonal projector on the null space: K + P is positive definite and is
// inverted by its Cholesky factor in double precision.  Returns 0, or -1
// when n > max_cells (at most COARSE_EXACT_MAX), the face bits of two
// neighbours disagree, or the factorization meets a non-positive pivot.
constexpr int COARSE_EXACT_MAX = 1024;
inline int coarse_exact_operator(const uint8_t* flags, int H, int W, double alpha, float* out,
                                 int max_cells = 256) {
  const int n = H * W;
  if (H <= 0 || W <= 0 || n > max_cells || n > COARSE_EXACT_MAX) return -1;
  std::vector<int> idx(n, -1), cell;
  for (int c = 0; c < n; ++c)
    if (flags[c] & F_FLUID) {
      idx[c] = (int)cell.size();
      cell.push_back(c);
    }
  const int m = (int)cell.size();
  for (int k = 0; k < n * n; ++k) out[k] = 0.0f;
  if (m == 0) return 0;

  // K over the fluid cells; leak[k] = open faces of cell k to non-fluid cells
  std::vector<double> K((size_t)m * m, 0.0);
  std::vector<int> leak(m, 0);
  const uint8_t bits[4] = {F_LEFT, F_RIGHT, F_DOWN, F_UP};
  const uint8_t back[4] = {F_RIGHT, F_LEFT, F_UP, F_DOWN};
  const int di[4] = {-1, 1, 0, 0}, dj[4] = {0, 0, -1, 1};
  for (int k = 0; k < m; ++k) {
    const int c = cell[k], i = c % W, j = c / W;
    for (int d = 0; d < 4; ++d) {
      if (!(flags[c] & bits[d])) continue;
      const int ii = i + di[d], jj = j + dj[d];
      if (ii < 0 || ii >= W || jj < 0 || jj >= H) return -1;
      const int nb = jj * W + ii;
      K[(size_t)k * m + k] += 1.0;
      if (idx[nb] < 0) {
        ++leak[k];
        continue;
      }
      if (!(flags[nb] & back[d])) return -1;  // one-sided face
      K[(size_t)k * m + idx[nb]] -= 1.0;
    }
  }

  // connected components (open faces between fluid cells); a component
  // without leak spans one null-space direction
  std::vector<int> comp(m, -1), stack;
  std::vector<int> size, leaky;
  for (int s = 0; s < m; ++s) {
    if (comp[s] >= 0) continue;
    const int id = (int)size.size();
    size.push_back(0);
    leaky.push_back(0);
    comp[s] = id;
    stack.push_back(s);
    while (!stack.empty()) {
      const int k = stack.back();
      stack.pop_back();
      ++size[id];
      leaky[id] |= leak[k] > 0;
      for (int q = 0; q < m; ++q)
        if (q != k && K[(size_t)k * m + q] != 0.0 && comp[q] < 0) {
          comp[q] = id;
          stack.push_back(q);
        }
    }
  }
  auto proj = [&](int a, int b) -> double {  // P_ab
    const int c = comp[a];
    return (comp[b] == c && !leaky[c]) ? 1.0 / (double)size[c] : 0.0;
  };
  for (int a = 0; a < m; ++a)
    for (int b = 0; b < m; ++b) K[(size_t)a * m + b] += proj(a, b);

  // K + P = L L^T (L in the lower triangle of K), then Z = L^-1 and
  // (K + P)^-1 = Z^T Z
  for (int c = 0; c < m; ++c) {
    double d = K[(size_t)c * m + c];
    for (int k = 0; k < c; ++k) d -= K[(size_t)c * m + k] * K[(size_t)c * m + k];
    if (!(d > 0.0)) return -1;
    const double lc = std::sqrt(d);
    K[(size_t)c * m + c] = lc;
    for (int r = c + 1; r < m; ++r) {
      double v = K[(size_t)r * m + c];
      for (int k = 0; k < c; ++k) v -= K[(size_t)r * m + k] * K[(size_t)c * m + k];
      K[(size_t)r * m + c] = v / lc;
    }
  }
  // Zt[c * m + r] = (L^-1)[r][c]: column c of L^-1 stored contiguously, so
  // both loops below run along memory (the sums are taken in the same order
  // k = c .. r-1 and k = b .. m-1 either way)
  std::vector<double> Zt((size_t)m * m, 0.0);
  for (int c = 0; c < m; ++c) {  // column c of L^-1: forward substitution
    double* z = Zt.data() + (size_t)c * m;
    z[c] = 1.0 / K[(size_t)c * m + c];
    for (int r = c + 1; r < m; ++r) {
      double v = 0.0;
      const double* kr = K.data() + (size_t)r * m;
      for (int k = c; k < r; ++k) v -= kr[k] * z[k];
      z[r] = v / K[(size_t)r * m + r];
    }
  }
  for (int a = 0; a < m; ++a)
    for (int b = a; b < m; ++b) {
      double v = 0.0;
      const double* za = Zt.data() + (size_t)a * m;
      const double* zb = Zt.data() + (size_t)b * m;
      for (int k = b; k < m; ++k) v += za[k] * zb[k];
      const float x = (float)(alpha * (v - proj(a, b)));
      out[(size_t)cell[b] * n + cell[a]] = x;
      out[(size_t)cell[a] * n + cell[b]] = x;
    }
  return 0;
}

}  // namespace smk
