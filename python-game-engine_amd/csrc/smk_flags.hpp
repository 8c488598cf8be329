// smk_flags.hpp -- the per-cell flag byte of the solvers (host and device).
#pragma once

#include <cstdint>

namespace smk {

// --------------------------------------------------------- cell flags
// One byte per cell replaces the three mask reads of the Poisson stencil
// (residual.py:35-38 / jacobi.py:18-21).  A neighbour bit is set only when
// the reference would read a *different* cell: an open face whose clamped
// neighbour index equals the cell itself reads the cell's own value, exactly
// like a closed face.
//
// Two fluid bits: F_FLUID is the is_fluid a smoother derives from the face
// masks it is handed (jacobi.py:46, red_black_sor_gauss_seidel.py:38),
// F_LFLUID the is_fluid a multigrid level holds (solver.py:38-47), which the
// residual, the restriction, the prolongation and the residual norms read.
// Both are set alike except on a cached multigrid level 0 whose face-mask
// arrays changed in place since the hierarchy was built: the reference
// keeps those arrays by reference and is_fluid by value, so its level-0
// smoother sees the new masks and its level-0 residual the old is_fluid.
enum : uint8_t {
  F_FLUID = 1u,
  F_LEFT = 2u,
  F_RIGHT = 4u,
  F_DOWN = 8u,
  F_UP = 16u,
  F_LFLUID = 32u,
};

}  // namespace smk
