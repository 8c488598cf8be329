// smk_batch.hpp -- many independent smoke scenes of uneven sizes advanced
// together: the per-scene descriptor the batched kernels read, and the
// transport launchers (defined in smk_ops.hip) the batch driver (in
// smk_solve.hip) calls.
//
// Work is described by packed job lists: a job is one 64 x 4 tile of one
// scene (int4 {scene, tile x, tile y, -}); every batched kernel is
// persistent -- a fixed grid of workgroups walks the job list -- so one
// launch serves all scenes whatever their sizes.
#pragma once

#include "smk_common.hpp"

namespace smk {

constexpr int BT_X = 64, BT_Y = 4, BT_NT = BT_X * BT_Y;  // transport / norm tile

struct BatchScene {
  int H, W;
  double h;                    // cell size
  double alpha0;               // -h^2 rho / dt (level-0 Poisson scale)
  double* rho[2];              // density ping-pong, current = rho[*parity]
  double* u;                   // u faces H x (W+1) (state)
  double* u_tmp;               // advected u
  double* v;                   // v faces (H+1) x W (state)
  double* v_tmp;               // advected v
  double* p;                   // pressure (float64)
  double* div;                 // divergence
  const uint8_t* fmu;
  const uint8_t* fmv;
  const uint8_t* fluid;
  const uint8_t* solid;
  const uint8_t* flags0;       // level-0 cell flags of the multigrid hierarchy
  float* defect;               // level-0 FP32 right-hand side of the correction solve
  const float* e;              // level-0 FP32 correction (result of the V-cycle)
  int box_x0, box_y0, box_x1, box_y1;
  double box_value;
  int part_base;               // first norm-partial slot of the scene
  int nbx, nby;                // norm tiles (cdiv(W, 64) x cdiv(H, 4))
};

struct BatchPhys {
  double dt;
  double gravity;
  double buoyancy;             // temperature coefficient (temperature is 0)
  double density_coefficient;
  double ambient_temperature;
  double fluid_density;
};

// transport launchers (persistent kernels over face jobs: tiles of the
// (H+1) x (W+1) face superset; cell jobs: tiles of H x W)
int batch_box_forces(const BatchScene* sc, const int4* jobs, int njobs, const int* parity,
                     const BatchPhys& ph, int blocks, hipStream_t s);
int batch_advect_velocity(const BatchScene* sc, const int4* jobs, int njobs, const BatchPhys& ph,
                          int blocks, hipStream_t s);
int batch_advect_density(const BatchScene* sc, const int4* jobs, int njobs, const int* parity,
                         const BatchPhys& ph, int blocks, hipStream_t s);
int batch_divergence_rhs_norm(const BatchScene* sc, const int4* cell_jobs, int njobs,
                              double* partial, int blocks, hipStream_t s);
int batch_gradient(const BatchScene* sc, const int4* jobs, int njobs, int* parity,
                   const BatchPhys& ph, int blocks, hipStream_t s);

}  // namespace smk
