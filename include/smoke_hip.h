/*
 * smoke_hip.h -- C ABI of libsmoke_hip.so, the gfx950 smoke-step library.
 *
 * Each entry point replaces one (operation, backend) implementation that the
 * reference registers behind its operator / solver registries
 * (src/cfd/operators/registrations.py:21-29, src/cfd/solvers/registrations.py:17-20);
 * the comment above each names the reference kernel it replaces
 * (paths relative to the reference root, LCEBurton/python-game-engine).
 *
 * Conventions
 *  - Pointers are device pointers (HBM) owned by the caller; the library never
 *    frees caller memory.  Cell fields are H x W, u faces H x (W+1), v faces
 *    (H+1) x W, row-major.  Masks are uint8 0/1 (torch.bool storage).
 *  - dtype: SMK_F64 or SMK_F32; every floating-point pointer of one call has
 *    that type.
 *  - stream: a hipStream_t (NULL = default stream).  Calls only enqueue work,
 *    except where a comment says the call reads a result back (solver
 *    tolerance checks, solve info).
 *  - Return: 0 on success, < 0 invalid argument (SMK_EINVAL), > 0 a
 *    hipError_t.  smk_last_error() describes the last failure of the calling
 *    thread.
 *  - Kernels are deterministic: the same inputs give bit-identical outputs.
 */
#ifndef SMOKE_HIP_H
#define SMOKE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SMK_F64 0
#define SMK_F32 1

#define SMK_OK 0
#define SMK_EINVAL (-1)

/* Thread-local description of the last error (never NULL). */
const char* smk_last_error(void);
/* ABI version, bumped on incompatible change. */
int smk_abi_version(void);

/* ---------------------------------------------------------------- masks */

/* simulation.py:106-108 / :200-202 and multigrid/masks.py:30-49:
 * face_mask_u[:,1:-1] = ~(solid[:,:-1] | solid[:,1:]), outer faces 0;
 * likewise v.  Writes every element of fmu and fmv. */
int smk_face_masks_from_solid(const uint8_t* solid, int H, int W, uint8_t* fmu,
                              uint8_t* fmv, void* stream);
/* multigrid/masks.py:52-74 (== jacobi.py:46): is_fluid = OR of 4 faces. */
int smk_is_fluid(const uint8_t* fmu, const uint8_t* fmv, int H, int W,
                 uint8_t* is_fluid, void* stream);
/* Packed per-cell stencil flags used by the solvers (no reference
 * counterpart: one byte replaces the three masks read by residual.py:35-38). */
int smk_cell_flags(const uint8_t* fmu, const uint8_t* fmv, int H, int W,
                   uint8_t* flags, void* stream);
/* multigrid/masks.py:3-27: coarse solid = any solid in the 2x2 block. */
int smk_coarsen_solid(const uint8_t* solid_fine, int H, int W,
                      uint8_t* solid_coarse, void* stream);

/* ------------------------------------------------------------ advection */

/* backends/numba/operators/advection.py:112-128 advect_scalar_field
 * (operator key "advection_scalar").  u is uh x uw, v is vh x vw (the
 * samplers clamp to each array's own shape, advection.py:24).  dst =
 * sample * scale; scale = 1 is the reference operator, scale != 1 fuses the
 * density dissipation of simulation.py:232. */
int smk_advect_scalar(const void* src, void* dst, int H, int W, const void* u,
                      int uh, int uw, const void* v, int vh, int vw,
                      const uint8_t* is_fluid, double dt, double cell_size,
                      double scale, int dtype, void* stream);
/* advection.py:130-154 advect_velocity_u ("advection_velocity_u"). */
int smk_advect_u(const void* u_src, void* u_dst, int uh, int uw, const void* v,
                 int vh, int vw, const uint8_t* fmu, double dt,
                 double cell_size, int dtype, void* stream);
/* advection.py:156-180 advect_velocity_v ("advection_velocity_v"). */
int smk_advect_v(const void* v_src, void* v_dst, int vh, int vw, const void* u,
                 int uh, int uw, const uint8_t* fmv, double dt,
                 double cell_size, int dtype, void* stream);
/* Both of the above in one launch on MAC-shaped arrays, as
 * simulation.py:280-286 calls them (both read the old velocity). */
int smk_advect_velocity(const void* u_src, void* u_dst, const void* v_src,
                        void* v_dst, const uint8_t* fmu, const uint8_t* fmv,
                        int H, int W, double dt, double cell_size, int dtype,
                        void* stream);

/* --------------------------------------------------------------- forces */

/* forces.py:78-105 apply_forces ("apply_forces"), including the float32
 * force scratch of forces.py:33,62 and the float32 sum of forces.py:105.
 * v is vh x vw, density H x W, temperature Ht x Wt or NULL (all zero). */
int smk_apply_forces(void* v, int vh, int vw, const void* density, int H, int W,
                     const void* temperature, int Ht, int Wt, double dt,
                     double gravity, double buoyancy_coefficient,
                     double density_coefficient, double ambient_temperature,
                     int dtype, void* stream);
/* forces.py:108-123 cool_temperature ("cool_temperature"); decay is
 * exp(-cooling_rate*dt), evaluated once by the caller as forces.py:120 does. */
int smk_cool_temperature(void* T, int H, int W, double ambient, double decay,
                         int dtype, void* stream);
/* simulation.py:232 density *= dissipation (x[k] *= factor, n elements). */
int smk_scale(void* x, long n, double factor, int dtype, void* stream);

/* ----------------------------------------------------------- projection */

/* divergence.py:4-25 compute_divergence ("compute_divergence"). */
int smk_divergence(const void* u, const void* v, void* div, const uint8_t* fmu,
                   const uint8_t* fmv, int H, int W, double cell_size,
                   int dtype, void* stream);
/* pressure_gradient.py:4-37 apply_pressure_gradient
 * ("apply_pressure_gradient").  If solid != NULL the boundary enforcement of
 * boundaries.py:7-39 is applied to the same faces in the same launch
 * (simulation.py:244-249 always runs it right after). */
int smk_pressure_gradient(void* u, void* v, const void* p, const uint8_t* fmu,
                          const uint8_t* fmv, const uint8_t* solid, int H,
                          int W, double fluid_density, double cell_size,
                          double dt, int dtype, void* stream);
/* boundaries.py:7-39 enforce_boundary_conditions
 * ("enforce_boundary_conditions"). */
int smk_enforce_boundaries(void* u, void* v, const uint8_t* solid, int H, int W,
                           int dtype, void* stream);
/* residual.py:6-42 compute_residual ("compute_residual"). */
int smk_residual(const void* p, const void* div, const uint8_t* is_fluid,
                 const uint8_t* fmu, const uint8_t* fmv, int H, int W,
                 double alpha, double reciprocal_beta, void* r, int dtype,
                 void* stream);
/* The step's metrics pass, simulation.py:251-265: smk_divergence of (u, v)
 * into div, then smk_residual of p against that div into r, in one launch
 * (each cell's residual reads only its own divergence); both outputs equal
 * the two calls', bit for bit. */
int smk_divergence_residual(const void* u, const void* v, const void* p,
                            const uint8_t* is_fluid, const uint8_t* fmu,
                            const uint8_t* fmv, int H, int W, double cell_size,
                            double alpha, double reciprocal_beta, void* div,
                            void* r, int dtype, void* stream);

/* ------------------------------------------------------------- emitters */

/* emitters.py:157-178 (Emitter2D._apply_scalar_field) with the falloffs of
 * emitters.py:339-411; approach 0 hard_cutoff, 1 linear, 2 quadratic,
 * 3 gaussian.  Only the bounding box of the disc is visited (cells outside
 * it get falloff 0, as in emitters.py:52-56). */
int smk_emit_scalar(void* field, int H, int W, double value, double dt,
                    double pos_x, double pos_y, double radius, int approach,
                    int dtype, void* stream);
/* Box source (no reference counterpart; the plume workloads need it):
 * field[y0:y1, x0:x1] = value (mode 0) or += value (mode 1). */
int smk_box_source(void* field, int H, int W, int x0, int y0, int x1, int y1,
                   double value, int mode, int dtype, void* stream);
/* n box sources on one field, applied in list order (no reference
 * counterpart; the plume workloads' emitters): rects[4k .. 4k+3] = x0, y0,
 * x1, y1 of box k, values[k], modes[k] as in smk_box_source.  When the
 * clamped boxes are pairwise disjoint (and n <= 8) one launch applies all of
 * them (each cell is written by at most one box, so the order cannot
 * matter); otherwise they are applied one launch after another. */
int smk_box_sources(void* field, int H, int W, int n, const int* rects,
                    const double* values, const int* modes, int dtype, void* stream);

/* -------------------------------------------------------------- solvers */

typedef struct smk_solve_info {
  int iterations;      /* sweeps (Jacobi / RB-SOR) or V-cycles executed   */
  int converged;       /* 1 if the tolerance was met (tolerance runs)     */
  double rel_residual; /* ||r|| / ||alpha*div|| over fluid cells, or -1   */
  int inner_cycles;    /* mixed precision: FP32 V-cycles in total         */
  int refinements;     /* mixed precision: FP64 defect-correction steps   */
} smk_solve_info;

/* Device workspace (bytes) the single-grid solvers need for an H x W grid. */
size_t smk_solver_workspace_bytes(int H, int W, int dtype);

/* jacobi.py:27-52 jacobi_pressure_solver (solver key "jacobi"):
 * `iterations` sweeps, non-fluid cells copied, result in p. */
int smk_jacobi_solve(void* p, const void* div, const uint8_t* fmu,
                     const uint8_t* fmv, int H, int W, double fluid_density,
                     double cell_size, double dt, int iterations, void* ws,
                     size_t ws_bytes, int dtype, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): smk_jacobi_solve without the workspace argument, same
 * kernels and results.  Its scratch (cell flags and the p ping-pong buffer)
 * is a static device array of SMK_NOALLOC_MAX_CELLS cells' worth, 9 MiB,
 * held on each GPU the library runs on until the process ends; grids of
 * more cells are rejected with SMK_EINVAL.  The call allocates no device or
 * pinned memory and only enqueues work on `stream` (no synchronisation),
 * which must belong to the current device.  Every call on one GPU
 * shares that GPU's scratch: calls that overlap on one GPU (on different
 * streams, or from different threads) race, so order them. */
#define SMK_NOALLOC_MAX_CELLS (1L << 20)
int smk_jacobi_solve_noalloc(void* p, const void* div, const uint8_t* fmu,
                             const uint8_t* fmv, int H, int W, double fluid_density,
                             double cell_size, double dt, int iterations, int dtype,
                             void* stream);

/* red_black_sor_gauss_seidel.py:30-42 (solver key "rb_sor_gauss_seidel"):
 * omega <= 0 or > 2 picks calculateOptimalOmega(max(H,W)) (common.py:30-39).
 * tol <= 0: exactly `iterations` sweeps (the reference).  tol > 0: sweeps in
 * blocks of check_every until ||r||/||alpha*div|| <= tol or `iterations`
 * sweeps (new; reads the norm back once per block).  info may be NULL. */
int smk_rbsor_solve(void* p, const void* div, const uint8_t* fmu,
                    const uint8_t* fmv, int H, int W, double fluid_density,
                    double cell_size, double dt, int iterations, double omega,
                    double tol, int check_every, void* ws, size_t ws_bytes,
                    smk_solve_info* info, int dtype, void* stream);

/* gauss_seidel.py:8-48 gauss_seidel_pressure_solver (solver key
 * "gauss_seidel"): `iterations` lexicographic sweeps (row by row, left to
 * right, in place), run as a wavefront on one workgroup; identical to the
 * sequential sweep.  Workspace as for the other single-grid solvers. */
int smk_gs_solve(void* p, const void* div, const uint8_t* fmu,
                 const uint8_t* fmv, int H, int W, double rho, double h,
                 double dt, int iterations, void* ws, size_t ws_bytes,
                 int dtype, void* stream);
/* The same solver with FP32 sweeps inside FP64 defect correction (no
 * reference counterpart): p and div are float64; every check_every sweeps
 * the FP64 residual (and its norm) is formed, the sweeps run in FP32 from
 * zero on the residual equation, and p += e in FP64.  Same tolerance rule
 * and sweep accounting as smk_rbsor_solve; workspace from
 * smk_solver_workspace_bytes(H, W, SMK_F64). */
int smk_rbsor_solve_mixed(void* p, const void* div, const uint8_t* fmu,
                          const uint8_t* fmv, int H, int W, double rho,
                          double h, double dt, int iterations, double omega,
                          double tol, int check_every, void* ws,
                          size_t ws_bytes, smk_solve_info* info, void* stream);
/* ||r||^2 over fluid cells of the residual of residual.py:6-42 with
 * reciprocal_beta = 1/4, reduced on the device in a fixed order and read
 * back into *norm2 (synchronises the stream). */
int smk_residual_norm2(const void* p, const void* div, const uint8_t* fmu,
                       const uint8_t* fmv, int H, int W, double alpha, void* ws,
                       size_t ws_bytes, double* norm2, int dtype, void* stream);

/* ------------------------------------------------------------ multigrid */

typedef struct smk_mg smk_mg;

#define SMK_SMOOTH_JACOBI 0
#define SMK_SMOOTH_GS 1 /* lexicographic Gauss-Seidel (a wavefront on the GPU) */
#define SMK_SMOOTH_RBSOR 2

#define SMK_COARSE_REFERENCE 0    /* restrict(r)/alpha(2h), solver.py:124-125 */
#define SMK_COARSE_REDISCRETIZED 1 /* restrict(r)/alpha(h) */

#define SMK_PREC_NATIVE 0 /* V-cycles in the handle's dtype               */
#define SMK_PREC_MIXED 1  /* FP32 V-cycles inside FP64 defect correction  */

#define SMK_INNER_F64 0 /* mixed: the FP32 V-cycles evaluate the reference's
                           float64 expression, as a float32 field's solve does */
#define SMK_INNER_F32 1 /* mixed: their tile passes and residuals in float32
                           arithmetic (the inner solve need only converge)   */

#define SMK_COARSEST_SWEEPS 0   /* coarsest level: coarsest_iterations sweeps */
#define SMK_COARSEST_OPERATOR 1 /* mixed precision, red-black, a coarsest level
                                   of at most 256 cells: the FP32 V-cycles apply
                                   those sweeps as their precomputed matrix
                                   (the sweeps from a zero guess are linear in
                                   the rhs); other solves use the sweeps     */
#define SMK_COARSEST_EXACT 2    /* as SMK_COARSEST_OPERATOR, but the matrix is
                                   the coarsest level's exact least-squares
                                   solution operator (the limit of the sweeps,
                                   minus its drift along the level's constant
                                   null space), built once on the host;
                                   coarsest_iterations and coarsest_omega do
                                   not enter (new, mixed precision only)     */
#define SMK_COARSEST_EXACT_1024 3 /* as SMK_COARSEST_EXACT, for the first
                                   level of at most 1024 cells (32x32 on a
                                   power-of-two grid): the FP32 V-cycles stop
                                   at that level and apply its exact
                                   operator, a 1024x1024 float32 matrix, with
                                   one multi-workgroup product; the levels
                                   below it are not visited (new, mixed
                                   precision, red-black only)                */

typedef struct smk_mg_params {
  int v_cycles;            /* fixed cycles (tolerance <= 0) or the cap    */
  int smoother;            /* SMK_SMOOTH_*  (MultigridParams.smoother_type) */
  int smoother_iterations; /* sweeps per pre/post smooth                  */
  int coarsest_iterations; /* sweeps on the coarsest level                */
  double smoother_omega;   /* <= 0: the RB-SOR smoother's automatic omega
                              per level (reference); > 0 forces it        */
  int coarse_scale;        /* SMK_COARSE_*                                 */
  double tolerance;        /* > 0: stop at ||r||/||alpha*div|| <= tolerance */
  int precision;           /* SMK_PREC_*                                   */
  int inner_cycles;        /* mixed: FP32 V-cycles per refinement          */
  double coarsest_omega;   /* coarsest level only: 0 = as smoother_omega
                              (reference); > 0 forces it; < 0 the optimal
                              SOR omega of the coarsest level's size (new) */
  int coarsest_solve;      /* SMK_COARSEST_* (new; 0 = the reference's)    */
  int inner_arithmetic;    /* SMK_INNER_* (new; mixed precision only)       */
  int zero_guess;          /* 1: p's content is ignored and taken as 0 (the
                              caller's p.zero_() of simulation.py:310 folded
                              in; every cell of p is written); 0: p is the
                              initial guess (reference)                    */
} smk_mg_params;

/* multigrid/solver.py:34-92 (_build_levels, _build_buffers, _ensure_cache):
 * builds the level hierarchy for these face masks (halving while both sizes
 * are even, up to num_levels) and allocates every level buffer. */
int smk_mg_create(const uint8_t* fmu, const uint8_t* fmv, int H, int W,
                  int num_levels, int dtype, void* stream, smk_mg** out);
/* multigrid/solver.py:34-47, :87-92: the mask side of the reference's level
 * cache, in place and without allocation.  all_levels != 0: every level's
 * masks are derived again from fmu/fmv (what _build_levels does on a rebuild,
 * rebuild_masks=True or a level count short of num_levels).  all_levels == 0:
 * level 0 follows face-mask arrays that changed in place -- the reference
 * keeps level 0's face-mask arrays by reference, so its level-0 smoother
 * sees their current content, while level 0's is_fluid (read by the
 * residual, restriction and prolongation) and the coarse levels stay as
 * built.  Stream-ordered on `stream`. */
int smk_mg_refresh_masks(smk_mg* mg, const uint8_t* fmu, const uint8_t* fmv,
                         int all_levels, void* stream);
/* Number of levels actually built. */
int smk_mg_num_levels(const smk_mg* mg);
/* multigrid/solver.py:141-176 MultigridSolver.__call__ (solver key
 * "multigrid"): p is read as the initial guess and overwritten with the
 * result; div is the right-hand side.  Tolerance and mixed-precision runs
 * read the residual norm back once per cycle. */
int smk_mg_solve(smk_mg* mg, void* p, const void* div, double fluid_density,
                 double cell_size, double dt, const smk_mg_params* params,
                 smk_solve_info* info, void* stream);
int smk_mg_destroy(smk_mg* mg);

/* ------------------------------------------- differentiable transport */
/* No reference counterpart (the reference has no gradients): forward pieces
 * whose controls live in device memory, and the adjoints (transposed
 * Jacobians) of the transport operators of SmokeSimulation2D.substep
 * (src/cfd/common/simulation.py:206-249).  u is H x (W+1), v (H+1) x W.
 * The projection's adjoint needs no entry point: it is the forward
 * divergence / solve / gradient applied after smk_enforce_boundaries.
 *
 * Adjoint modes for the transposed bilinear gather of advection:
 *   SMK_ADJ_GATHER   deterministic: each source visits the targets that can
 *                    reach it (window = max backtrace displacement + 1)
 *   SMK_ADJ_SCATTER  each target adds its weights with atomics
 * Where a backtrace lies exactly on a grid line the slope used is the mean
 * of the one-sided slopes (what a central difference measures). */
#define SMK_ADJ_GATHER 0
#define SMK_ADJ_SCATTER 1

/* Bytes of scratch the two advection adjoints need for an H x W grid. */
size_t smk_adjoint_workspace_bytes(int H, int W, int dtype);

/* field[y0:y1, x0:x1] = values[index] (a device array): the box source with
 * a strength that is a control. */
int smk_box_set_indexed(void* field, int H, int W, int x0, int y0, int x1,
                        int y1, const void* values, int index, int dtype,
                        void* stream);
/* Adjoint of smk_box_set_indexed: values_bar[index] += sum of
 * field_bar over the box, then field_bar[box] = 0. */
int smk_box_set_adjoint(void* field_bar, int H, int W, int x0, int y0, int x1,
                        int y1, void* values_bar, int index, int dtype,
                        void* stream);

/* forces.py:19-46 buoyancy (temperature term 0) with the force kept in the
 * field type instead of the reference's float32 scratch (forces.py:33), so
 * that it is differentiable: v[j] += dt * (-dc * rho_face(j) * g), j < H. */
int smk_apply_buoyancy(void* v, const void* rho, int H, int W, double dt,
                       double gravity, double density_coefficient, int dtype,
                       void* stream);
/* rho_bar += (d v / d rho)^T v_bar for smk_apply_buoyancy. */
int smk_apply_buoyancy_adjoint(const void* v_bar, void* rho_bar, int H, int W,
                               double dt, double gravity,
                               double density_coefficient, int dtype,
                               void* stream);

/* Adjoint of smk_advect_scalar (advection.py:112-128, scale 1) for the
 * gradient dst_bar of its output: src_bar = (d dst / d src)^T dst_bar
 * (overwritten); when u_bar / v_bar are given, the backtrace sensitivity
 * (d dst / d u, d v)^T dst_bar is added to them. */
int smk_advect_scalar_adjoint(const void* dst_bar, const void* src,
                              const void* u, const void* v,
                              const uint8_t* is_fluid, int H, int W,
                              double dt, double cell_size, void* src_bar,
                              void* u_bar, void* v_bar, int mode,
                              void* workspace, size_t workspace_bytes,
                              int dtype, void* stream);
/* Adjoint of smk_advect_velocity (advection.py:130-180: u and v both
 * advected by the old velocity (us, vs)): us_bar, vs_bar (overwritten) =
 * (d (ud, vd) / d (us, vs))^T (ud_bar, vd_bar). */
int smk_advect_velocity_adjoint(const void* ud_bar, const void* vd_bar,
                                const void* us, const void* vs,
                                const uint8_t* fmu, const uint8_t* fmv, int H,
                                int W, double dt, double cell_size,
                                void* us_bar, void* vs_bar, int mode,
                                void* workspace, size_t workspace_bytes,
                                int dtype, void* stream);

/* ------------------------------------------------------ batched scenes */
/* No reference counterpart (the reference steps one simulation at a time):
 * many independent plume scenes of uneven sizes advanced together.  Each
 * step is SmokeSimulation2D.substep (simulation.py:186-249) per scene with
 * a box source (density := box_value), buoyancy (forces.py:78-105, no
 * temperature), RK1 advection, the divergence, a multigrid solve of FP32
 * V-cycles inside FP64 defect correction to `tolerance` (per scene), and
 * the pressure gradient with boundaries; the per-scene results are the
 * ones the single-scene entry points give.  All scenes' tiles go through
 * one persistent launch per stage; with use_graphs the stages are replayed
 * from HIP graphs.  Fields are float64 and owned by the caller. */
typedef struct smk_batch_scene {
  int H, W;
  double cell_size;
  int num_levels;              /* multigrid levels (halving while even)   */
  void* density[2];            /* ping-pong; density[0] is current first */
  void* u;                     /* H x (W+1) state                         */
  void* u_tmp;                 /* H x (W+1) scratch (advected u)          */
  void* v;                     /* (H+1) x W state                         */
  void* v_tmp;                 /* (H+1) x W scratch                       */
  void* pressure;              /* H x W                                   */
  void* divergence;            /* H x W                                   */
  const uint8_t* face_mask_u;
  const uint8_t* face_mask_v;
  const uint8_t* is_fluid;
  const uint8_t* solid;        /* boundary enforcement mask               */
  int box_x0, box_y0, box_x1, box_y1;
  double box_value;
} smk_batch_scene;

typedef struct smk_batch_params {
  double dt, gravity, buoyancy, density_coefficient, ambient_temperature;
  double fluid_density;
  int smoother_iterations;     /* V(s,s), 1..3                            */
  int coarsest_iterations;
  int coarse_scale;            /* SMK_COARSE_*                            */
  double coarsest_omega;       /* as smk_mg_params.coarsest_omega         */
  double tolerance;            /* relative residual per scene             */
  int max_cycles;              /* V-cycles per scene per step             */
  int use_graphs;              /* replay stages from HIP graphs           */
  int tile_rows;               /* smoother tile height: 48 or 32          */
  int persistent_blocks;       /* workgroups per persistent launch (0: auto) */
  double smoother_omega;       /* SOR weight of the smoother on every level
                                  but the coarsest; <= 0: 1.0 (Gauss-Seidel)
                                  (new in ABI 6)                           */
} smk_batch_params;

typedef struct smk_batch_info {
  int iterations;              /* refinement iterations of the slowest scene */
  int cycles_total;            /* V-cycles summed over scenes             */
  double max_rel_residual;     /* worst scene's final relative residual   */
} smk_batch_info;

typedef struct smk_batch smk_batch;
int smk_batch_create(const smk_batch_scene* scenes, int n,
                     const smk_batch_params* params, void* stream,
                     smk_batch** out);
/* One step of every scene.  Reads a 4-byte active-scene count back once
 * per refinement iteration; fills info when non-NULL. */
int smk_batch_step(smk_batch* b, smk_batch_info* info, void* stream);
/* Which density buffer (0 or 1) holds the current density. */
int smk_batch_density_index(const smk_batch* b);
/* Per-scene V-cycles and relative residual of the last step (n entries,
 * either may be NULL; synchronises the stream). */
int smk_batch_scene_stats(const smk_batch* b, int* cycles, double* rel,
                          void* stream);
/* Change tile height (48/32), V(s,s) sweeps (1..3) and persistent grid
 * size; rebuilds the job lists and graphs. */
int smk_batch_set_tuning(smk_batch* b, int tile_rows, int smoother_iterations,
                         int persistent_blocks, void* stream);
int smk_batch_destroy(smk_batch* b);

/* ------------------------------------------------- profiling metrics */
/* profiling/config.py:268-301 reducers over profiling/metrics.py:7-28
 * scopes, on the device: out18[scope * 6 + reducer] (float64, device
 * memory), scopes all / interior ([1:-1, 1:-1]) / boundary, reducers rms /
 * mean / mean_abs / max / max_abs / std (numpy semantics, ddof 0; NaN for
 * an empty scope).  field is rows x cols. */
size_t smk_field_stats_workspace_bytes(void);
int smk_field_stats(const void* field, int rows, int cols, void* out18,
                    void* workspace, size_t workspace_bytes, int dtype,
                    void* stream);

/* ------------------------------------------------- profiling / tuning */
/* No reference counterpart: measurement and tuning hooks.
 * smk_profile_enable(1) makes every solver launch record a HIP event pair on
 * its stream, tagged with the launch's algorithmic bytes; with on = k > 1 the
 * finest-level smoothing launches (idempotent) are repeated k times (at most
 * 64) inside their pair, so the pair's own overhead is spread over k
 * launches.  While profiling, gated solves wait for each iteration's check
 * before queueing the next (no gated-off launches in the records).
 * smk_profile_read synchronises, aggregates by kernel family (first-seen
 * order) into up to n records, optionally clears, and returns the number of
 * families. */
typedef struct smk_profile_record {
  char name[48];
  int launches;
  double total_ms;
  double alg_bytes;
} smk_profile_record;
int smk_profile_enable(int on);
int smk_profile_read(smk_profile_record* out, int n, int reset);

/* Red-black smoother tile variant per dtype for the fine level (which = 0:
 * finest multigrid level and the single-grid solvers) and the coarse levels
 * (which = 1): 0 = 64x48 LDS tile, 1 = 64x32 LDS tile, 2 = 64x64
 * register-resident region, 3 = 64x48 tile with only p in LDS (rhs and
 * flags in registers), 4 = the same at 64x32, 5 = the same at 64x64 with
 * 512 threads, 6 = 64x16 LDS tile, 7 = 64x32 tile with p staged by LDS-DMA
 * (rhs and flags in registers), 8 = the same at 64x64 with 512 threads,
 * 9 = the same at 64x16.
 * which = 3 selects the tile variant of
 * the coarse levels of at most 512x512 cells (default 6; which = 1 then
 * covers the larger coarse levels).  which = 2 selects the threads of the
 * single-workgroup small-level V-cycle: 0 = 256, 1 = 512, 2 = 1024; 3 = 256
 * and 4 = 512 also run its red-black levels of at most 32x32 cells, the
 * coarsest included, on one wavefront with the cells in registers; 5 = 256,
 * 6 = 512 and 7 = 1024 run only a coarsest level of at most 16x16 cells
 * that way.
 * Results do not depend on the variant. */
int smk_set_smoother_variant(int dtype, int which, int variant);
int smk_get_smoother_variant(int dtype, int which);

#ifdef __cplusplus
}
#endif
#endif
