// smk_batch.inl -- batched scenes: the multigrid side and the step driver.
// Included at the end of smk_solve.hip (it reuses that file's tile and
// small-level kernel bodies and the per-scene multigrid hierarchy).
//
// One step of n scenes:
//   pre   init flags | box + forces | advect u, v | advect rho |
//         divergence + p := 0 + ||alpha div||^2 partials | per-scene b2
//   iter  (repeated until no scene is active)
//         FP64 defect + ||r||^2 partials | per-scene check (active flag,
//         count) | V-cycle: pre-smooth+restrict by depth, one small-level
//         workgroup per scene, prolong+post-smooth by depth | p += e
//   post  gradient + boundaries, density parity flip
// Every launch is persistent over a packed job list and skips the tiles of
// inactive scenes, so each stage is one launch for all scenes and the
// iteration can be replayed from a HIP graph; the host reads one int (the
// active count) per iteration.  For each scene the kernels, their
// arguments and their order are those of the single-scene FP32-in-FP64
// solve (mg_solve_mixed), so a scene's result does not depend on the batch.

#include "smk_batch.hpp"

namespace smk {

template <int RH>
__global__ __launch_bounds__(TILE_NT) void k_b_smooth(const SmoothArgs<float>* __restrict__ args,
                                                      const int4* __restrict__ jobs, int njobs,
                                                      const int* __restrict__ active) {
  for (int q = blockIdx.x; q < njobs; q += gridDim.x) {
    const int4 jb = jobs[q];
    if (!active[jb.w]) continue;
    smooth_tile64_body<float, MODE_RB, RH, TILE_NT>(args[jb.x], jb.y, jb.z, 0);
    __syncthreads();
  }
}

__global__ __launch_bounds__(SMALL_NT) void k_b_vcycle_small(
    const SmallArgs<float>* __restrict__ args, const int* __restrict__ scene_of,
    const int* __restrict__ active) {
  if (!active[scene_of[blockIdx.x]]) return;
  vcycle_small_body<float, MODE_RB, SMALL_NT, 0>(args[blockIdx.x]);
}

__global__ void k_b_init(int n, int* __restrict__ active, int* __restrict__ cycles) {
  for (int s = threadIdx.x; s < n; s += blockDim.x) {
    active[s] = 1;
    cycles[s] = 0;
  }
}

// k_residual_norm (T = double, with the FP32 defect) per 64 x 4 tile.
__global__ __launch_bounds__(NB_T) void k_b_defect(const BatchScene* __restrict__ sc,
                                                   const int4* __restrict__ jobs, int njobs,
                                                   const int* __restrict__ active,
                                                   double* __restrict__ partial,
                                                   int* __restrict__ count) {
  __shared__ double red[NB_T / 64];
  if (blockIdx.x == 0 && threadIdx.x == 0) *count = 0;
  for (int q = blockIdx.x; q < njobs; q += gridDim.x) {
    const int4 jb = jobs[q];
    if (!active[jb.x]) continue;
    const BatchScene& d = sc[jb.x];
    const int H = d.H, W = d.W;
    const int i = jb.y * NB_X + (threadIdx.x & (NB_X - 1));
    const int j = jb.z * NB_Y + (threadIdx.x / NB_X);
    double acc = 0.0;
    if (i < W && j < H) {
      const size_t o = (size_t)j * W + i;
      const uint8_t f = d.flags0[o];
      const double* p = d.p;
      // operands loaded (clamped indices) before the flags are tested
      const size_t n = (size_t)H * W;
      const double ce = p[o], pl = p[o > 0 ? o - 1 : 0], pr = p[o + 1 < n ? o + 1 : o];
      const double pd = p[o >= (size_t)W ? o - W : o], pu = p[o + W < n ? o + W : o];
      const double b = d.div[o];
      double rv = 0.0;
      if (f & F_LFLUID) {
        const double l = (f & F_LEFT) ? pl : ce;
        const double rr = (f & F_RIGHT) ? pr : ce;
        const double dn = (f & F_DOWN) ? pd : ce;
        const double up = (f & F_UP) ? pu : ce;
        rv = poisson_residual(l, rr, dn, up, ce, b, d.alpha0);
        acc = rv * rv;
      }
      d.defect[o] = (f & F_LFLUID) ? (float)(rv / d.alpha0) : 0.0f;
    }
    const double tot = block_sum<NB_T>(acc, red);
    if (threadIdx.x == 0) partial[d.part_base + jb.z * d.nbx + jb.y] = tot;
  }
}

// Per scene (one workgroup each): the k_sum_partials reduction of its tiles;
// mode 0 stores ||alpha div||^2, mode 1 turns ||r||^2 into the relative
// residual and decides whether the scene runs another V-cycle.
__global__ __launch_bounds__(1024) void k_b_reduce(const BatchScene* __restrict__ sc,
                                                   const double* __restrict__ partial, int mode,
                                                   double* __restrict__ b2,
                                                   double* __restrict__ rel_out,
                                                   int* __restrict__ active,
                                                   int* __restrict__ cycles,
                                                   int* __restrict__ count, double tol,
                                                   int max_cycles) {
  __shared__ double red[1024 / 64];
  const int s = blockIdx.x;
  if (mode == 1 && !active[s]) return;
  const BatchScene& d = sc[s];
  const int np = d.nbx * d.nby;
  double acc = 0.0;
  for (int k = threadIdx.x; k < np; k += 1024) acc += partial[d.part_base + k];
  const double tot = block_sum<1024>(acc, red);
  if (threadIdx.x != 0) return;
  if (mode == 0) {
    b2[s] = tot;
    return;
  }
  const double bb = b2[s];
  const double rel = bb > 0.0 ? sqrt(tot / bb) : 0.0;
  rel_out[s] = rel;
  const int c = cycles[s];
  const bool done = (tol > 0.0 ? rel <= tol : c >= max_cycles) || c >= max_cycles;
  if (done) {
    active[s] = 0;
  } else {
    cycles[s] = c + 1;
    atomicAdd(count, 1);
  }
}

// k_add_correction: p += e (float64 += float32)
__global__ __launch_bounds__(NB_T) void k_b_correct(const BatchScene* __restrict__ sc,
                                                    const int4* __restrict__ jobs, int njobs,
                                                    const int* __restrict__ active) {
  for (int q = blockIdx.x; q < njobs; q += gridDim.x) {
    const int4 jb = jobs[q];
    if (!active[jb.x]) continue;
    const BatchScene& d = sc[jb.x];
    const int i = jb.y * NB_X + (threadIdx.x & (NB_X - 1));
    const int j = jb.z * NB_Y + (threadIdx.x / NB_X);
    if (i < d.W && j < d.H) {
      const size_t o = (size_t)j * d.W + i;
      d.p[o] = d.p[o] + (double)d.e[o];
    }
  }
}

struct BatchDepth {
  int npre = 0, npost = 0;
  int4* pre = nullptr;   // device job lists {args index, tile x, tile y, scene}
  int4* post = nullptr;
};

}  // namespace smk

struct smk_batch {
  int n = 0;
  smk_batch_params prm{};
  smk::BatchPhys ph{};
  std::vector<smk_mg*> mg;
  std::vector<smk::BatchScene> hs;
  std::vector<int> small_from;
  smk::BatchScene* d_sc = nullptr;
  int4* d_face = nullptr;
  int4* d_cell = nullptr;
  int n_face = 0, n_cell = 0;
  double* d_partial = nullptr;
  int npart = 0;
  double* d_b2 = nullptr;
  double* d_rel = nullptr;
  int* d_active = nullptr;
  int* d_cycles = nullptr;
  int* d_count = nullptr;
  int* d_parity = nullptr;
  // pinned ring of active counts, one slot per queued iteration (4 slots),
  // and its device address (publications numbered by next_ring_sequence)
  smk::CountSlot* h_ring = nullptr;
  smk::CountSlot* d_ring = nullptr;
  // V-cycle
  std::vector<smk::BatchDepth> depth;
  smk::SmoothArgs<float>* d_args = nullptr;
  smk::SmallArgs<float>* d_small = nullptr;
  int* d_small_scene = nullptr;
  size_t small_lds = 0;
  int blocks = 0;
  // graphs
  hipStream_t cap = nullptr;
  hipGraphExec_t g_pre = nullptr, g_iter = nullptr, g_post = nullptr;
  int parity = 0;
  int last_iterations = 0;
};

namespace smk {

static void batch_free_vcycle(smk_batch* b) {
  for (BatchDepth& d : b->depth) {
    if (d.pre) (void)hipFree(d.pre);
    if (d.post) (void)hipFree(d.post);
  }
  b->depth.clear();
  if (b->d_args) (void)hipFree(b->d_args);
  if (b->d_small) (void)hipFree(b->d_small);
  if (b->d_small_scene) (void)hipFree(b->d_small_scene);
  b->d_args = nullptr;
  b->d_small = nullptr;
  b->d_small_scene = nullptr;
}

static void batch_free_graphs(smk_batch* b) {
  for (hipGraphExec_t* g : {&b->g_pre, &b->g_iter, &b->g_post}) {
    if (*g) (void)hipGraphExecDestroy(*g);
    *g = nullptr;
  }
}

template <typename V>
static int upload(V** dst, const std::vector<V>& src, hipStream_t s) {
  *dst = nullptr;
  if (src.empty()) return SMK_OK;
  int rc = mg_alloc((void**)dst, src.size() * sizeof(V));
  if (rc) return rc;
  SMK_CHECK_HIP(hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(V), hipMemcpyHostToDevice, s));
  SMK_CHECK_HIP(hipStreamSynchronize(s));
  return SMK_OK;
}

static bool g_batch_attr_set = false;

// Arguments and job lists of the FP32 V-cycle for the current tuning, built
// exactly as mg_vcycle / mg_smooth / mg_small build them for one scene.
static int batch_build_vcycle(smk_batch* b, hipStream_t s) {
  batch_free_vcycle(b);
  const int si = b->prm.smoother_iterations, ci = b->prm.coarsest_iterations;
  const int rh = b->prm.tile_rows;
  const int npass = 2 * si;
  const bool redisc = b->prm.coarse_scale == SMK_COARSE_REDISCRETIZED;
  const double rho = b->ph.fluid_density, dt = b->ph.dt;
  std::vector<SmoothArgs<float>> args;
  std::vector<std::vector<int4>> pre, post;
  std::vector<SmallArgs<float>> small;
  std::vector<int> small_scene;
  size_t lds = 0;
  for (int sc = 0; sc < b->n; ++sc) {
    smk_mg* mg = b->mg[sc];
    const int f = b->small_from[sc];
    if ((int)pre.size() < f) {
      pre.resize(f);
      post.resize(f);
    }
    double h = b->hs[sc].h;
    for (int k = 0; k < f; ++k, h *= 2.0) {
      Level& L = mg->lv[k];
      Level& C = mg->lv[k + 1];
      const double alpha = calc_alpha(h, rho, dt);
      SmoothArgs<float> a{};
      a.p_in = reinterpret_cast<const float*>(L.p32[0]);
      a.p_out = reinterpret_cast<float*>(L.p32[1]);
      a.rhs = reinterpret_cast<const float*>(L.rhs32);
      a.flags = L.flags;
      a.H = L.H;
      a.W = L.W;
      a.alpha = alpha;
      a.omega = b->prm.smoother_omega > 0.0 ? b->prm.smoother_omega : 1.0;
      a.npass = npass;
      a.first_colour = 0;
      a.R = npass + 1;
      a.zero_init = 1;
      a.coarse_rhs = reinterpret_cast<float*>(C.rhs32);
      a.coarse_div = redisc ? alpha : calc_alpha(h * 2.0, rho, dt);
      const int ipre = (int)args.size();
      args.push_back(a);
      SmoothArgs<float> q{};
      q.p_in = reinterpret_cast<const float*>(L.p32[1]);
      q.p_out = reinterpret_cast<float*>(L.p32[0]);
      q.rhs = a.rhs;
      q.flags = L.flags;
      q.H = L.H;
      q.W = L.W;
      q.alpha = alpha;
      q.omega = a.omega;
      q.npass = npass;
      q.first_colour = 0;
      q.R = npass;
      q.zero_init = 0;
      q.coarse_e = reinterpret_cast<const float*>(C.p32[0]);
      q.Hc = C.H;
      q.Wc = C.W;
      q.coarse_div = 1.0;
      const int ipost = (int)args.size();
      args.push_back(q);
      for (int pass = 0; pass < 2; ++pass) {
        const int R = pass == 0 ? a.R : q.R;
        const int tx = 64 - 2 * R, ty = rh - 2 * R;
        if (tx <= 0 || ty <= 0) return fail(SMK_EINVAL, "tile of %d rows too small for halo %d", rh, R);
        const int gx = (int)cdiv(L.W, tx), gy = (int)cdiv(L.H, ty);
        auto& list = pass == 0 ? pre[k] : post[k];
        for (int y = 0; y < gy; ++y)
          for (int x = 0; x < gx; ++x) list.push_back(make_int4(pass == 0 ? ipre : ipost, x, y, sc));
      }
    }
    // levels f .. nl-1 in one small-level workgroup (mg_small)
    SmallArgs<float> m{};
    m.nl = mg->nl - f;
    double hl = b->hs[sc].h * (double)(1 << f);
    for (int k = 0; k < m.nl; ++k) {
      const Level& L = mg->lv[f + k];
      m.H[k] = L.H;
      m.W[k] = L.W;
      m.flags[k] = L.flags;
      m.alpha[k] = calc_alpha(hl, rho, dt);
      m.omega[k] = b->prm.smoother_omega > 0.0 ? b->prm.smoother_omega : 1.0;
      if (k == m.nl - 1 && b->prm.coarsest_omega != 0.0)
        m.omega[k] = b->prm.coarsest_omega > 0.0 ? b->prm.coarsest_omega
                                                 : optimal_omega(L.H > L.W ? L.H : L.W);
      m.coarse_div[k] = redisc ? m.alpha[k] : calc_alpha(hl * 2.0, rho, dt);
      hl *= 2.0;
    }
    Level& T0 = mg->lv[f];
    m.rhs_top = reinterpret_cast<const float*>(T0.rhs32);
    m.p_top_in = reinterpret_cast<const float*>(T0.p32[0]);
    m.p_top_out = reinterpret_cast<float*>(T0.p32[0]);
    m.zero_top = 1;
    m.npass_smooth = npass;
    m.npass_coarsest = 2 * ci;
    const size_t l = small_lds_bytes(m.H, m.W, m.nl, sizeof(float), MODE_RB);
    lds = l > lds ? l : lds;
    small.push_back(m);
    small_scene.push_back(sc);
  }
  const size_t lim = small_lds_limit<float>();
  if (lds > lim) return fail(SMK_EINVAL, "small-level LDS %zu too large", lds);
  if (!g_batch_attr_set && lim > 64 * 1024) {
    SMK_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_b_vcycle_small),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lim));
    g_batch_attr_set = true;
  }
  b->small_lds = lds;
  int rc = upload(&b->d_args, args, s);
  if (!rc) rc = upload(&b->d_small, small, s);
  if (!rc) rc = upload(&b->d_small_scene, small_scene, s);
  b->depth.assign(pre.size(), BatchDepth{});
  for (size_t k = 0; k < pre.size() && !rc; ++k) {
    b->depth[k].npre = (int)pre[k].size();
    b->depth[k].npost = (int)post[k].size();
    rc = upload(&b->depth[k].pre, pre[k], s);
    if (!rc) rc = upload(&b->depth[k].post, post[k], s);
  }
  return rc;
}

static int batch_launch_smooth(smk_batch* b, const int4* jobs, int njobs, hipStream_t s) {
  if (njobs == 0) return SMK_OK;
  const int cap = b->blocks > 0 ? b->blocks : 2048;
  const unsigned g = (unsigned)(njobs < cap ? njobs : cap);
  if (b->prm.tile_rows == 32)
    hipLaunchKernelGGL((k_b_smooth<32>), dim3(g), dim3(TILE_NT), 0, s, b->d_args, jobs, njobs,
                       b->d_active);
  else
    hipLaunchKernelGGL((k_b_smooth<48>), dim3(g), dim3(TILE_NT), 0, s, b->d_args, jobs, njobs,
                       b->d_active);
  return check_launch("k_b_smooth");
}

static int batch_stage_pre(smk_batch* b, hipStream_t s) {
  const int nb = b->blocks;
  hipLaunchKernelGGL(k_b_init, dim3(1), dim3(256), 0, s, b->n, b->d_active, b->d_cycles);
  int rc = check_launch("k_b_init");
  if (!rc) rc = batch_box_forces(b->d_sc, b->d_face, b->n_face, b->d_parity, b->ph, nb, s);
  if (!rc) rc = batch_advect_velocity(b->d_sc, b->d_face, b->n_face, b->ph, nb, s);
  if (!rc) rc = batch_advect_density(b->d_sc, b->d_cell, b->n_cell, b->d_parity, b->ph, nb, s);
  if (!rc) rc = batch_divergence_rhs_norm(b->d_sc, b->d_cell, b->n_cell, b->d_partial, nb, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_b_reduce, dim3(b->n), dim3(1024), 0, s, b->d_sc, b->d_partial, 0, b->d_b2,
                     b->d_rel, b->d_active, b->d_cycles, b->d_count, b->prm.tolerance,
                     b->prm.max_cycles);
  return check_launch("k_b_reduce");
}

static int batch_stage_iter(smk_batch* b, hipStream_t s) {
  const int cap = b->blocks > 0 ? b->blocks : 2048;
  const unsigned gc = (unsigned)(b->n_cell < cap ? b->n_cell : cap);
  hipLaunchKernelGGL(k_b_defect, dim3(gc), dim3(NB_T), 0, s, b->d_sc, b->d_cell, b->n_cell,
                     b->d_active, b->d_partial, b->d_count);
  int rc = check_launch("k_b_defect");
  if (rc) return rc;
  hipLaunchKernelGGL(k_b_reduce, dim3(b->n), dim3(1024), 0, s, b->d_sc, b->d_partial, 1, b->d_b2,
                     b->d_rel, b->d_active, b->d_cycles, b->d_count, b->prm.tolerance,
                     b->prm.max_cycles);
  rc = check_launch("k_b_reduce");
  for (size_t k = 0; k < b->depth.size() && !rc; ++k)
    rc = batch_launch_smooth(b, b->depth[k].pre, b->depth[k].npre, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_b_vcycle_small, dim3(b->n), dim3(SMALL_NT), b->small_lds, s, b->d_small,
                     b->d_small_scene, b->d_active);
  rc = check_launch("k_b_vcycle_small");
  for (size_t k = b->depth.size(); k-- > 0 && !rc;)
    rc = batch_launch_smooth(b, b->depth[k].post, b->depth[k].npost, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_b_correct, dim3(gc), dim3(NB_T), 0, s, b->d_sc, b->d_cell, b->n_cell,
                     b->d_active);
  return check_launch("k_b_correct");
}

static int batch_stage_post(smk_batch* b, hipStream_t s) {
  return batch_gradient(b->d_sc, b->d_face, b->n_face, b->d_parity, b->ph, b->blocks, s);
}

static int batch_capture(smk_batch* b, int (*stage)(smk_batch*, hipStream_t),
                         hipGraphExec_t* out) {
  hipGraph_t g = nullptr;
  SMK_CHECK_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeRelaxed));
  int rc = stage(b, b->cap);
  hipError_t e = hipStreamEndCapture(b->cap, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail((int)e, "graph capture failed: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail((int)e, "graph instantiate failed: %s", hipGetErrorString(e));
  return SMK_OK;
}

static int batch_build_graphs(smk_batch* b) {
  batch_free_graphs(b);
  if (!b->prm.use_graphs) return SMK_OK;
  const bool prof = g_prof_on;
  g_prof_on = false;  // no event records inside captured graphs
  int rc = batch_capture(b, batch_stage_pre, &b->g_pre);
  if (!rc) rc = batch_capture(b, batch_stage_iter, &b->g_iter);
  if (!rc) rc = batch_capture(b, batch_stage_post, &b->g_post);
  g_prof_on = prof;
  return rc;
}

// The active count of one batch iteration into a slot of the pinned host
// ring, then the slot's sequence number (after a system-scope fence), so the
// host can poll the slot.
__global__ void k_publish_count(const int* __restrict__ count, CountSlot* slot, unsigned seq) {
  if (threadIdx.x != 0) return;
  slot->count = *count;
  __threadfence_system();
  slot->seq = seq;
}

static void batch_release(smk_batch* b) {
  if (!b) return;
  batch_free_graphs(b);
  batch_free_vcycle(b);
  for (void* p : {(void*)b->d_sc, (void*)b->d_face, (void*)b->d_cell, (void*)b->d_partial,
                  (void*)b->d_b2, (void*)b->d_rel, (void*)b->d_active, (void*)b->d_cycles,
                  (void*)b->d_count, (void*)b->d_parity})
    if (p) (void)hipFree(p);
  if (b->h_ring) (void)hipHostFree(b->h_ring);
  for (smk_mg* m : b->mg)
    if (m) smk_mg_destroy(m);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  delete b;
}

}  // namespace smk

extern "C" {

int smk_batch_create(const smk_batch_scene* scenes, int n, const smk_batch_params* prm,
                     void* stream, smk_batch** out) {
  SMK_REQUIRE(scenes && prm && out && n > 0, "smk_batch_create: bad arguments");
  SMK_REQUIRE(prm->smoother_iterations >= 1 && prm->smoother_iterations <= 3,
              "smk_batch_create: smoother_iterations must be 1..3");
  SMK_REQUIRE(prm->tile_rows == 48 || prm->tile_rows == 32,
              "smk_batch_create: tile_rows must be 48 or 32");
  SMK_REQUIRE(prm->coarsest_iterations >= 1 && prm->max_cycles >= 1 && prm->dt > 0.0 &&
                  prm->fluid_density != 0.0,
              "smk_batch_create: bad solver parameters");
  *out = nullptr;
  hipStream_t s = as_stream(stream);
  smk_batch* b = new smk_batch();
  b->n = n;
  b->prm = *prm;
  b->ph = BatchPhys{prm->dt, prm->gravity, prm->buoyancy, prm->density_coefficient,
                    prm->ambient_temperature, prm->fluid_density};
  b->blocks = prm->persistent_blocks;
  b->mg.assign(n, nullptr);
  b->hs.assign(n, BatchScene{});
  b->small_from.assign(n, -1);
  std::vector<int4> face, cell;
  int rc = SMK_OK;
  int part = 0;
  for (int k = 0; k < n && !rc; ++k) {
    const smk_batch_scene& sc = scenes[k];
    if (sc.H < 2 || sc.W < 2 || sc.cell_size <= 0.0 || !sc.density[0] || !sc.density[1] ||
        !sc.u || !sc.u_tmp || !sc.v || !sc.v_tmp || !sc.pressure || !sc.divergence ||
        !sc.face_mask_u || !sc.face_mask_v || !sc.is_fluid || !sc.solid) {
      rc = fail(SMK_EINVAL, "smk_batch_create: scene %d: bad shape or null pointer", k);
      break;
    }
    rc = smk_mg_create(sc.face_mask_u, sc.face_mask_v, sc.H, sc.W, sc.num_levels, SMK_F64,
                       stream, &b->mg[k]);
    if (rc) break;
    smk_mg* mg = b->mg[k];
    rc = mg_ensure32(mg);
    if (rc) break;
    mg_prepare_small<float>(mg, 1, MODE_RB);
    b->small_from[k] = mg->small_from[1];
    if (b->small_from[k] < 0) {
      rc = fail(SMK_EINVAL, "smk_batch_create: scene %d: coarsest level %dx%d does not fit one "
                "workgroup; use more levels", k, mg->lv[mg->nl - 1].H, mg->lv[mg->nl - 1].W);
      break;
    }
    mg->lv[0].p[0] = sc.pressure;
    mg->lv[0].rhs = sc.divergence;
    BatchScene& d = b->hs[k];
    d.H = sc.H;
    d.W = sc.W;
    d.h = sc.cell_size;
    d.alpha0 = calc_alpha(sc.cell_size, prm->fluid_density, prm->dt);
    d.rho[0] = static_cast<double*>(sc.density[0]);
    d.rho[1] = static_cast<double*>(sc.density[1]);
    d.u = static_cast<double*>(sc.u);
    d.u_tmp = static_cast<double*>(sc.u_tmp);
    d.v = static_cast<double*>(sc.v);
    d.v_tmp = static_cast<double*>(sc.v_tmp);
    d.p = static_cast<double*>(sc.pressure);
    d.div = static_cast<double*>(sc.divergence);
    d.fmu = sc.face_mask_u;
    d.fmv = sc.face_mask_v;
    d.fluid = sc.is_fluid;
    d.solid = sc.solid;
    d.flags0 = mg->lv[0].flags;
    d.defect = static_cast<float*>(mg->lv[0].rhs32);
    d.e = static_cast<const float*>(mg->lv[0].p32[0]);
    d.box_x0 = std::max(sc.box_x0, 0);
    d.box_y0 = std::max(sc.box_y0, 0);
    d.box_x1 = std::min(sc.box_x1, sc.W);
    d.box_y1 = std::min(sc.box_y1, sc.H);
    d.box_value = sc.box_value;
    d.nbx = (int)cdiv(sc.W, NB_X);
    d.nby = (int)cdiv(sc.H, NB_Y);
    d.part_base = part;
    part += d.nbx * d.nby;
    const int fx = (int)cdiv(sc.W + 1, BT_X), fy = (int)cdiv(sc.H + 1, BT_Y);
    for (int y = 0; y < fy; ++y)
      for (int x = 0; x < fx; ++x) face.push_back(make_int4(k, x, y, 0));
    for (int y = 0; y < d.nby; ++y)
      for (int x = 0; x < d.nbx; ++x) cell.push_back(make_int4(k, x, y, 0));
  }
  b->npart = part;
  b->n_face = (int)face.size();
  b->n_cell = (int)cell.size();
  if (!rc) rc = upload(&b->d_sc, b->hs, s);
  if (!rc) rc = upload(&b->d_face, face, s);
  if (!rc) rc = upload(&b->d_cell, cell, s);
  if (!rc) rc = mg_alloc((void**)&b->d_partial, (size_t)part * sizeof(double));
  if (!rc) rc = mg_alloc((void**)&b->d_b2, (size_t)n * sizeof(double));
  if (!rc) rc = mg_alloc((void**)&b->d_rel, (size_t)n * sizeof(double));
  if (!rc) rc = mg_alloc((void**)&b->d_active, (size_t)n * sizeof(int));
  if (!rc) rc = mg_alloc((void**)&b->d_cycles, (size_t)n * sizeof(int));
  if (!rc) rc = mg_alloc((void**)&b->d_count, sizeof(int));
  if (!rc) rc = mg_alloc((void**)&b->d_parity, sizeof(int));
  if (!rc) {
    hipError_t e = hipHostMalloc((void**)&b->h_ring, 4 * sizeof(CountSlot), hipHostMallocDefault);
    if (e == hipSuccess) {
      memset(b->h_ring, 0, 4 * sizeof(CountSlot));
      e = hipHostGetDevicePointer((void**)&b->d_ring, b->h_ring, 0);
    }
    if (e != hipSuccess) rc = fail((int)e, "smk_batch_create: %s", hipGetErrorString(e));
  }
  if (!rc) {
    hipError_t e = hipMemsetAsync(b->d_parity, 0, sizeof(int), s);
    if (e == hipSuccess) e = hipMemsetAsync(b->d_rel, 0, (size_t)n * sizeof(double), s);
    if (e == hipSuccess) e = hipMemsetAsync(b->d_cycles, 0, (size_t)n * sizeof(int), s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = fail((int)e, "smk_batch_create: %s", hipGetErrorString(e));
  }
  if (!rc) {
    hipError_t e = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking);
    if (e != hipSuccess) rc = fail((int)e, "hipStreamCreate failed: %s", hipGetErrorString(e));
  }
  if (!rc) rc = batch_build_vcycle(b, s);
  if (!rc) rc = batch_build_graphs(b);
  if (rc) {
    batch_release(b);
    return rc;
  }
  *out = b;
  return SMK_OK;
}

int smk_batch_step(smk_batch* b, smk_batch_info* info, void* stream) {
  SMK_REQUIRE(b, "smk_batch_step: null handle");
  hipStream_t s = as_stream(stream);
  const bool graphs = b->prm.use_graphs && b->g_iter;
  int rc = SMK_OK;
  if (graphs) {
    SMK_CHECK_HIP(hipGraphLaunch(b->g_pre, s));
  } else {
    rc = batch_stage_pre(b, s);
    if (rc) return rc;
  }
  // the active count of iteration k-1 is read while iteration k is queued
  // (an iteration after the last useful one is a no-op: every launch skips
  // inactive scenes).  Each iteration ends with k_publish_count, which
  // writes the count and then its sequence number into a slot of the pinned
  // ring; the host polls the slot of iteration k-1 -- no copy, event or
  // wait on the stream per iteration.
  volatile CountSlot* ring = b->h_ring;
  const unsigned base = next_ring_sequence((unsigned)b->prm.max_cycles + 4u);
  int it = 0, done_at = -1;
  for (; it <= b->prm.max_cycles + 1; ++it) {
    if (graphs) {
      SMK_CHECK_HIP(hipGraphLaunch(b->g_iter, s));
    } else {
      rc = batch_stage_iter(b, s);
      if (rc) return rc;
    }
    hipLaunchKernelGGL(k_publish_count, dim3(1), dim3(64), 0, s, b->d_count,
                       b->d_ring + (it & 3), base + (unsigned)it);
    rc = check_launch("k_publish_count");
    if (rc) return rc;
    if (it >= 1) {
      const int prev = (it - 1) & 3;
      const unsigned want = base + (unsigned)(it - 1);
      for (unsigned spin = 1; ring[prev].seq != want; ++spin) {
        if ((spin & 1023u) == 0u) {
          const hipError_t e = hipStreamQuery(s);
          if (e == hipSuccess && ring[prev].seq != want)
            return fail((int)hipErrorUnknown, "batch step: count %u never published", want);
          if (e != hipSuccess && e != hipErrorNotReady)
            return fail((int)e, "batch step: %s", hipGetErrorString(e));
        }
      }
      __atomic_thread_fence(__ATOMIC_ACQUIRE);
      if (ring[prev].count == 0) {
        done_at = it - 1;
        break;
      }
    }
  }
  if (done_at >= 0) it = done_at;
  if (graphs) {
    SMK_CHECK_HIP(hipGraphLaunch(b->g_post, s));
  } else {
    rc = batch_stage_post(b, s);
    if (rc) return rc;
  }
  b->parity ^= 1;
  b->last_iterations = it;
  if (info) {
    std::vector<int> cyc(b->n);
    std::vector<double> rel(b->n);
    SMK_CHECK_HIP(hipMemcpyAsync(cyc.data(), b->d_cycles, b->n * sizeof(int),
                                 hipMemcpyDeviceToHost, s));
    SMK_CHECK_HIP(hipMemcpyAsync(rel.data(), b->d_rel, b->n * sizeof(double),
                                 hipMemcpyDeviceToHost, s));
    SMK_CHECK_HIP(hipStreamSynchronize(s));
    info->iterations = it;
    info->cycles_total = 0;
    info->max_rel_residual = 0.0;
    for (int k = 0; k < b->n; ++k) {
      info->cycles_total += cyc[k];
      info->max_rel_residual = std::max(info->max_rel_residual, rel[k]);
    }
  }
  return SMK_OK;
}

int smk_batch_density_index(const smk_batch* b) { return b ? b->parity : -1; }

int smk_batch_scene_stats(const smk_batch* b, int* cycles, double* rel, void* stream) {
  SMK_REQUIRE(b, "smk_batch_scene_stats: null handle");
  hipStream_t s = as_stream(stream);
  if (cycles)
    SMK_CHECK_HIP(hipMemcpyAsync(cycles, b->d_cycles, b->n * sizeof(int), hipMemcpyDeviceToHost, s));
  if (rel)
    SMK_CHECK_HIP(hipMemcpyAsync(rel, b->d_rel, b->n * sizeof(double), hipMemcpyDeviceToHost, s));
  SMK_CHECK_HIP(hipStreamSynchronize(s));
  return SMK_OK;
}

int smk_batch_set_tuning(smk_batch* b, int tile_rows, int smoother_iterations,
                         int persistent_blocks, void* stream) {
  SMK_REQUIRE(b, "smk_batch_set_tuning: null handle");
  SMK_REQUIRE(tile_rows == 48 || tile_rows == 32, "smk_batch_set_tuning: tile_rows 48 or 32");
  SMK_REQUIRE(smoother_iterations >= 1 && smoother_iterations <= 3,
              "smk_batch_set_tuning: smoother_iterations 1..3");
  SMK_REQUIRE(persistent_blocks >= 0, "smk_batch_set_tuning: persistent_blocks >= 0");
  hipStream_t s = as_stream(stream);
  SMK_CHECK_HIP(hipStreamSynchronize(s));
  b->prm.tile_rows = tile_rows;
  b->prm.smoother_iterations = smoother_iterations;
  b->prm.persistent_blocks = persistent_blocks;
  b->blocks = persistent_blocks;
  int rc = batch_build_vcycle(b, s);
  if (!rc) rc = batch_build_graphs(b);
  return rc;
}

int smk_batch_destroy(smk_batch* b) {
  if (b) (void)hipDeviceSynchronize();
  batch_release(b);
  return SMK_OK;
}

}  // extern "C"
