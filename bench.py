"""Benchmark: smoke cell-steps/s on one MI355X, 2048x2048 multigrid plume.

Workload (plume_2048_mg): 2048x2048 MAC grid, float64 fields, dx = 1/2048,
dt = 0.005, buoyancy f_y = 1.0*density, 4 density boxes of 48x48 cells
(strength U(0.5,1.5), lower third), initial velocity = curl of 32 Gaussian
blobs (seed 3).  Pressure: geometric multigrid, V(2,2) red-black
Gauss-Seidel (omega = 1), full-weighting (masked 2x2 mean) restriction,
bilinear prolongation, 8 levels down to 16x16 solved by 50 RB-GS sweeps,
V-cycles until ||r|| / ||alpha*div|| <= 1e-8.

A step is one full SmokeSimulation2D.substep (sources, buoyancy, advection of
u, v and density, divergence, pressure solve, projection + walls, metrics
divergence and residual) with all fields resident in HBM.  Two variants:
``fp64`` (the FP64 V-cycles above) and ``mixed`` (FP32 V-cycles inside FP64
defect correction, same FP64 tolerance; by default their 16x16 level is
solved exactly, by its precomputed least-squares solution operator, in place
of the 50 sweeps, whose partial solve is what bounds the V-cycles'
convergence).  ``value`` is the faster variant's cell-steps/s
(``config.variant`` names it); both are reported.

The same line carries the two secondary workloads: ``batched_scenes`` (96
uneven scenes stepped in one batch, FP32 V-cycles in FP64 refinement to
1e-6) and ``differentiable`` (the 256^2 64-step rollout: forward vs forward +
adjoint time, FP32 and FP64, and a batch of seeded targets); ``--only``
runs one of them alone.

Roofline: ``roofline`` puts the whole step against the HBM roofline (the
north-star target): algorithmic bytes per step (one compulsory read or
write per field per logical pass, ``algorithmic_bytes_per_cell``) over the
HIP-event time of a step, as a fraction of 8 TB/s, with the step's HBM bytes
measured by counters as ``traffic`` (profiles/traffic.json, written by
``--traffic-from`` from ``--pmc-run`` counter runs).  ``roofline.kernel``
names the family carrying most algorithmic bytes with its launch time and
its counter-measured HBM rate as a fraction of the peak; its algorithmic
rate counts each of a multi-pass launch's passes and is not a fraction.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "python-game-engine_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

PEAK_HBM_GBS = 8000.0  # MI355X HBM3E, GB/s


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes_per_cell(v_cycles_per_step: float, mixed: bool, levels: int,
                               refinements: float = 0.0, coarsest_op: bool = False,
                               exact_level: int = 0, exact_cells: int = 0) -> float:
    """Bytes per cell-step by logical pass (float64 8 B, float32 4 B):
    non-solve passes 168, metrics divergence + residual 48, per FP64 V-cycle
    176.06 (V(2,2) smoothing 4x24, residual+restrict 18, prolong+correct 18,
    summed over levels x 4/3, plus 50 coarsest sweeps) + 16 for the norm
    check; mixed: per FP32 V-cycle 88.03, per refinement 40.  With
    ``coarsest_op`` the coarsest level is one read of its float32 256x256
    operator (4 x 256 B per coarsest cell) in place of the sweeps.  With
    ``exact_cells`` > 0 (mixed, coarsest_solve "exact_1024") the FP32
    V-cycles smooth levels 0 .. exact_level - 1 and read the float32 exact
    operator of level exact_level, exact_cells^2 x 4 B, in place of that
    level and those below."""
    if mixed and exact_cells > 0:
        geo = sum(0.25 ** l for l in range(exact_level))
        op_per_cell = 4.0 * exact_cells * 0.25 ** exact_level  # 4 n^2 B over n / 0.25^l cells
        return (168.0 + 48.0 + v_cycles_per_step * (66.0 * geo + op_per_cell)
                + refinements * 40.0 + 16.0)
    geo = sum(0.25 ** l for l in range(levels - 1))
    coarse_cells = 0.25 ** (levels - 1)
    coarsest = 4.0 * 256 * coarse_cells if coarsest_op else 50 * 24.0 * coarse_cells
    per_vc64 = 132.0 * geo + coarsest
    if mixed:
        per_vc = 66.0 * geo + (coarsest if coarsest_op else coarsest / 2.0)
        return 168.0 + 48.0 + v_cycles_per_step * per_vc + refinements * 40.0 + 16.0
    return 168.0 + 48.0 + v_cycles_per_step * (per_vc64 + 16.0) + 16.0


def exact_level_of(N: int, max_cells: int = 1024) -> tuple:
    """(level, cells) of the first level of an N x N hierarchy with at most
    max_cells cells (sizes halve while even), as the solver picks it."""
    n, level = N, 0
    while n * n > max_cells and n % 2 == 0:
        n //= 2
        level += 1
    return (level, n * n) if n * n <= max_cells else (0, 0)


# finest-level launches repeated inside their event pair while profiling
PROFILE_REPEAT = 8


def profile_steps(sim, scene, n: int) -> list:
    """Per kernel family over n extra (untimed) steps: launches, ms, algorithmic
    bytes, recorded by the library with HIP events on the launch stream."""
    import torch

    from cfd_amd.backends.hip import _lib

    torch.cuda.synchronize()
    _lib.profile_read(reset=True)
    _lib.profile_enable(True, repeat=PROFILE_REPEAT)
    try:
        for _ in range(n):
            sim.substep(scene.dt)
        recs = _lib.profile_read(reset=True)
    finally:
        _lib.profile_enable(False)
    for r in recs:
        r["avg_launch_us"] = 1e3 * r["total_ms"] / max(r["launches"], 1)
        r["alg_bytes_per_launch"] = r["alg_bytes"] / max(r["launches"], 1)
        r["alg_GBps"] = r["alg_bytes"] / max(r["total_ms"], 1e-12) / 1e6
    return sorted(recs, key=lambda r: -r["total_ms"])


def calibrate_copy(nbytes: int = 1 << 30) -> float:
    """A plain device copy of known size (nbytes read + nbytes written), run
    once so that HBM byte counters (FETCH_SIZE / WRITE_SIZE) of a profiled run
    can be scaled; returns its GB/s.  1 GiB per buffer is far above the
    256 MiB last-level (MALL) cache, so the counters see HBM traffic."""
    import torch

    src = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    return 2 * nbytes / (e0.elapsed_time(e1) * 1e6)


# kernel-name fragment -> the library's profiling family (the names
# smk_profile_read reports), for the PMC traffic table
TRAFFIC_FAMILIES = [
    ("k_smooth_fine_d<double", "rb_smooth_fine_f64"), ("k_smooth_fine_r<double", "rb_smooth_fine_f64"),
    ("k_smooth_fine<double", "rb_smooth_fine_f64"),
    ("k_smooth_fine_d<float", "rb_smooth_fine_f32"), ("k_smooth_fine_r<float", "rb_smooth_fine_f32"),
    ("k_smooth_fine<float", "rb_smooth_fine_f32"),
    ("k_smooth_tile64<float, 0", "rb_smooth_coarse_f32"),
    ("k_smooth_tile64<double, 0", "rb_smooth_coarse_f64"),
    ("k_smooth_tile64_r<float", "rb_smooth_coarse_f32"), ("k_smooth_tile64_d<float", "rb_smooth_coarse_f32"),
    ("k_smooth_tile64_r<double", "rb_smooth_coarse_f64"),
    ("k_smooth_tile64_d<double", "rb_smooth_coarse_f64"),
    ("k_vcycle_small<float", "vcycle_small_f32"), ("k_vcycle_small<double", "vcycle_small_f64"),
    ("k_apply_exact", "exact_apply_f32"),
    ("k_residual_norm<double", "residual_norm_f64"), ("k_correct_defect<double", "correct_defect_fp64"),
    ("k_advect_velocity<double", "advect_velocity_f64"), ("k_advect_scalar<double", "advect_scalar_f64"),
    ("k_pressure_gradient<double", "pressure_gradient_f64"), ("k_divergence<double", "divergence_f64"),
    ("k_residual_masks<double", "residual_f64"), ("k_apply_forces<double", "apply_forces_f64"),
    ("k_div_residual<double", "div_residual_f64")]


def traffic_from_pmc(fetch_csv: str, write_csv: str, copy_bytes: int = 1 << 30,
                     variant: str | None = None, steps: int = 0) -> dict:
    """HBM bytes per launch by kernel family from two rocprofv3 --pmc runs of
    this script (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch).  Each counter
    is scaled by what it reads for the run's own 1 GiB copy (calibrate_copy:
    copy_bytes read and written), so the table is in true bytes.  With
    ``variant`` and ``steps`` (a ``--pmc-run`` of that variant, which runs
    ``steps`` substeps and nothing else on the GPU but the copy), the bytes of
    every library kernel over the run divided by ``steps`` are recorded as
    that variant's HBM bytes per step."""
    import collections
    import csv

    def family(name):
        for frag, fam in TRAFFIC_FAMILIES:
            if frag in name:
                return fam
        return None

    per, factor, lib_total = {}, {}, {}
    for counter, path in (("FETCH_SIZE", fetch_csv), ("WRITE_SIZE", write_csv)):
        rows = list(csv.DictReader(open(path)))
        lib_total[counter] = sum(float(r["Counter_Value"]) for r in rows
                                 if "smk::k_" in r["Kernel_Name"])
        agg = collections.defaultdict(lambda: [0, 0.0])
        for r in rows:
            f = family(r["Kernel_Name"])
            if f is not None:
                agg[f][0] += 1
                agg[f][1] += float(r["Counter_Value"])
        per[counter] = {f: (n, v / n) for f, (n, v) in agg.items()}
        big = [float(r["Counter_Value"]) for r in rows
               if "copyBuffer" in r["Kernel_Name"] and float(r["Counter_Value"]) * 1024 > copy_bytes / 4]
        factor[counter] = (copy_bytes / 1024) / (sum(big) / len(big))
    out = {"calibration": {"copy_bytes_each_way": copy_bytes, "fetch_factor": factor["FETCH_SIZE"],
                           "write_factor": factor["WRITE_SIZE"]}}
    if variant and steps > 0:
        out["per_step"] = {variant: {
            "steps": steps,
            "hbm_bytes_per_step": round((lib_total["FETCH_SIZE"] * factor["FETCH_SIZE"]
                                         + lib_total["WRITE_SIZE"] * factor["WRITE_SIZE"])
                                        * 1024 / steps)}}
    for f, (n, fk) in sorted(per["FETCH_SIZE"].items()):
        wk = per["WRITE_SIZE"].get(f, (0, 0.0))[1]
        out[f] = {"dispatches": n, "fetch_KiB_per_launch": round(fk, 1),
                  "write_KiB_per_launch": round(wk, 1),
                  "hbm_bytes_per_launch": round(fk * 1024 * factor["FETCH_SIZE"]
                                                + wk * 1024 * factor["WRITE_SIZE"])}
    return out


def run_variant(scene, precision: str, steps: int, warmup: int, profile: int = 3,
                inner: int = 1, coarsest: str = "sweeps", arith: str = "float64",
                tune: bool = True, graphs: bool = True, omega: float = None):
    """Warm-up steps, the autotuner's choice of tile variants (same results,
    timed on this simulation from a saved state), ``steps`` timed steps,
    then as many steps with the stages around the solve replayed from HIP
    graphs, and the profiled steps."""
    import torch

    from cfd_amd import scenes, tuning

    over = {"precision": precision, "inner_cycles": inner, "coarsest_solve": coarsest,
            "inner_arithmetic": arith}
    if omega is not None:  # the smoother's SOR weight (else the scene's)
        over["smoother_omega"] = omega
    sim = scenes.make_simulation(scene, dtype="float64", device="cuda", solver_overrides=over)
    torch.cuda.synchronize()
    for _ in range(warmup):
        sim.substep(scene.dt)
    tuned = tuning.autotune(sim, scene.dt, steps=2, repeats=2) if tune else None
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    cycles, refinements = [], []
    start.record(stream)
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.substep(scene.dt)
        info = sim.last_solve_info
        cycles.append(info.get("iterations", 0))
        refinements.append(info.get("refinements", 0))
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = start.elapsed_time(end)
    graph_ms = None
    if graphs:
        sim.use_graphs = True
        for _ in range(2):  # captures both buffer assignments
            sim.substep(scene.dt)
        torch.cuda.synchronize()
        start.record(stream)
        for _ in range(steps):
            sim.substep(scene.dt)
        end.record(stream)
        torch.cuda.synchronize()
        graph_ms = start.elapsed_time(end) / steps
        sim.use_graphs = False
    kernels = profile_steps(sim, scene, profile) if profile > 0 else []
    return {
        "kernels": kernels,
        "ms_per_step": ms / steps,
        "wall_ms_per_step": wall * 1e3 / steps,
        "graph_ms_per_step": graph_ms,
        "autotune": None if tuned is None else tuned["chosen"],
        "cell_steps_per_s": scene.cells * steps / (ms / 1e3),
        "cells": scene.cells,
        "v_cycles_per_step": sum(cycles) / steps,
        "refinements_per_step": sum(refinements) / steps,
        "rel_residual_last": sim.last_solve_info.get("rel_residual"),
        "levels": sim.last_solve_info.get("levels"),
    }, sim


def step_roofline(r: dict, tmap: dict, var: str) -> dict:
    """The whole step against the HBM roofline: algorithmic bytes per step
    (algorithmic_bytes_per_cell x cells) over the HIP-event time of a step;
    ``traffic`` is the step's HBM bytes measured by counters (FETCH_SIZE +
    WRITE_SIZE of every library kernel over a --pmc-run of this variant,
    profiles/traffic.json), null without such a run.  ``kernel``: the family
    carrying most algorithmic bytes, with its HIP-event launch time, its
    algorithmic rate (a multi-pass kernel counts every pass, so this rate can
    exceed the HBM peak and is not a fraction of it) and its counter-measured
    HBM bytes and rate, the latter as a fraction of the peak."""
    alg_step = r["alg_bytes_per_cell_step"] * r["cells"]
    achieved = alg_step / (r["ms_per_step"] * 1e6)
    per_step = tmap.get("per_step", {}).get(var, {}).get("hbm_bytes_per_step")
    out = {"bound": "hbm", "scope": "whole step", "achieved": achieved, "peak": PEAK_HBM_GBS,
           "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS, "traffic": per_step,
           "alg_bytes_per_step": alg_step, "ms_per_step": r["ms_per_step"],
           "traffic_GBps": None if per_step is None else per_step / (r["ms_per_step"] * 1e6),
           "traffic_source": tmap.get("source") if per_step is not None else None}
    if r["kernels"]:
        k = max(r["kernels"], key=lambda q: q["alg_bytes"])
        hbm = tmap.get(k["name"], {}).get("hbm_bytes_per_launch")
        kg = None if hbm is None else hbm / (k["avg_launch_us"] * 1e3)
        out["kernel"] = {"name": k["name"], "avg_launch_us": k["avg_launch_us"],
                         "alg_bytes_per_launch": k["alg_bytes_per_launch"],
                         "alg_GBps": k["alg_GBps"], "hbm_bytes_per_launch": hbm,
                         "hbm_GBps": kg, "hbm_frac": None if kg is None else kg / PEAK_HBM_GBS}
    return out


def mixed_vs_fp64(sims: dict, scene) -> dict:
    """Relative L2 difference of one mixed-variant step from one FP64 step
    taken from the same state (the FP64 run's fields after its timed and
    profiled steps, copied into the mixed simulation); pressure modulo its
    mean (the all-fluid Neumann problem fixes it up to a constant).  Rollouts
    of the two variants drift apart over many steps: at this workload's CFL
    numbers the flow amplifies any difference at the solver's tolerance, as
    it does between FP64 runs at two tolerances."""
    import torch

    if "fp64" not in sims or "mixed" not in sims:
        return {}
    a, b = sims["mixed"], sims["fp64"]
    for f in ("density", "velocity_u", "velocity_v"):
        getattr(a, f).copy_(getattr(b, f))
    a.substep(scene.dt)
    b.substep(scene.dt)
    torch.cuda.synchronize()
    out = {"mixed_rel_residual": a.last_solve_info.get("rel_residual"),
           "fp64_rel_residual": b.last_solve_info.get("rel_residual")}
    for f in ("density", "velocity_u", "velocity_v", "pressure"):
        x, y = getattr(a, f), getattr(b, f)
        if f == "pressure":
            x, y = x - x.mean(), y - y.mean()
        out[f] = float((x - y).norm() / y.norm().clamp_min(1e-300))
    return out


def _cpu_model() -> str | None:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline_in_child(N: int, sample_steps: int) -> dict:
    """cpu_baseline in a child process started with OMP_PROC_BIND=close: the
    OpenMP threads stay on the cores they start on, and this process (whose
    torch runtime would pin its main thread under that setting) is not
    bound."""
    import subprocess

    env = dict(os.environ, OMP_PROC_BIND="close")
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-child",
                        str(N), str(sample_steps)], env=env, stdout=subprocess.PIPE, text=True,
                       check=True)
    return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_baseline(scene, sample_steps: int):
    import oracle
    import oracle.sim

    cpus = len(os.sched_getaffinity(0))  # before the OpenMP runtime binds this thread
    ref = oracle.sim.from_scene(scene)
    ref.substep(scene.dt)  # warm-up (page-in, thread start)
    t0 = time.perf_counter()
    cyc = 0
    for _ in range(sample_steps):
        ref.substep(scene.dt)
        cyc += ref.last_solve["iterations"]
    dt = time.perf_counter() - t0
    return {"value": scene.cells * sample_steps / dt, "unit": "cell-steps/s",
            "cores": oracle.num_threads(), "kind": "port", "cpu_model": _cpu_model(),
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "cpus_available": cpus,
            "sample": (f"{sample_steps} steps of the same {scene.H}x{scene.W} workload after 1 "
                       f"warm-up step, FP64 oracle (C/OpenMP restatement), "
                       f"{cyc / sample_steps:.1f} V-cycles/step, {dt:.1f} s")}


def _time_steps(sim, scene, steps: int, warmup: int) -> tuple:
    import torch

    for _ in range(warmup):
        sim.substep(scene.dt)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    its = []
    e0.record()
    for _ in range(steps):
        sim.substep(scene.dt)
        its.append(sim.last_solve_info.get("iterations", 0))
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps, sum(its) / max(len(its), 1)


def bench_small_configs(steps_c1: int = 100, steps_c2: int = 20) -> dict:
    """The 128^2 Jacobi-200 plume (100 steps) and the 512^2 RB-SOR plume
    (omega 1.9 to 1e-6, FP64 and FP32 sweeps in FP64 defect correction)."""
    from cfd_amd import scenes

    out = {}
    sc = scenes.plume_128_jacobi()
    sim = scenes.make_simulation(sc, dtype="float64", device="cuda")
    ms, its = _time_steps(sim, sc, steps_c1, 2)
    out["plume_128_jacobi"] = {"steps": steps_c1, "ms_per_step": ms,
                               "cell_steps_per_s": sc.cells / (ms / 1e3),
                               "jacobi_iterations_per_step": its}
    sc = scenes.plume_512_rbsor()
    res = {}
    for prec in ("native", "mixed"):
        sim = scenes.make_simulation(sc, dtype="float64", device="cuda",
                                     solver_overrides={"precision": prec})
        ms, its = _time_steps(sim, sc, steps_c2, 2)
        res["fp64" if prec == "native" else "mixed"] = {
            "ms_per_step": ms, "cell_steps_per_s": sc.cells / (ms / 1e3),
            "sweeps_per_step": its, "rel_residual_last": sim.last_solve_info.get("rel_residual")}
    res["steps"] = steps_c2
    res["mixed_speedup_vs_fp64"] = res["fp64"]["ms_per_step"] / res["mixed"]["ms_per_step"]
    out["plume_512_rbsor"] = res
    return out


def bench_batched(count: int, steps: int, omega: float = 1.0) -> dict:
    """The uneven-scenes workload: ``count`` scenes (N log-uniform in
    [64, 1024], multiple of 16, aspect 1:1 or 2:1, seeds 4000+s), one
    batch, FP32 V-cycles in FP64 defect correction to 1e-6 per scene; the
    autotuner picks tile height, sweeps and grid size first."""
    import torch

    from cfd_amd import scenes
    from cfd_amd.batch import BatchedSmoke

    sl = scenes.uneven_scenes(count=count, steps=steps)
    b = BatchedSmoke(sl, tolerance=1e-6, smoother_omega=omega)
    b.step()
    tune = b.autotune(tile_rows=(48, 32), smoother_iterations=(1, 2, 3),
                      persistent_blocks=(2048, 4096), repeats=1)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    infos = []
    e0.record()
    for _ in range(steps):
        infos.append(b.step())
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    cyc, rel = b.scene_stats()
    out = {"workload": f"uneven_{count}", "scenes": count, "cells": b.cells, "steps": steps,
           "ms_per_step": ms, "cell_steps_per_s": b.cells / (ms / 1e3),
           "tuning": tune["chosen"], "iterations_per_step": [i["iterations"] for i in infos],
           "v_cycles_last_step": sum(cyc), "max_rel_residual": max(rel)}
    b.close()
    # the same scenes one at a time (single-scene path), a few steps
    sims = [scenes.make_simulation(s, dtype="float64", device="cuda",
                                   solver_overrides={"precision": "mixed", "inner_cycles": 1,
                                                     "smoother_iterations":
                                                         tune["chosen"]["smoother_iterations"]})
            for s in sl]
    for sim, s in zip(sims, sl):
        sim.compute_metrics = False
        sim.substep(s.dt)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(2):
        for sim, s in zip(sims, sl):
            sim.substep(s.dt)
    e1.record()
    torch.cuda.synchronize()
    out["one_at_a_time_ms_per_step"] = e0.elapsed_time(e1) / 2
    out["speedup_vs_one_at_a_time"] = out["one_at_a_time_ms_per_step"] / ms
    return out


def bench_differentiable(steps: int, batch: int, grad_check: bool = True) -> dict:
    """The 256^2 differentiable rollout: 64 steps, MG to 1e-8, controls u0,
    v0 (zero) and per-step strengths (1.0), loss 0.5||rho_T - target||^2
    with a Gaussian target; forward-only vs forward + adjoint (checkpoint
    every 8 steps, recompute between) in FP32 and FP64, and a batch of
    seeded targets."""
    import torch

    from cfd_amd import diff, scenes

    sc = scenes.diff_rollout_256(steps=steps)
    N = sc.H
    out = {"workload": f"diff_{N}", "steps": steps}

    def ev():
        return torch.cuda.Event(enable_timing=True)

    for name, dt_ in (("fp32", torch.float32), ("fp64", torch.float64)):
        m = diff.from_scene(sc, dtype=dt_)
        tgt = torch.as_tensor(diff.gaussian_target(N, N, sc.cell_size), dtype=dt_, device="cuda")

        def controls(grad):
            z = dict(dtype=dt_, device="cuda", requires_grad=grad)
            return (torch.zeros((N, N + 1), **z), torch.zeros((N + 1, N), **z),
                    torch.ones(steps, **z))

        def fwd():
            u0, v0, s = controls(False)
            return m.forward_only(u0, v0, s, steps)

        def fb(target=tgt):
            u0, v0, s = controls(True)
            rho, _, _ = m.rollout(u0, v0, s, steps, checkpoint_every=8)
            diff.density_loss(rho, target).backward()
            return u0.grad, v0.grad, s.grad

        fwd()
        fb()
        res = {}
        for key, f in (("forward_ms", fwd), ("forward_adjoint_ms", fb)):
            a, c = ev(), ev()
            a.record()
            f()
            c.record()
            torch.cuda.synchronize()
            res[key] = a.elapsed_time(c)
        res["grad_fwd_ratio"] = res["forward_adjoint_ms"] / res["forward_ms"]
        if name == "fp32" and batch > 0:
            targets = [torch.as_tensor(t, dtype=dt_, device="cuda")
                       for t in diff.seeded_targets(N, N, sc.cell_size, range(500, 500 + batch))]
            a, c = ev(), ev()
            a.record()
            for t in targets:
                fb(t)
            c.record()
            torch.cuda.synchronize()
            res["batch_targets"] = batch
            res["batch_ms"] = a.elapsed_time(c)
        out[name] = res
    if grad_check:
        out["gradient_check"] = _gradient_check(sc, steps)
    return out


def _gradient_check(sc, steps: int, directions: int = 32, eps: float = 1e-7) -> dict:
    """FP64 gradient at the rollout's starting controls vs central differences
    along seeded random directions (solves run a fixed 30 V-cycles, converged
    to round-off), and the FP32 gradient's relative error against the FP64
    one.  From zero velocity every backtrace sits on a grid line, a kink of
    the bilinear interpolation: the difference quotient approaches the
    adjoint's mean slope with an error that falls with eps (max 3.0e-6 at
    1e-6, 3.2e-7 at 1e-7 over these directions), so eps is 1e-7."""
    import numpy as np
    import torch

    from cfd_amd import diff

    N = sc.H
    grads = {}
    for name, dt_, kw in (("fp64", torch.float64, {"tolerance": 0.0, "v_cycles": 30}),
                          ("fp64_tol", torch.float64, {}), ("fp32", torch.float32, {})):
        m = diff.from_scene(sc, dtype=dt_, **kw)
        tgt = torch.as_tensor(diff.gaussian_target(N, N, sc.cell_size), dtype=dt_, device="cuda")
        z = dict(dtype=dt_, device="cuda", requires_grad=True)
        u0 = torch.zeros((N, N + 1), **z)
        v0 = torch.zeros((N + 1, N), **z)
        s = torch.ones(steps, **z)
        rho, _, _ = m.rollout(u0, v0, s, steps, checkpoint_every=8)
        diff.density_loss(rho, tgt).backward()
        grads[name] = (u0.grad.double(), v0.grad.double(), s.grad.double())
        if name == "fp64":
            model, target = m, tgt
    rng = np.random.default_rng(50)
    errs = []
    zero = dict(dtype=torch.float64, device="cuda")
    for _ in range(directions):
        du = torch.as_tensor(rng.standard_normal((N, N + 1)), **zero)
        dv = torch.as_tensor(rng.standard_normal((N + 1, N)), **zero)
        ds = torch.as_tensor(rng.standard_normal(steps), **zero)
        gu, gv, gs = grads["fp64"]
        adj = float((gu * du).sum() + (gv * dv).sum() + (gs * ds).sum())
        fd = diff.directional_fd(model, torch.zeros((N, N + 1), **zero),
                                 torch.zeros((N + 1, N), **zero), torch.ones(steps, **zero),
                                 steps, target, du, dv, ds, eps)
        errs.append(abs(fd - adj) / abs(fd))

    def rel(a, b):
        """max |a - b| / max |b| per control (u0, v0, strengths)"""
        return {k: float((x - y).abs().max() / y.abs().max())
                for k, x, y in zip(("u0", "v0", "strengths"), a, b)}

    return {"directions": directions, "eps": eps,
            "fd_max_rel_err": max(errs), "fd_median_rel_err": float(np.median(errs)),
            "tolerance_1e-8_vs_converged_rel_err": rel(grads["fp64_tol"], grads["fp64"]),
            "fp32_vs_fp64_rel_err": rel(grads["fp32"], grads["fp64"])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--N", type=int, default=2048)
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default tile variants (profiled runs: a profiler's overhead "
                         "distorts the autotuner's timings)")
    ap.add_argument("--variants", default="fp64,mixed")
    ap.add_argument("--mixed-omega", type=float, default=1.2,
                    help="SOR weight of the mixed variant's FP32 V-cycle smoother (the FP64 "
                         "variant keeps the workload's red-black Gauss-Seidel, 1.0)")
    ap.add_argument("--batch-omega", type=float, default=1.2,
                    help="SOR weight of the uneven-batch workload's FP32 V-cycle smoother")
    ap.add_argument("--mixed-inner", type=int, default=2,
                    help="FP32 V-cycles per FP64 refinement in the mixed variant")
    ap.add_argument("--mixed-coarsest", choices=["sweeps", "operator", "exact", "exact_1024"],
                    default="exact_1024",
                    help="coarsest level of the mixed variant's FP32 V-cycles: the 50 sweeps, "
                         "those sweeps applied as their precomputed 256x256 operator, the "
                         "level's exact least-squares operator, or the exact operator of the "
                         "32x32 level (the cycles stop there)")
    ap.add_argument("--mixed-arith", choices=["float64", "float32"], default="float32",
                    help="arithmetic of the mixed variant's FP32 V-cycle passes: the reference's "
                         "float64 expression, or float32")
    ap.add_argument("--batched-scenes", type=int, default=96,
                    help="scenes of the uneven-batch workload (0 skips it)")
    ap.add_argument("--diff-steps", type=int, default=64,
                    help="steps of the differentiable rollout (0 skips it)")
    ap.add_argument("--diff-batch", type=int, default=16)
    ap.add_argument("--no-small-configs", action="store_true",
                    help="skip the 128^2 Jacobi and 512^2 RB-SOR lines")
    ap.add_argument("--only", choices=("batched", "diff", "small"), default=None,
                    help="run one secondary workload alone and print its JSON (profiling runs)")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="merge into profiles/traffic.json the HBM bytes from the "
                         "counter_collection.csv files of two rocprofv3 --pmc runs (FETCH_SIZE, "
                         "WRITE_SIZE) of `bench.py --pmc-run VARIANT`, then exit")
    ap.add_argument("--traffic-variant", default=None,
                    help="with --traffic-from: the variant the counter runs ran")
    ap.add_argument("--cpu-baseline-child", nargs=2, type=int, metavar=("N", "STEPS"),
                    help=argparse.SUPPRESS)
    ap.add_argument("--pmc-run", choices=("fp64", "mixed"), default=None,
                    help="counter run: only this variant's warm-up and timed steps and the "
                         "calibration copy, no profiling, other workloads or CPU baseline; "
                         "prints the step count --traffic-from needs")
    args = ap.parse_args()

    if args.traffic_from:
        tfile = os.path.join(ROOT, "profiles", "traffic.json")
        table = {}
        if os.path.exists(tfile):
            with open(tfile) as f:
                table = json.load(f)
        steps = args.warmup + args.steps if args.traffic_variant else 0
        new = traffic_from_pmc(*args.traffic_from, variant=args.traffic_variant, steps=steps)
        per_step = dict(table.get("per_step", {}))
        per_step.update(new.pop("per_step", {}))
        table.update(new)
        table["per_step"] = per_step
        table["source"] = ("rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, of `python3 bench.py "
                           "--pmc-run VARIANT --steps K --warmup W`; per_step = all library "
                           "kernels of the run / (W + K)")
        with open(tfile, "w") as f:
            json.dump(table, f, indent=1)
            f.write("\n")
        print(json.dumps(table, indent=1))
        return

    if args.cpu_baseline_child:
        # the CPU baseline, alone in a child process that never touches the GPU
        n, k = args.cpu_baseline_child
        from cfd_amd import scenes

        print(json.dumps(cpu_baseline(scenes.plume_2048_mg(N=n), k)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU; --gpus {args.gpus} is not supported "
              "(it measures one GPU; nothing here runs across GPUs)", file=sys.stderr)
        sys.exit(2)

    import torch

    from cfd_amd import scenes

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)

    if args.only == "batched":
        print(json.dumps(bench_batched(args.batched_scenes or 96, args.steps, args.batch_omega)),
              flush=True)
        return
    if args.only == "small":
        print(json.dumps(bench_small_configs()), flush=True)
        return
    if args.only == "diff":
        print(json.dumps(bench_differentiable(args.diff_steps or 64, args.diff_batch)), flush=True)
        return

    log(f"generating plume_{args.N}_mg")
    scene = scenes.plume_2048_mg(N=args.N)
    if args.pmc_run:
        var = args.pmc_run
        run_variant(scene, "mixed" if var == "mixed" else "native", args.steps, args.warmup,
                    profile=0, tune=False, graphs=False,
                    inner=args.mixed_inner if var == "mixed" else 1,
                    coarsest=args.mixed_coarsest if var == "mixed" else "sweeps",
                    arith=args.mixed_arith if var == "mixed" else "float64",
                    omega=args.mixed_omega if var == "mixed" else None)
        calibrate_copy()
        print(json.dumps({"pmc_run": var, "steps": args.warmup + args.steps}), flush=True)
        return
    results, sims = {}, {}
    for var in args.variants.split(","):
        log(f"variant {var}: warmup {args.warmup}, steps {args.steps}")
        inner = args.mixed_inner if var == "mixed" else 1
        coarsest = args.mixed_coarsest if var == "mixed" else "sweeps"
        arith = args.mixed_arith if var == "mixed" else "float64"
        omega = args.mixed_omega if var == "mixed" else None
        r, sims[var] = run_variant(scene, "mixed" if var == "mixed" else "native", args.steps,
                                   args.warmup, inner=inner, coarsest=coarsest, arith=arith,
                                   tune=not args.no_tune, omega=omega)
        r["inner_cycles_per_refinement"] = inner
        r["inner_smoother_omega"] = omega
        r["coarsest_solve"] = coarsest
        r["inner_arithmetic"] = arith if var == "mixed" else None
        ex_level, ex_cells = (exact_level_of(args.N) if var == "mixed" and
                              coarsest == "exact_1024" else (0, 0))
        r["alg_bytes_per_cell_step"] = algorithmic_bytes_per_cell(
            r["v_cycles_per_step"], var == "mixed", r["levels"] or 8, r["refinements_per_step"],
            coarsest_op=var == "mixed" and coarsest in ("operator", "exact"),
            exact_level=ex_level, exact_cells=ex_cells)
        r["alg_GBps_whole_step"] = r["alg_bytes_per_cell_step"] * r["cell_steps_per_s"] / 1e9
        results[var] = r
        log(f"variant {var}: {json.dumps(r)}")

    # the headline variant is the faster one; both are reported
    main_var = min(results, key=lambda k: results[k]["ms_per_step"])
    m = results[main_var]
    accuracy = mixed_vs_fp64(sims, scene)
    sims.clear()
    copy_gbps = calibrate_copy()
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    tmap = {}
    if os.path.exists(tfile):
        with open(tfile) as f:
            tmap = json.load(f)
    roof = {var: step_roofline(r, tmap, var) for var, r in results.items()}
    cpu = None
    if not args.no_cpu_baseline:
        log(f"cpu baseline: {args.cpu_steps} oracle steps")
        cpu = cpu_baseline_in_child(args.N, args.cpu_steps)
        log(f"cpu baseline: {cpu}")

    extra = {}
    from cfd_amd import tuning

    # the 2048^2 autotuning chose variants for that grid; the other workloads
    # run with the library's defaults
    tuning.restore_defaults()
    if not args.no_small_configs:
        log("128^2 Jacobi and 512^2 RB-SOR plumes")
        extra.update(bench_small_configs())
        log(f"small configs: {json.dumps({k: extra[k] for k in ('plume_128_jacobi', 'plume_512_rbsor')})}")
    if args.batched_scenes > 0:
        log(f"batched scenes: {args.batched_scenes}")
        extra["batched_scenes"] = bench_batched(args.batched_scenes, 20, args.batch_omega)
        log(f"batched scenes: {json.dumps(extra['batched_scenes'])}")
    if args.diff_steps > 0:
        log(f"differentiable rollout: {args.diff_steps} steps")
        extra["differentiable"] = bench_differentiable(args.diff_steps, args.diff_batch)
        log(f"differentiable rollout: {json.dumps(extra['differentiable'])}")

    line = {
        "metric": "smoke cell-steps/s & %HBM roofline at 2048² MG (mixed vs FP64); grad/fwd ratio",
        "value": m["cell_steps_per_s"],
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64/f32-mixed" if main_var == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": f"plume_{args.N}_mg", "grid": [args.N, args.N],
                   "solver": "multigrid V(2,2) RB-GS, 8 levels, coarsest 50 sweeps, tol 1e-8, "
                             "rediscretized coarse scale" + (
                                 {"exact": "; mixed: FP32 V-cycles, 16x16 level solved exactly",
                                  "exact_1024": "; mixed: FP32 V(2,2) cycles, red-black SOR "
                                                f"omega {args.mixed_omega:g}, to the 32x32 level, "
                                                "solved exactly"}.get(
                                     results[main_var].get("coarsest_solve"), "")
                                 if main_var == "mixed" else ""),
                   "variant": main_var,
                   "coarsest_solve": results[main_var].get("coarsest_solve", "sweeps"),
                   "inner_arithmetic": results[main_var].get("inner_arithmetic")},
        "variants": results,
        "roofline": dict(roof[main_var], copy_kernel_GBps=copy_gbps),
        "roofline_by_variant": roof,
        "cpu_baseline": cpu,
    }
    if accuracy:
        line["mixed_vs_fp64_rel_err"] = accuracy
    if "fp64" in results and "mixed" in results:
        line["mixed_speedup_vs_fp64"] = results["fp64"]["ms_per_step"] / results["mixed"]["ms_per_step"]
    if "differentiable" in extra:
        line["grad_fwd_ratio"] = extra["differentiable"]["fp32"]["grad_fwd_ratio"]
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
