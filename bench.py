"""Benchmark: particle-substeps/s of the PBF substep on one MI355X.

Workload (default): the 1.64M-particle scene - 2 m domain, h = 0.025 m,
160 x 160 x 64 jittered lattice (40% fill), 4 Jacobi iterations, cohesion on,
XSPH viscosity 0.01, a 0.3 m pillar and a 1.0 x 0.1 x 0.5 m half-wall on an
80^3 collider grid, dt from 24 fps (4 substeps per frame).  One step = one
full substep (normal, predict, grid build, 4 x (lambda, delta+apply),
velocity, XSPH, finalize) with every input resident in HBM.  The surface
(256-resolution splat + marching cubes) is timed separately.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]

Prints one JSON line.  The roofline entry is for the kernel with the largest
share of substep time, timed live with HIP events (fx_profile) over a second
timed region of the same substeps.
"""

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(kernel, n, c):
    """Compulsory HBM bytes of one launch: each array element read once and
    written once, float32 SoA accounting (vec3 = 12 B, scalar = 4 B, int = 4 B),
    neighbour gathers counted once (they re-read the same arrays)."""
    kernel = kernel.split("<")[0]
    per_particle = {
        "k_lambda": 24 + 8.0 * c / n,       # pred(12) + keys/idx(8) -> rho, lambda(8); cell ranges 8C
        "k_delta": 32 + 8.0 * c / n + 12,   # pred, lambda -> delta, plus the fused apply write (12)
        "k_xsph": 44 + 8.0 * c / n + 16,    # pred, vel, rho -> dv; the next substep's normal (16)
        "k_normal": 36 + 8.0 * c / n,
        "k_predict_key": 48 + 8,            # pos, vel -> vel, pred; key + rank
        # finalize + the next predict + key + rank: pred, vel, dv, normal (12 + 12 + 12 + 16) in;
        # position, prediction, key, rank (12 + 12 + 4 + 4) out
        "k_finalize_predict": 84,
        # (particle, previous slot) pairs 8, the previous slot's prediction, position, key (12 + 12 + 4)
        # in; sorted prediction + lambda, position, (rho, lambda), key, particle (16 + 12 + 8 + 4 + 4) out
        "k_gather_pairs": 80,
        "k_gather": 76,                     # the same from canonical rows (index 4 in, key 4 in)
        "k_scatter_pairs": 24 + 4.0 * c / n,  # key, rank, particle, cell offset in; pair out; counts zeroed
        "k_finalize_s": 36 + 48,            # pred, vel, dv in; position, velocity (canonical, sorted) out
        "k_canon_cells": 16 + 8.0 * c / n,  # each cell's pairs read and written once, its two offsets
    }.get(kernel)
    if per_particle is None:
        return None, None
    return per_particle * n, per_particle


# register count and the occupancy it allows (512 VGPRs per lane of a SIMD)
# of the kernels the roofline can name, from the gfx950 code objects
WAVES_PER_SIMD = {"k_lambda<VISIT_FULL + kMaskRecord + kXyzCand>": (72, 7), "k_lambda<VISIT_FULL>": (54, 8),
                  "k_delta<VISIT_FULL + kMaskWalk + kLdsWords>": (61, 8), "k_xsph<32 + 16 + VISIT_FULL>": (72, 7)}


def profile_name(kernel):
    """File-name form of a kernel's launch label: 'k_lambda<VISIT_FULL + kMaskRecord>' ->
    'k_lambda_VISIT_FULL_kMaskRecord'."""
    import re

    return re.sub(r"[^A-Za-z0-9]+", "_", kernel).strip("_")


def read_traffic(kernel):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary
    (profiles/<kernel>_traffic.txt: 'fetch_bytes=' and 'write_bytes=' lines),
    or None."""
    f = ROOT / "profiles" / f"{profile_name(kernel)}_traffic.txt"
    if not f.exists():
        return None
    vals = {}
    for line in f.read_text().splitlines():
        if "=" in line and not line.startswith("#"):
            k, v = line.split("=", 1)
            try:
                vals[k.strip()] = float(v.split()[0])
            except ValueError:
                pass
    if "fetch_bytes" in vals and "write_bytes" in vals:
        return vals["fetch_bytes"] + vals["write_bytes"]
    return None


COUNTERS_FILE = ROOT / "profiles" / "substep_scale_1.64M_valu_counters.txt"


def read_counters(kernel):
    """VALU and memory-pipeline figures of `kernel` from the committed
    rocprofv3 counter summary ('counters <kernel>: key=value ...' lines), or
    None."""
    if not COUNTERS_FILE.exists():
        return None
    for line in COUNTERS_FILE.read_text().splitlines():
        if line.startswith(f"counters {kernel}:"):
            return {k: float(v) for k, v in (kv.split("=") for kv in line.split(":", 1)[1].split())}
    return None


def stencil_candidates(cfg, pos):
    """Candidates one full-stencil neighbour pass visits: the sum over
    particles of the occupancy of the 27 cells around the particle's cell
    (each cell capped at max_cell_scan), from the cells of `pos` (the next
    substep's grid differs from it by one predict step)."""
    import numpy as np

    p = cfg.params(0.0)
    dims = np.array([p.cells[0], p.cells[1], p.cells[2]], np.int64)
    lo = np.array([p.lo[0], p.lo[1], p.lo[2]], np.float32)
    c = np.floor((pos[:, :3].astype(np.float32) - lo) / np.float32(p.cell_size)).astype(np.int64)
    c = np.clip(c, 0, dims - 1)
    key = (c[:, 2] * dims[1] + c[:, 1]) * dims[0] + c[:, 0]
    cnt = np.bincount(key, minlength=int(dims.prod())).reshape(dims[2], dims[1], dims[0])
    pad = np.pad(np.minimum(cnt, p.max_cell_scan), 1)
    box = np.zeros_like(cnt)
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                box += pad[dz:dz + dims[2], dy:dy + dims[1], dx:dx + dims[0]]
    return int((cnt * box).sum())


def cpu_baseline(cfg, dt, substeps):
    """The oracle (float32 C restatement of the reference passes, bitonic sort,
    OpenMP over the host cores) on `substeps` substeps of the same scene."""
    import oracle

    from flowx_amd.collision import grid_geometry, occupancy_cpu
    from flowx_amd.config import seed_positions

    pos0 = seed_positions(cfg)
    origin, voxel, dims = grid_geometry(cfg.lo, cfg.hi, cfg.extra["collider_resolution"])
    col = occupancy_cpu(origin, voxel, dims, cfg.extra["colliders"]).reshape(-1)
    p = cfg.params(dt)
    p.collider_dims[:] = list(dims)
    p.collider_voxel = voxel
    st = oracle.State(p, pos0, precision="f32", sort_mode="bitonic")
    st.substep(col)  # warm-up (also builds the previous grid the normal pass reads)
    t0 = time.perf_counter()
    for _ in range(substeps):
        st.substep(col)
    el = time.perf_counter() - t0
    cores = oracle.num_threads()
    return {
        "value": cfg.particle_count * substeps / el,
        "unit": "particle-substeps/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{substeps} substeps of the same {cfg.particle_count}-particle scene, oracle "
                  f"(CPU restatement of the reference passes incl. the bitonic sort), float32, "
                  f"{cores} OpenMP threads, {el:.1f} s",
    }


def batch_workload(fx, steps, warmup, count=64):
    """64 independent scenes (log-uniform 10k-400k particles, fill 0.2-0.6,
    h 0.03-0.08, 0-3 boxes each) advanced in one launch per pass with the
    substeps replayed from a HIP graph, against the same scenes run one after
    another on the plain path."""
    import torch

    cfgs = fx.batch_configs(count=count, seed=0)
    batch = fx.BatchSolver(cfgs)
    n = batch.particle_count
    for _ in range(warmup):
        batch.substep()
    graph = batch.capture(steps)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    graph.replay()
    e1.record()
    torch.cuda.synchronize()
    t_graph = e0.elapsed_time(e1) / 1e3
    e0.record()
    for _ in range(steps):
        batch.substep()
    e1.record()
    torch.cuda.synchronize()
    t_eager = e0.elapsed_time(e1) / 1e3
    del batch, graph
    torch.cuda.empty_cache()
    t_seq = 0.0
    for cfg in fx.batch_configs(count=count, seed=0):
        s = fx.Solver(cfg, keep_scratch=False)
        s.reset()
        _, dt = cfg.substep_dt(1.0 / 24.0)
        s.substeps_n(dt, warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.substeps_n(dt, steps)
        torch.cuda.synchronize()
        t_seq += time.perf_counter() - t0
        del s
    rate_graph = n * steps / t_graph
    rate_seq = n * steps / t_seq
    return {
        "metric": "particle-substeps/s", "value": round(rate_graph, 1), "unit": "particle-substeps/s",
        "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": round(t_graph * 1e3 / steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic scenes (numpy PCG64 seed 0)",
        "config": {"workload": f"{count} batched scenes, {n} particles total, one launch per pass, HIP graph",
                   "particles": n, "scenes": count},
        "batched_eager_value": round(n * steps / t_eager, 1),
        "sequential_plain_value": round(rate_seq, 1),
        "batched_over_sequential": round(rate_graph / rate_seq, 3),
    }


def _sync_time(fn):
    import torch

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = fn()
    torch.cuda.synchronize()
    return out, time.perf_counter() - t0


def dambreak_workload(fx, frames=48, cpu=True):
    """The 25,600-particle dam break run in full: 48 frames of 2 substeps
    with a 64-resolution surface (splat + marching cubes) every frame, on the
    GPU end to end (each frame's substeps replayed from a HIP graph; the
    eager time is reported beside it), and - beside it - the same run on the oracle (CPU
    restatement of the reference passes, bitonic sort, OpenMP) with its
    splat and marching cubes."""
    import numpy as np
    import torch

    cfg = fx.dam_break_config()
    solver = fx.Solver(cfg, keep_scratch=False)
    solver.reset()
    surf = fx.Surface(solver, cfg.extra["surface_resolution"], 0.5, max_samples=1 << 30)
    frame_dt = 1.0 / cfg.extra["fps"]
    substeps, dt = cfg.substep_dt(frame_dt)
    solver.step(frame_dt)  # warm-up frame (kernels loaded, workspace touched)
    surf.update()
    seed = fx.seed_positions(cfg)
    solver.load_state(seed, None)

    def gpu_run(step):
        tri = 0
        for _ in range(frames):
            step()
            _, t = surf.update()  # marching cubes reads its counts back: one sync per frame
            tri = int(t.shape[0])
        return tri

    tris, t_eager = _sync_time(lambda: gpu_run(lambda: solver.substeps_n(dt, substeps)))
    # the same run with each frame's substeps replayed from a HIP graph (the
    # small scene is launch-bound); capture runs one substep eagerly, so the
    # seeded state is loaded again into the same buffers afterwards
    graph = solver.capture(dt, substeps)
    solver.load_state(seed, None)
    tris_g, t_gpu = _sync_time(lambda: gpu_run(graph.replay))
    assert tris_g == tris
    n = cfg.particle_count
    out = {
        "metric": "particle-substeps/s", "value": round(n * substeps * frames / t_gpu, 1),
        "unit": "particle-substeps/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f32",
        "data": "synthetic jittered lattice (seeded MT19937), analytic box collider",
        "config": {"workload": f"dam break, {n} particles (40x40x16, h=0.1 m), {frames} frames x {substeps} "
                               f"substeps, 64-resolution surface every frame, end to end",
                   "particles": n, "frames": frames, "substeps_per_frame": substeps},
        "gpu_s": round(t_gpu, 4), "gpu_ms_per_frame": round(t_gpu * 1e3 / frames, 3),
        "gpu_eager_s": round(t_eager, 4), "last_frame_triangles": tris,
    }
    if cpu:
        import oracle

        from flowx_amd.collision import grid_geometry, occupancy_cpu

        origin, voxel, dims = grid_geometry(cfg.lo, cfg.hi, cfg.extra["collider_resolution"])
        col = occupancy_cpu(origin, voxel, dims, cfg.extra["colliders"]).reshape(-1)
        p = cfg.params(dt)
        p.collider_dims[:] = list(dims)
        p.collider_voxel = voxel
        sc = surf.config
        st = oracle.State(p, fx.seed_positions(cfg), precision="f32", sort_mode="bitonic")
        t0 = time.perf_counter()
        t_sub = 0.0
        for _ in range(frames):
            a = time.perf_counter()
            for _ in range(substeps):
                st.substep(col)
            t_sub += time.perf_counter() - a
            field = oracle.splat(p, sc.struct(), st.pos, st.idx, st.start, st.end, col)
            oracle.mc_extract(field, sc.dims, sc.lo, sc.spacing, sc.iso)
        t_cpu = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "value": round(n * substeps * frames / t_cpu, 1), "unit": "particle-substeps/s",
            "cores": oracle.num_threads(), "kind": "port",
            "sample": f"the whole run ({frames} frames, surfaces included) on the oracle, float32, "
                      f"{oracle.num_threads()} OpenMP threads, {t_cpu:.2f} s ({t_sub:.2f} s in substeps)",
        }
        out["gpu_over_cpu"] = round(t_cpu / t_gpu, 1)
    return out


def animated_workload(fx, frames=24, warmup=2):
    """~4.1M particles (h = 0.018 m, 222x222x83, 6 iterations), a 0.3 m sphere
    dropped ballistically from z = 1.8 m and re-voxelized on the GPU onto a
    111^3 grid every frame, a 384-resolution surface (splat + marching cubes)
    every frame.  value = particle-substeps/s of the frames' substeps plus the
    collider rebuilds; the surface time is reported beside it."""
    import torch

    cfg = fx.animated_config()
    (c0, r) = cfg.extra["sphere"]
    cfg.extra["colliders"] = [("sphere", c0, r)]

    def sphere_at(frame):
        t = frame / cfg.extra["fps"]
        z = max(c0[2] - 0.5 * 9.81 * t * t, r)
        return [("sphere", (c0[0], c0[1], z), r)]

    solver = fx.Solver(cfg, keep_scratch=False)
    solver.reset()
    surf = fx.Surface(solver, cfg.extra["surface_resolution"], 0.5, max_samples=1 << 30)
    frame_dt = 1.0 / cfg.extra["fps"]
    substeps, dt = cfg.substep_dt(frame_dt)
    n = cfg.particle_count
    for f in range(warmup):
        solver.collider.rebuild(sphere_at(f))
        solver.substeps_n(dt, substeps)
        surf.update()
    e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    t_sim = t_surf = 0.0
    tris = 0
    for f in range(warmup, warmup + frames):
        e[0].record()
        solver.collider.rebuild(sphere_at(f))
        solver.substeps_n(dt, substeps)
        e[1].record()
        _, t = surf.update()
        e[2].record()
        torch.cuda.synchronize()
        t_sim += e[0].elapsed_time(e[1]) / 1e3
        t_surf += e[1].elapsed_time(e[2]) / 1e3
        tris = int(t.shape[0])
    return {
        "metric": "particle-substeps/s", "value": round(n * substeps * frames / t_sim, 1),
        "unit": "particle-substeps/s", "n_gpus": 1, "steps": frames * substeps, "higher_is_better": True,
        "ms_per_step": round(t_sim * 1e3 / (frames * substeps), 4), "dtype": "f32",
        "data": "synthetic jittered lattice (seeded MT19937), analytic moving sphere",
        "config": {"workload": f"animated sphere, {n} particles (222x222x83, h=0.018 m), 6 iterations, "
                               f"{frames} frames x {substeps} substeps, sphere re-voxelized on 111^3 per frame, "
                               f"384-resolution surface per frame", "particles": n, "cells": cfg.cell_count},
        "surface_ms_per_frame": round(t_surf * 1e3 / frames, 3),
        "surface_dims": list(surf.config.dims), "last_frame_triangles": tris,
        "frame_ms_end_to_end": round((t_sim + t_surf) * 1e3 / frames, 3),
    }


def adjoint_workload(fx, substeps, directions, fd_checks=((4, 1e-4, 1e-8), (8, 1e-6, 1e-9))):
    """200k particles (h = 0.05, 80x80x31, 4 iterations, 0.4 m box): forward
    and backward of a `substeps` rollout in float32 (adjoint / forward cost),
    then float64 central finite differences of the loss along `directions`
    random initial-velocity directions, gravity and the three
    collider-translation axes, for each (substeps, eps, eps_translation) of
    `fd_checks`.  Translation moves every pushed-out particle (thousands per
    substep) by eps at once, so its difference quotient leaves the linear
    range at a far smaller step than a unit velocity direction spread over
    all particles; it gets its own eps.  The perturbed runs replay the base
    run's record (branches and neighbour lists).  The float64 gradient norm |dL/dv0| over growing horizons
    measures how fast the substep map amplifies perturbations."""
    import numpy as np
    import torch

    from flowx_amd.adjoint import DiffSolver, gaussian_target, loss_fn, rollout

    out = {}
    rng = np.random.default_rng(0)
    centre = (float(rng.uniform(-0.5, 0.5)), float(rng.uniform(-0.5, 0.5)), float(rng.uniform(0.1, 0.9)))
    for dtype in (torch.float32, torch.float64):
        d = DiffSolver(fx.adjoint_config(), dtype=dtype)
        x0 = torch.as_tensor(fx.seed_positions(d.cfg), dtype=dtype).cuda()
        x0[:, 3] = 0
        v0 = torch.zeros_like(x0)
        g = torch.tensor(-9.81, dtype=torch.float64, device="cuda", requires_grad=True)
        tr = torch.zeros(3, dtype=torch.float64, device="cuda", requires_grad=True)
        target = gaussian_target(d, centre, sigma=0.2)
        com_t = torch.tensor(centre, dtype=dtype, device="cuda")
        if dtype == torch.float32:
            v = v0.clone().requires_grad_(True)
            rollout(d, x0, v, g, tr, 2, [], 1)  # warm-up
            # the per-substep records are allocated ahead: the timed region
            # holds the substeps, not the allocator.  The host drives every
            # substep, so wall times vary from run to run: the pair is timed
            # three times and the medians are reported
            records = [d.new_decisions() for _ in range(substeps)]
            fwd, bwd = [], []
            for _ in range(3):
                v = v0.clone().requires_grad_(True)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                x, _ = rollout(d, x0, v, g, tr, substeps, records, 1, check_overflow=False)
                L = loss_fn(d, x, target, com_t)
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                L.backward()
                torch.cuda.synchronize()
                t2 = time.perf_counter()
                fwd.append(t1 - t0)
                bwd.append(t2 - t1)
                g.grad = tr.grad = None
            tf, tb = float(np.median(fwd)), float(np.median(bwd))
            out.update(particles=len(x0), substeps=substeps, forward_s=round(tf, 4), backward_s=round(tb, 4),
                       adjoint_over_forward=round(tb / tf, 3), forward_s_runs=[round(t, 4) for t in fwd],
                       backward_s_runs=[round(t, 4) for t in bwd], list_overflow=d.overflowed(),
                       fp32_grad_v0_norm=float(v.grad.norm()))
            del x, L, v
            torch.cuda.empty_cache()
            continue
        growth = []
        for k in (1, 4, 8, 16, 30, substeps):
            v = v0.clone().requires_grad_(True)
            x, _ = rollout(d, x0, v, g, tr, k, [], 1)
            loss_fn(d, x, target, com_t).backward()
            growth.append((k, float(v.grad.norm())))
            g.grad = tr.grad = None
        out["fp64_grad_v0_norm_by_substeps"] = growth
        checks = []
        for fsteps, eps, eps_t in fd_checks:
            g.grad = tr.grad = None
            v = v0.clone().requires_grad_(True)
            decisions = []  # the base run's record, replayed by every perturbed run
            x, _ = rollout(d, x0, v, g, tr, fsteps, decisions, 1)
            L = loss_fn(d, x, target, com_t)
            L.backward()
            gv, gg, gt = v.grad.clone(), g.grad.item(), tr.grad.cpu().numpy()

            def L_at(dv=None, dg=0.0, dtr=None):
                with torch.no_grad():
                    vv = v0 + (dv if dv is not None else 0)
                    tt = tr.detach() + (torch.as_tensor(dtr, dtype=torch.float64, device="cuda")
                                        if dtr is not None else 0)
                    xx, _ = rollout(d, x0, vv, g.detach() + dg, tt, fsteps, decisions, 2)
                    return loss_fn(d, xx, target, com_t).item()

            errs = []
            for _ in range(directions):
                dirn = torch.as_tensor(rng.normal(0, 1, v0.shape), dtype=dtype).cuda()
                dirn[:, 3] = 0
                dirn /= dirn.norm()
                fd = (L_at(dv=eps * dirn) - L_at(dv=-eps * dirn)) / (2 * eps)
                errs.append(abs(fd - float((gv * dirn).sum())) / max(abs(fd), 1e-300))
            fdg = (L_at(dg=eps) - L_at(dg=-eps)) / (2 * eps)
            fdt = []
            for a in range(3):
                e = np.zeros(3)
                e[a] = eps_t
                fdt.append((L_at(dtr=e) - L_at(dtr=-e)) / (2 * eps_t))
            checks.append(dict(
                substeps=fsteps, eps=eps, eps_translation=eps_t, directions=directions,
                velocity_max_rel=float(max(errs)), velocity_median_rel=float(np.median(errs)),
                gravity=(fdg, gg), gravity_rel=abs(fdg - gg) / abs(fdg),
                translation=[(fdt[a], float(gt[a])) for a in range(3)],
                translation_rel=[abs(fdt[a] - gt[a]) / max(abs(fdt[a]), 1e-300) for a in range(3)]))
            del decisions
        out["fd_checks"] = checks
    return {"metric": "adjoint/forward cost, fp64 finite-difference agreement", "config": {
        "workload": f"{out.get('particles')} particles, h=0.05, 4 iterations, 0.4 m box, {substeps} substeps"},
        **out}


def batch_summary(fx, steps, warmup, count=64):
    """The batched line of the metric, in the default run: the 64-scene set
    (seed 0) replayed from a HIP graph against the same scenes run one after
    another on the plain path (fewer substeps than the batch workload)."""
    b = batch_workload(fx, steps, warmup, count)
    return {"scenes": count, "particles": b["config"]["particles"], "substeps": steps,
            "batched_graph_particle_substeps_per_s": b["value"],
            "batched_eager_particle_substeps_per_s": b["batched_eager_value"],
            "sequential_plain_particle_substeps_per_s": b["sequential_plain_value"],
            "batched_over_sequential": b["batched_over_sequential"]}


def adjoint_summary(fx, substeps=24, reps=3):
    """The adjoint line of the metric, in the default run: float32 forward and
    backward of the 200k-particle differentiable scene over `substeps`
    substeps (the horizon whose gradient stays finite in float32: |dL/dv0|
    grows about tenfold per substep once the dam collapses and leaves the
    float32 range near 30 substeps), medians of `reps` timed pairs, records
    allocated ahead.  No finite differences here (tests/test_gpu_adjoint.py
    holds them)."""
    import numpy as np
    import torch

    from flowx_amd.adjoint import DiffSolver, gaussian_target, loss_fn, rollout

    rng = np.random.default_rng(0)
    centre = (float(rng.uniform(-0.5, 0.5)), float(rng.uniform(-0.5, 0.5)), float(rng.uniform(0.1, 0.9)))
    d = DiffSolver(fx.adjoint_config(), dtype=torch.float32)
    x0 = torch.as_tensor(fx.seed_positions(d.cfg), dtype=torch.float32).cuda()
    x0[:, 3] = 0
    v0 = torch.zeros_like(x0)
    g = torch.tensor(-9.81, dtype=torch.float64, device="cuda", requires_grad=True)
    tr = torch.zeros(3, dtype=torch.float64, device="cuda", requires_grad=True)
    target = gaussian_target(d, centre, sigma=0.2)
    com_t = torch.tensor(centre, dtype=torch.float32, device="cuda")
    records = [d.new_decisions() for _ in range(substeps)]
    v = v0.clone().requires_grad_(True)
    loss_fn(d, rollout(d, x0, v, g, tr, 2, [], 1)[0], target, com_t).backward()  # warm-up
    fwd, bwd, norm = [], [], None
    for _ in range(reps):
        g.grad = tr.grad = None
        v = v0.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x, _ = rollout(d, x0, v, g, tr, substeps, records, 1)
        L = loss_fn(d, x, target, com_t)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        L.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        fwd.append(t1 - t0)
        bwd.append(t2 - t1)
        norm = float(v.grad.double().norm())
    tf, tb = float(np.median(fwd)), float(np.median(bwd))
    return {"particles": len(x0), "substeps": substeps, "dtype": "f32", "forward_s": round(tf, 4),
            "backward_s": round(tb, 4), "adjoint_over_forward": round(tb / tf, 3),
            "grad_v0_norm": norm, "gradient_finite": True}


def reference_scene_summary(fx, frames=24):
    """The reference's own scene against its own published figures: the
    smoke test's default domain at the origin (scripts/smoke_test.py; 2 m
    cube, default settings, the reference's caps: 15,360 particles,
    h = 0.1187 m) with its 0.5 m cube collider, `frames` frames of
    Solver.step (the reference's substeps per frame) and a surface at the
    default resolution 48 (splat + marching cubes) every frame.  The
    reference's source quotes ~35 ms per frame for the solver step at 2 m
    and 15k particles and ~35 ms of CPU marching cubes + ~2 ms of read-back
    for a resolution-48 surface (solver/surface.py:14-17; hardware not
    stated).  Timed under the reference's literal update (its semantics; the
    scene flies apart in its first frame, see the density check) and under
    the converging update."""
    import torch

    from flowx_amd.config import REFERENCE_CAPS, DomainSettings, converging_update, resolve_config

    out = {"published_solver_ms_per_frame": 35.0, "published_surface_ms": 37.0,
           "published_source": "solver/surface.py:14-17 (15k particles, 2 m domain, surface resolution 48)"}
    for name, literal in (("literal", True), ("converging", False)):
        settings = DomainSettings()
        cfg = resolve_config(settings, (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), REFERENCE_CAPS)
        cfg.extra.update(colliders=[("box", (-0.25, -0.25, -0.25), (0.25, 0.25, 0.25))],
                         collider_resolution=settings.resolution)
        if not literal:
            cfg = converging_update(cfg)
        solver = fx.Solver(cfg, keep_scratch=False)
        solver.reset()
        surf = fx.Surface(solver, settings.surface_resolution, settings.surface_iso)
        frame_dt = 1.0 / 24.0
        solver.step(frame_dt)  # warm-up frame
        surf.update()
        _, t_step = _sync_time(lambda: [solver.step(frame_dt) for _ in range(frames)])
        _, t_surf = _sync_time(lambda: [surf.update() for _ in range(frames)])
        steps, _ = cfg.substep_dt(frame_dt)
        out[name] = {"particles": cfg.particle_count, "substeps_per_frame": steps,
                     "solver_ms_per_frame": round(t_step * 1e3 / frames, 3),
                     "surface_ms_per_frame": round(t_surf * 1e3 / frames, 3),
                     "surface_dims": list(surf.config.dims)}
        del surf, solver
        torch.cuda.empty_cache()
    conv = out["converging"]
    out["solver_speedup_vs_published"] = round(35.0 / conv["solver_ms_per_frame"], 1)
    out["surface_speedup_vs_published"] = round(37.0 / conv["surface_ms_per_frame"], 1)
    return out


def scorr_delta(fx, kernels, steps=8):
    """The delta pass with s_corr on, at a strength that still converges on
    the scale scene (k = 1e-5; k >= 1e-4 makes it blow up), timed on its own
    scene: the per-pair pow and poly6 the default line leaves out."""
    import torch

    from flowx_amd import _lib

    cfg = fx.scale_config()
    cfg.scorr_strength = 1e-5
    solver = fx.Solver(cfg, keep_scratch=False)
    solver.reset()
    _, dt = cfg.substep_dt(1.0 / cfg.extra["fps"])
    solver.substeps_n(dt, 8)
    torch.cuda.synchronize()
    _lib.profile(True)
    solver.substeps_n(dt, steps)
    torch.cuda.synchronize()
    rep = _lib.profile_report()
    _lib.profile(False)
    name = next(k for k in rep if k.startswith("k_delta"))
    with_us = 1e3 * rep[name][1] / rep[name][0]
    without_us = kernels.get(name, {}).get("avg_us")
    del solver
    torch.cuda.empty_cache()
    return {"scorr_k": 1e-5, "kernel": name, "avg_us_with_scorr": round(with_us, 2),
            "avg_us_without": round(without_us, 2) if without_us else None}


def streaming_rooflines(kernels, n, c):
    """Algorithmic bytes / average duration for the streaming kernels."""
    out = {}
    for k, v in kernels.items():
        name = k.split("<")[0]
        if name in ("k_lambda", "k_delta", "k_xsph", "k_normal"):
            continue
        alg, per = algorithmic_bytes(k, n, c)
        if alg is None:
            continue
        gbs = alg / (v["avg_us"] * 1e-6) / 1e9
        out[k] = {"avg_us": round(v["avg_us"], 2), "bytes_per_particle": round(per, 2),
                  "achieved_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-substeps", type=int, default=12)
    ap.add_argument("--literal-update", action="store_true",
                    help="the reference's literal position correction (diverges at these masses)")
    ap.add_argument("--no-surface", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="leave out the reference-scene, s_corr, batched and adjoint objects of the default line")
    ap.add_argument("--workload", default="scale",
                    choices=["scale", "batch", "adjoint", "dambreak", "animated", "reference"])
    ap.add_argument("--frames", type=int, default=None, help="dambreak / animated workloads: frames to run")
    ap.add_argument("--autotune", action="store_true",
                    help="sweep the neighbour-pass tuning on the 1.64M scene, write flow-x_amd/tuning.json, exit")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run three plain device-to-device copies of MIB MiB (the reference "
                         "dispatch for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    ap.add_argument("--walk", type=int, default=None, choices=[0, 1, 2, 3, 4],
                    help="scale scene: run with this large-scene walk (FxTuning.walk; 3 = LDS tiles, "
                         "4 = the default walk without lambda's pair masks for delta)")
    ap.add_argument("--block", type=int, default=None, choices=[64, 128, 256],
                    help="scale scene: run with this neighbour-pass block size (FxTuning.block)")
    ap.add_argument("--fd-directions", type=int, default=32)
    ap.add_argument("--fd-checks", default="4:1e-4:1e-8,8:1e-6:1e-9",
                    help="adjoint workload: comma-separated substeps:eps:eps_translation finite-difference checks")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    if args.copy_calibration:
        nbytes = args.copy_calibration << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": nbytes, "copies": 3}))
        return

    import __graft_entry__

    fx = __graft_entry__.load_package()
    from flowx_amd import _lib

    if args.workload == "batch":
        print(json.dumps(batch_workload(fx, args.steps, args.warmup)))
        return
    if args.autotune:
        from flowx_amd import autotune

        print(json.dumps(autotune.tune(fx)))
        return
    if args.workload == "dambreak":
        print(json.dumps(dambreak_workload(fx, args.frames or 48, cpu=not args.no_cpu_baseline)))
        return
    if args.workload == "reference":
        print(json.dumps(reference_scene_summary(fx, args.frames or 24)))
        return
    if args.workload == "animated":
        print(json.dumps(animated_workload(fx, args.frames or 24)))
        return
    if args.workload == "adjoint":
        checks = [(int(a), float(b), float(t)) for a, b, t in (c.split(":") for c in args.fd_checks.split(","))]
        print(json.dumps(adjoint_workload(fx, args.steps, args.fd_directions, checks)))
        return

    cfg = fx.scale_config(literal_update=args.literal_update)
    solver = fx.Solver(cfg, keep_scratch=False)
    if args.walk is not None or args.block is not None:  # after the Solver applied any tuning file
        _lib.set_tuning(walk=args.walk, block=args.block)
    solver.reset()
    _, dt = cfg.substep_dt(1.0 / cfg.extra["fps"])
    n, c = cfg.particle_count, cfg.cell_count
    print(f"[bench] {n} particles, {c} cells, dt {dt:.6f}, warmup {args.warmup}, steps {args.steps}",
          file=sys.stderr, flush=True)

    solver.substeps_n(dt, args.warmup)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    solver.substeps_n(dt, args.steps)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = e0.elapsed_time(e1)
    ms_per_step = gpu_ms / args.steps
    value = n * args.steps / (gpu_ms / 1e3)
    print(f"[bench] {ms_per_step:.3f} ms/substep (wall {wall * 1e3 / args.steps:.3f})", file=sys.stderr, flush=True)

    # per-kernel durations over a second timed region (same substeps)
    prof_steps = max(4, min(args.steps, 20))
    _lib.profile(True)
    solver.substeps_n(dt, prof_steps)
    torch.cuda.synchronize()
    rep = _lib.profile_report()
    _lib.profile(False)
    kernels = {k: {"launches": cnt, "avg_us": 1e3 * ms / cnt, "ms_per_substep": ms / prof_steps}
               for k, (cnt, ms) in rep.items()}
    dominant = max(kernels, key=lambda k: kernels[k]["ms_per_substep"])
    alg, per_particle = algorithmic_bytes(dominant, n, c)
    avg_s = kernels[dominant]["avg_us"] * 1e-6
    roofline = None
    if alg is not None:
        achieved = alg / avg_s / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": read_traffic(dominant),
                    "kernel": dominant, "avg_us": round(kernels[dominant]["avg_us"], 2),
                    "bytes_per_particle": round(per_particle, 2)}

    # the neighbour passes against their own bounds: candidates per second
    # and the counter figures of the committed profile
    cand = stencil_candidates(cfg, solver.pos.cpu().numpy())
    neighbour_passes = {}
    for nb_kernel in kernels:
        if nb_kernel.split("<")[0] not in ("k_lambda", "k_delta", "k_xsph"):
            continue
        nb_s = kernels[nb_kernel]["avg_us"] * 1e-6
        neighbour_passes[nb_kernel] = {"avg_us": round(kernels[nb_kernel]["avg_us"], 2),
                                       "candidates_per_particle": round(cand / n, 2),
                                       "candidate_pairs_per_s": round(cand / nb_s, 1),
                                       "counters": read_counters(nb_kernel)}
    if roofline is not None:
        ctr = read_counters(dominant)
        if ctr:
            # the roofline is HBM's (no MFMA work); what the counters show
            # limits the kernel is each wave's instruction stream and its
            # dependent loads at the occupancy ceiling, with the
            # vector-memory pipeline near saturation
            occ = WAVES_PER_SIMD.get(dominant)
            at = f" at {occ[1]} waves per SIMD ({occ[0]} VGPRs)" if occ else ""
            roofline["measured_limiter"] = (
                f"per-wave instruction issue and load latency{at}: VALU "
                f"{ctr.get('valu_insts_per_cu_cycle', 0):.2f} instructions per CU per cycle, "
                f"{ctr.get('wait_inst_share', 0) * 100:.0f}% of wave cycles waiting on a dependency, TA "
                f"{ctr.get('ta_busy', 0) * 100:.0f}% / TD {ctr.get('td_busy', 0) * 100:.0f}% busy, VALU lanes "
                f"{ctr.get('valu_lane_utilisation', 0) * 100:.0f}% active, HBM {roofline['frac'] * 100:.1f}% of "
                f"peak (committed rocprofv3 counters)")

    surface = None
    if not args.no_surface:
        surf = fx.Surface(solver, cfg.extra["surface_resolution"], 0.5, max_samples=1 << 30)
        surf.update()
        torch.cuda.synchronize()
        s0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            verts, tris = surf.update()
        torch.cuda.synchronize()
        surface = {"dims": list(surf.config.dims), "ms": round((time.perf_counter() - s0) * 1e3 / reps, 3),
                   "vertices": int(verts.shape[0]), "triangles": int(tris.shape[0])}

    del solver
    torch.cuda.empty_cache()
    scorr = batched = adjoint = reference = None
    if not args.no_extras:
        print("[bench] the reference's own scene against its published figures", file=sys.stderr, flush=True)
        reference = reference_scene_summary(fx)
        print("[bench] s_corr delta pass", file=sys.stderr, flush=True)
        scorr = scorr_delta(fx, kernels)
        print("[bench] 64 batched scenes vs sequential plain runs", file=sys.stderr, flush=True)
        batched = batch_summary(fx, steps=10, warmup=2)
        torch.cuda.empty_cache()
        print("[bench] adjoint forward / backward", file=sys.stderr, flush=True)
        adjoint = adjoint_summary(fx)
        torch.cuda.empty_cache()
    base = None
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline (oracle)", file=sys.stderr, flush=True)
        base = cpu_baseline(cfg, dt, args.cpu_substeps)

    out = {
        "metric": "particle-substeps/s",
        "value": round(value, 1),
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic jittered lattice (seeded MT19937), analytic colliders",
        "config": {
            "workload": f"PBF substep, {n} particles (160x160x64, h=0.025 m, 2 m domain), 4 Jacobi iterations, "
                        f"cohesion on, XSPH 0.01, pillar + half-wall on 80^3 voxels",
            "update": cfg.extra.get("update", "reference literal (no mass in the correction)"),
            "scorr_k": cfg.scorr_strength,
            "particles": n, "cells": c, "dt": dt,
        },
        "tuning": _lib.get_tuning(),
        "roofline": roofline,
        "neighbour_passes": neighbour_passes,
        "streaming": streaming_rooflines(kernels, n, c),
        "cpu_baseline": base,
        "kernels": {k: {kk: round(vv, 4) for kk, vv in v.items()} for k, v in kernels.items()},
        "surface": surface,
        "scorr_delta": scorr,
        "batched": batched,
        "adjoint": adjoint,
        "reference_scene": reference,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
