#!/usr/bin/env python
"""Benchmark of the per-step contact pipeline on one MI355X.

Workload (default): 4096 independent mixed worlds (body counts uniform in
[32, 512], seed 1; boxes / circles / convex polygons 40/30/30 above a static
edge), dt = 1/60, 8 velocity and 3 position iterations, continuous physics
off.  A step is one World.step of every world.

    python bench.py [--gpus 1] [--steps 1000] [--warmup 10]

Prints one JSON line: value = dynamic bodies x timed steps / wall seconds of
the timed loop (FP32-bulk pipeline, state resident in HBM), plus the FP64
pipeline's rate, the dominant kernel's roofline, one-step errors against the
FP64 oracle on a sample of worlds, long-horizon statistics, and the oracle's
rate on the host cores for the same worlds and step count.  Timed steps are
launched with HIP events between their kernels (for the per-kernel averages);
the rate of HIP-graph-captured steps, measured right after, is reported
beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

from pypybox2d_amd import builder, layout as L, scenes  # noqa: E402

HBM_PEAK_GBS = 8000.0
GRAPH_CHECK_STEPS = 100
KERNELS = ("k_collide", "k_islands", "k_solve", "k_sync", "k_pairs")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload_counts(state: L.BatchArrays):
    """Per-step unit counts of a state: bodies, fixtures, live and touching contacts."""
    bi = state.body_i
    ci = state.con_i
    dyn = int(np.count_nonzero(bi[:, L.BI_TYPE] == L.DYNAMIC))
    touching = int(np.count_nonzero(ci[:, L.CI_FLAGS] & L.CF_TOUCHING))
    points = int(ci[(ci[:, L.CI_FLAGS] & L.CF_TOUCHING) != 0, L.CI_PC].sum())
    island_bodies = int(np.count_nonzero(bi[:, L.BI_ISLAND] >= 0))
    nverts = state.fix_i[:, L.FI_NVERTS]
    shape = state.fix_i[:, L.FI_SHAPE]
    geo = np.where(shape == L.SHAPE_POLYGON, 16 * nverts + 4, np.where(shape == L.SHAPE_CIRCLE, 12, 20))
    return {"bodies": int(bi.shape[0]), "dynamic_bodies": dyn, "fixtures": int(state.fix_i.shape[0]),
            "contacts": int(ci.shape[0]), "touching": touching, "points": points,
            "island_bodies": island_bodies, "geometry_bytes": int(geo.sum())}


def algorithmic_bytes(kernel: str, c: dict, fp64: bool) -> float:
    """Minimum HBM bytes one launch of `kernel` must move: the per-unit figures below times the
    counts of the end state."""
    t = 8 if fp64 else 4
    man = 4 + 4 * 2 * t + 4 * t + 8      # state word + ln,lp,p0,p1 + 4 impulses + 2 ids
    if kernel == "k_collide":
        # per contact: pair + manifold read + write; per body: xf and fat AABB and geometry once
        return c["contacts"] * (8 + 2 * man) + c["dynamic_bodies"] * (32 + 32) + c["geometry_bytes"]
    if kernel == "k_solve":
        # per island body: c, a, v, w read + write, c0/a0 write, constants; per touching contact:
        # pair + manifold read, impulses write
        return c["island_bodies"] * (48 + 48 + 24 + 4 * t + 16) + c["touching"] * (8 + man + 4 * t)
    if kernel == "k_islands":
        return c["contacts"] * (8 + 4) + c["bodies"] * 8 + c["touching"] * 8
    if kernel == "k_sync":
        return c["fixtures"] * (32 + 32) + c["island_bodies"] * (56 + 32) + c["geometry_bytes"]
    if kernel == "k_pairs":
        return c["contacts"] * 12 + c["fixtures"] * 4
    return 0.0


def run_gpu(batch: L.BatchArrays, precision: str, steps: int, warmup: int, tune: bool = True,
            solver_times: bool = True, graph_check: bool = True):
    """Times `steps` steps and the average duration of each kernel over them.

    Every timed step is launched with HIP events between its kernels on the
    library's stream (b2g_profile_step), so the per-kernel averages cover exactly
    the timed launches.  Every 200 steps the launch-shape tuner takes the next
    steps, one per candidate shape (real steps, timed with the rest).
    """
    from pypybox2d_amd._native import Engine
    from pypybox2d_amd.autotune import LaunchTuner
    eng = Engine(precision)
    eng.upload(batch)
    tuner = LaunchTuner(eng) if tune else None
    log("[%s] uploaded; warmup %d steps" % (precision, warmup))
    if warmup > 0:
        eng.step_n(warmup, graph=True)
    eng.sync()
    log("[%s] timing %d steps%s" % (precision, steps, " (launch shape tuned at the start, after 50, 50 and 100 steps, then every 200)" if tune else ""))
    sums = np.zeros(5)
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        left = steps - done
        if tuner is not None and left > tuner.steps_per_tune:
            tuner.tune(1.0 / 60, 8, 3)
            sums += tuner.last_ms.sum(axis=0)
            done += tuner.steps_per_tune
            n = min(tuner.interval() - tuner.steps_per_tune, steps - done)
        else:
            n = min(200, left)
        seg = None
        if n > 0:
            seg = eng.profile(n)
            sums += seg * n
            done += n
        log("[%s] %d/%d steps, %.2f s%s%s" % (
            precision, done, steps, time.perf_counter() - t0,
            (", launch shape %s (solve ms per candidate %s)" % (tuner.best, np.round(tuner.last_times, 2).tolist()))
            if tuner and tuner.last_times is not None else "",
            (", solve %.2f ms over the next %d steps" % (seg[2], n)) if seg is not None else ""))
    eng.sync()
    elapsed = time.perf_counter() - t0
    state = eng.download()
    # the same launch sequence captured in a HIP graph, continuing from here
    graph_ms = None
    if graph_check:
        g0 = time.perf_counter()
        eng.step_n(GRAPH_CHECK_STEPS, graph=True)
        eng.sync()
        graph_ms = 1e3 * (time.perf_counter() - g0) / GRAPH_CHECK_STEPS
    # one more (untimed) step with per-island timing, for the solver's latency report (last,
    # since switching the timing on drops the captured graph)
    solver = None
    if solver_times:
        eng.set_param("island_times", 1)
        eng.profile(1)
        solver = solver_report(eng.island_times(), eng.get_param("solve_waves_per_cu"))
    eng.close()
    return elapsed, sums / steps, state, (dict(tuner.best) if tuner and tuner.best else None), graph_ms, solver


def solver_report(rows: np.ndarray, waves_per_cu: int) -> dict:
    """The island solver as occupancy plus per-island latency (it is bound by the dependency
    chains inside islands, not by bytes): one step's islands grouped by the lanes that solved
    them (a team's size; 1: one lane per island; 0: a team island solved by one lane)."""
    start, end = rows[:, 0], rows[:, 1]
    lat = (end - start) / 1e3
    span = (end.max() - start.min()) / 1e3
    groups = {}
    for lanes in sorted(set(rows[:, 4].tolist()), reverse=True):
        m = rows[:, 4] == lanes
        groups[str(int(lanes))] = {"islands": int(m.sum()), "contacts": [int(rows[m, 2].min()), int(rows[m, 2].max())],
                                   "sweep_period": [int(rows[m, 5].min()), int(rows[m, 5].max())],
                                   "latency_us": {"mean": float(lat[m].mean()), "p50": float(np.median(lat[m])),
                                                  "max": float(lat[m].max())}}
    c = int(np.argmax(end))
    return {"islands": int(rows.shape[0]), "span_us": float(span), "mean_islands_in_flight": float(lat.sum() / span),
            "resident_wavefronts_per_cu": int(waves_per_cu), "by_lanes": groups,
            "last_island": {"contacts": int(rows[c, 2]), "bodies": int(rows[c, 3]), "lanes": int(rows[c, 4]),
                            "sweep_period": int(rows[c, 5]),
                            "start_us": float((start[c] - start.min()) / 1e3), "latency_us": float(lat[c])}}


G_DT = 10.0 / 60


def one_step_error(state: L.BatchArrays, precision: str, worlds):
    """Relative one-step error of a GPU pipeline vs the FP64 oracle from an identical state.

    Per body and quantity |gpu - ref| / max(|ref|, floor) with floors 1 m, 1 rad,
    |g| dt and |g| dt / 1 m; accumulated impulses of contacts present in both
    with floor m_min |g| dt.  Reports max, 99.9th and 99th percentiles and the
    fraction of bodies above 1e-5.
    """
    from oracle import oracle as O
    from pypybox2d_amd._native import Engine
    sample = L.concat([state.world(w) for w in worlds])
    sample.world_i[:, L.WI_FLAGS] &= ~L.WF_NEW_FIXTURE
    o = O.Oracle(sample)
    o.step(1.0 / 60, 8, 3)
    ref = o.state()
    e = Engine(precision)
    e.upload(sample)
    e.step(1.0 / 60, 8, 3)
    got = e.download()
    e.close()
    out = {}
    g, r = got.body_d, ref.body_d
    dyn = sample.body_i[:, L.BI_TYPE] == L.DYNAMIC
    for name, cols, floor in (("position", [L.BD_C_X, L.BD_C_Y], 1.0), ("angle", [L.BD_A], 1.0),
                              ("velocity", [L.BD_V_X, L.BD_V_Y], G_DT),
                              ("angular_velocity", [L.BD_W], G_DT)):
        e_ = np.max(np.abs(g[:, cols] - r[:, cols]) / np.maximum(np.abs(r[:, cols]), floor), axis=1)[dyn]
        out[name] = {"max": float(e_.max()), "p99.9": float(np.percentile(e_, 99.9)),
                     "p99": float(np.percentile(e_, 99)), "frac_over_1e-5": float(np.mean(e_ > 1e-5))}
    # accumulated impulses of contacts present in both with the same discrete state
    imp = []
    flips = 0
    bm = ref.body_d[:, L.BD_MASS]
    for w in range(sample.n_worlds):
        rc = {tuple(ref.con_i[k, :2]): k for k in range(ref.con_off[w], ref.con_off[w + 1])}
        b0, f0 = ref.body_off[w], ref.fix_off[w]
        for k in range(got.con_off[w], got.con_off[w + 1]):
            j = rc.get(tuple(got.con_i[k, :2]))
            if j is None or got.con_i[k, L.CI_PC] != ref.con_i[j, L.CI_PC] or \
                    (got.con_i[k, L.CI_FLAGS] & 3) != (ref.con_i[j, L.CI_FLAGS] & 3):
                flips += 1
                continue
            ms = [m for m in (bm[b0 + ref.fix_i[f0 + ref.con_i[j, L.CI_FA], L.FI_BODY]],
                              bm[b0 + ref.fix_i[f0 + ref.con_i[j, L.CI_FB], L.FI_BODY]]) if m > 0]
            floor = (min(ms) if ms else 1.0) * G_DT
            for col in ((L.CD_P0_NI, L.CD_P0_TI, L.CD_P1_NI, L.CD_P1_TI)[:2 * int(ref.con_i[j, L.CI_PC])]):
                imp.append(abs(got.con_d[k, col] - ref.con_d[j, col]) / max(abs(ref.con_d[j, col]), floor))
    imp = np.array(imp) if imp else np.zeros(1)
    out["impulse"] = {"max": float(imp.max()), "p99.9": float(np.percentile(imp, 99.9)),
                      "p99": float(np.percentile(imp, 99)), "frac_over_1e-5": float(np.mean(imp > 1e-5))}
    out["contact_discrete_flips"] = flips
    same = all(np.array_equal(got.con_i[got.con_off[w]:got.con_off[w + 1], :2],
                              ref.con_i[ref.con_off[w]:ref.con_off[w + 1], :2]) for w in range(sample.n_worlds))
    out["pair_lists_equal"] = bool(same)
    out["worlds"] = len(worlds)
    out["dynamic_bodies"] = int(dyn.sum())
    return out


def sampled_errors(batch: L.BatchArrays, steps: int, every: int) -> dict:
    """One-step errors of the FP32 pipeline against the FP64 oracle from the FP64 pipeline's state
    every `every` steps (a separate, untimed run)."""
    from pypybox2d_amd._native import Engine
    e = Engine("fp64")
    e.upload(batch)
    samples = []
    for s in range(every, steps + 1, every):
        e.step_n(every, graph=True)
        st = e.download()
        err = one_step_error(st, "fp32", list(range(st.n_worlds)))
        samples.append({"step": s, **{k: (v["max"] if isinstance(v, dict) else v) for k, v in err.items()}})
    e.close()
    worst = {}
    for q in ("position", "angle", "velocity", "angular_velocity", "impulse"):
        worst[q] = max(smp[q] for smp in samples)
    return {"samples": samples, "max": worst}


def long_horizon(state: L.BatchArrays) -> dict:
    """Kinetic energy and mean height of the dynamic bodies (stacking is chaotic, so trajectories
    of different precisions are compared through such statistics)."""
    bd = state.body_d
    dyn = state.body_i[:, L.BI_TYPE] == L.DYNAMIC
    m, I = bd[dyn, L.BD_MASS], bd[dyn, L.BD_I]
    ke = 0.5 * (m * (bd[dyn, L.BD_V_X] ** 2 + bd[dyn, L.BD_V_Y] ** 2) + I * bd[dyn, L.BD_W] ** 2)
    return {"kinetic_energy_per_body": float(ke.sum() / max(1, dyn.sum())),
            "mean_height": float(bd[dyn, L.BD_C_Y].mean()),
            "awake_fraction": float(np.mean((state.body_i[dyn, L.BI_FLAGS] & L.BF_AWAKE) != 0))}


_CPU_BATCH = None


def _cpu_worker(args):
    """Whole trajectories of worlds p, p + P, ... (the GPU's step count each) until `seconds` pass."""
    p, nproc, steps, seconds = args
    from oracle import oracle as O
    t0 = time.perf_counter()
    body_steps, worlds = 0, 0
    for w in range(p, _CPU_BATCH.n_worlds, nproc):
        sample = _CPU_BATCH.world(w)
        o = O.Oracle(sample)
        for _ in range(steps):
            o.step(1.0 / 60, 8, 3)
        body_steps += int(np.count_nonzero(sample.body_i[:, L.BI_TYPE] == L.DYNAMIC)) * steps
        worlds += 1
        if time.perf_counter() - t0 >= seconds:
            break
    return body_steps, time.perf_counter() - t0, worlds


def cpu_baseline(batch: L.BatchArrays, steps: int, seconds: float, procs: int):
    """The FP64 oracle (oracle/b2oracle.c) on the host cores: one world per process at a time, each
    world stepped from the same initial state for the same number of steps as the GPU run."""
    import multiprocessing as mp
    global _CPU_BATCH
    _CPU_BATCH = batch
    procs = max(1, min(procs, batch.n_worlds))  # one world per process
    jobs = [(p, procs, steps, seconds) for p in range(procs)]
    if procs > 1:
        with mp.get_context("fork").Pool(procs) as pool:
            res = pool.map(_cpu_worker, jobs)
    else:
        res = [_cpu_worker(j) for j in jobs]
    rate = sum(bs / el for bs, el, _ in res)
    nw = sum(r[2] for r in res)
    return {"value": rate, "unit": "body-steps/s", "cores": procs, "kind": "port",
            "sample": "FP64 oracle (oracle/b2oracle.c, a C restatement of the reference's step) on %d host cores, "
                      "one process per core, each stepping whole worlds of the same batch (worlds p, p+%d, ...) "
                      "for %d steps from the initial state until %.0f s had passed: %d worlds in total"
                      % (procs, procs, steps, seconds, nw)}


# the kernels of each timed stage (the stage names are those of the stage's main kernel)
STAGE_KERNELS = {"k_collide": ("k_reset_queue", "k_collide"), "k_islands": ("k_islands", "k_order_islands"),
                 "k_solve": ("k_prep_bodies", "k_prep_contacts", "k_solve"), "k_sync": ("k_sync",),
                 "k_pairs": ("k_pairs",)}


def measured_traffic(kernel: str, workload: str):
    """HBM bytes per step of a stage (the sum over its kernels and their dispatches) from the
    committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary of the default workload (FP32 pipeline),
    or None."""
    path = os.path.join(HERE, "profiles", "traffic_mixed4096_fp32.md")
    if workload != "mixed" or not os.path.exists(path):
        return None
    per = {}
    for line in open(path):
        if line.startswith("traffic_bytes_per_step "):
            name, val = line[len("traffic_bytes_per_step "):].split(":", 1)
            per[name.strip()] = float(val)
    parts = STAGE_KERNELS.get(kernel, (kernel,))
    if kernel not in per:
        return None
    return sum(per.get(k, 0.0) for k in parts)


def roofline(kernel: str, ms: float, counts: dict, fp64: bool, workload: str = "mixed") -> dict:
    ab = algorithmic_bytes(kernel, counts, fp64)
    achieved = ab / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    out = {"bound": "hbm", "kernel": kernel, "kernels": list(STAGE_KERNELS.get(kernel, (kernel,))),
           "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(kernel, workload), "algorithmic_bytes": ab,
           "avg_ms": ms}
    if kernel in LATENCY_LIMITS:
        out["limit"] = LATENCY_LIMITS[kernel]
    return out


# Kernels whose time is set by serial dependency chains rather than by bytes or flops; their HBM
# fraction is reported for completeness and the "solver" report gives the solve's own measure.
LATENCY_LIMITS = {
    "k_solve": "dependency chains inside the largest islands: passes (about the Gauss-Seidel dependency depth "
               "of all sweeps) x pass latency; see the solver report (occupancy and per-island latency)",
    "k_islands": "one lane walks each world's (or island's) depth-first search",
}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="default 1000 (600 for pyramids)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", choices=("mixed", "lognormal", "pyramids"), default="mixed",
                    help="mixed: 4096 worlds, body counts U[32,512]; lognormal: 16384 worlds, median 64, sigma 0.75, "
                         "clipped to [16,1024]; pyramids: one world of 200 ten-row box pyramids")
    ap.add_argument("--worlds", type=int, default=None, help="default 4096 (mixed) or 16384 (lognormal)")
    ap.add_argument("--seed", type=int, default=None, help="default 1 (mixed) or 2 (lognormal)")
    ap.add_argument("--fp64-steps", type=int, default=None, help="timed steps of the FP64 run (default: --steps)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default solver launch shape")
    ap.add_argument("--error-worlds", type=int, default=64)
    ap.add_argument("--no-solver-report", action="store_true",
                    help="skip the extra step with per-island solver timing (used for profiler runs)")
    ap.add_argument("--no-graph-check", action="store_true",
                    help="skip the graph-replayed steps after the timed region (used for profiler runs)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-procs", type=int, default=min(16, os.cpu_count() or 1))
    args = ap.parse_args()
    world_size = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world_size != 1:
        print("bench.py runs on one GPU only (the batch is one GPU's share; worlds would shard across GPUs "
              "with no communication); got --gpus %d, WORLD_SIZE=%d" % (args.gpus, world_size))
        sys.exit(2)

    if args.worlds is None:
        args.worlds = {"mixed": 4096, "lognormal": 16384, "pyramids": 1}[args.workload]
    if args.seed is None:
        args.seed = 1 if args.workload == "mixed" else 2
    if args.steps is None:
        args.steps = 600 if args.workload == "pyramids" else 1000
    t0 = time.perf_counter()
    if args.workload == "pyramids":
        args.worlds = 1
        batch = builder.scene_batch(scenes.pyramids(200))
        desc = ("pyramids: one world of 200 ten-row box pyramids (11 000 boxes, half-extent 0.5) on 200 static "
                "ground edges 30 m apart, dt 1/60, 8/3 iterations")
    else:
        if args.workload == "mixed":
            counts = scenes.mixed_batch_counts(args.worlds, args.seed)
            desc = ("mixed_batch: %d worlds, body counts U[32,512] (seed %d), boxes/circles/polygons 40/30/30, "
                    "dt 1/60, 8/3 iterations" % (args.worlds, args.seed))
        else:
            counts = scenes.lognormal_batch_counts(args.worlds, args.seed)
            desc = ("lognormal_batch: %d worlds, body counts clip(round(64 exp(0.75 N(0,1))), 16, 1024) (seed %d), "
                    "boxes/circles/polygons 40/30/30, dt 1/60, 8/3 iterations" % (args.worlds, args.seed))
        log("building %d worlds (%d dynamic bodies) on the host" % (args.worlds, int(counts.sum())))
        batch = builder.batch_from_counts(counts, args.seed)
    log("built in %.1f s" % (time.perf_counter() - t0))
    dyn = int(np.count_nonzero(batch.body_i[:, L.BI_TYPE] == L.DYNAMIC))

    el32, ms32, st32, tuned32, gms32, solver32 = run_gpu(batch, "fp32", args.steps, args.warmup, not args.no_tune,
                                                         not args.no_solver_report, not args.no_graph_check)
    value = dyn * args.steps / el32
    log("fp32: %.3e body-steps/s (%.3f ms/step)" % (value, 1e3 * el32 / args.steps))
    c = workload_counts(st32)
    k = int(np.argmax(ms32))
    line = {
        "metric": "body-steps/sec, 4096 mixed worlds, FP32 vs FP64 on GPU; one-step err vs FP64",
        "value": value, "unit": "body-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * el32 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 bulk (f64 positions/angles/velocities/impulses/AABBs)",
        "data": "synthetic",
        "config": {"workload": desc,
                   "worlds": args.worlds, "dynamic_bodies": dyn, "contacts_at_end": c["contacts"],
                   "touching_at_end": c["touching"]},
        "kernel_ms": {n: float(v) for n, v in zip(KERNELS, ms32)},
        "roofline": roofline(KERNELS[k], float(ms32[k]), c, False, args.workload),
        "roofline_other": {n: roofline(n, float(ms32[i]), c, False, args.workload) for i, n in enumerate(KERNELS)
                           if n in ("k_collide", "k_pairs", "k_sync")},
        "solver": solver32,
        "launch_shape": tuned32,
        "graph_ms_per_step_next_%d" % GRAPH_CHECK_STEPS: gms32,
        "timing": "wall clock of the timed steps; per-kernel averages from HIP events around every timed launch",
    }
    if not args.no_fp64:
        n64 = args.fp64_steps or args.steps
        el64, ms64, st64, tuned64, gms64, solver64 = run_gpu(batch, "fp64", n64, args.warmup, not args.no_tune,
                                                             not args.no_solver_report, not args.no_graph_check)
        v64 = dyn * n64 / el64
        line["fp64"] = {"value": v64, "ms_per_step": 1e3 * el64 / n64, "steps": n64,
                        "kernel_ms": {n: float(v) for n, v in zip(KERNELS, ms64)},
                        "launch_shape": tuned64, "graph_ms_per_step_next_%d" % GRAPH_CHECK_STEPS: gms64,
                        "solver": solver64}
        line["fp32_over_fp64"] = value / v64
        log("fp64: %.3e body-steps/s; fp32/fp64 = %.2f" % (v64, value / v64))
        if n64 == args.steps:
            line["long_horizon"] = {"fp32": long_horizon(st32), "fp64": long_horizon(st64),
                                    "steps": args.warmup + args.steps}
    if args.workload == "pyramids" and args.error_worlds > 0:
        line["one_step_error_fp32_vs_fp64_oracle_every_50"] = sampled_errors(batch, args.warmup + args.steps, 50)
        log("sampled one-step errors:", line["one_step_error_fp32_vs_fp64_oracle_every_50"]["max"])
    rng = np.random.default_rng(0)
    worlds = sorted(rng.choice(args.worlds, size=min(args.error_worlds, args.worlds), replace=False).tolist())
    for prec, st in ((("fp32", st32),) + ((("fp64", st64),) if not args.no_fp64 else ())) if worlds else ():
        key = "one_step_error_%s_vs_fp64_oracle" % prec
        try:
            line[key] = one_step_error(st, prec, worlds)
            line[key]["from_step"] = args.warmup + ((args.fp64_steps or args.steps) if prec == "fp64" else args.steps)
            log("%s one-step error:" % prec, line[key])
        except Exception as exc:  # the error check never hides the measurement
            line[key] = {"error": str(exc)}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(batch, args.warmup + args.steps, args.cpu_seconds, args.cpu_procs)
        log("cpu baseline: %.3e body-steps/s" % line["cpu_baseline"]["value"])
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
