#!/usr/bin/env python
"""Benchmark of the per-step contact pipeline on one MI355X.

Workload (default): 4096 independent mixed worlds (body counts uniform in
[32, 512], seed 1; boxes / circles / convex polygons 40/30/30 above a static
edge), dt = 1/60, 8 velocity and 3 position iterations, continuous physics
off.  A step is one World.step of every world.

    python bench.py [--gpus 1] [--steps 100] [--warmup 10] [--settle 1000]

The worlds are first stepped --settle times (untimed, graph-captured, FP32) so
that the timed window sees settled piles, not bodies in free fall.  From that
state each pipeline (FP32 bulk, then all-FP64) takes one launch-tuning pass
and --warmup graph steps, then --steps graph-captured steps are timed by the
wall clock.  value = dynamic bodies x timed steps / seconds (FP32 pipeline,
state resident in HBM).  The same window is then re-run eagerly from its saved
start with HIP events between the kernels for the per-stage times (the
pipeline is deterministic; the re-run's end state is checked to be identical).
The line also gives the dominant stage's roofline, the whole step's
algorithmic bytes from the device's work counters, one-step errors of both
pipelines against the FP64 oracle on 64 worlds of the settled state, and the
oracle's rate on the host cores over the same worlds and window.

Profiles: every step of the process runs k_reset_queue first, so a rocprofv3
trace of this command splits into steps at those dispatches; the line's
"profile_window" gives the step ordinals (0-based, counted over the whole
process) of each pipeline's timed window and eager re-run.
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


def per_step(stats: dict) -> dict:
    """Work counters of the timed window, per step."""
    n = max(1, stats["steps"])
    return {k: v / n for k, v in stats.items() if k != "steps"}


def step_bytes(c: dict, w: dict) -> float:
    """Algorithmic bytes of one whole step by the survey's definition: per dynamic body 148 + G
    (state read and write, sweep write, constants, geometry, fat AABB), per persistent contact
    128 (pair, manifold read and write), per new contact 8, per re-inserted proxy 32."""
    return (148 * c["dynamic_bodies"] + c["geometry_bytes"] + 128 * w["live_contacts"] + 8 * w["new_contacts"]
            + 32 * w["moved_proxies"])


def stage_bytes(stage: str, c: dict, w: dict, fp64: bool) -> float:
    """Minimum HBM bytes of one launch of a stage's kernels, from the per-step work counters `w`
    and the batch's counts `c`.  t = bytes of the bulk type."""
    t = 8 if fp64 else 4
    man = 8 * t + 32 + 8 + 4  # local normal/point and two points (T), four FP64 impulses, two ids, state
    geo_per_fix = c["geometry_bytes"] / max(1, c["fixtures"])
    if stage == "k_collide":
        # live contacts: pair + state; evaluated: manifold read and write, both fixtures' records
        # and bodies' transforms; geometry of the fixtures once
        return (w["live_contacts"] * (8 + 8) + w["evaluated_contacts"] * (2 * man + 2 * (16 + 32))
                + c["geometry_bytes"])
    if stage == "k_solve":
        # island body: c, a, v, w read and write, c0/a0 write, inverse mass/inertia and local centre;
        # island contact: fixture/body refs, manifold and material read, impulses read and written
        return (w["island_bodies"] * (48 + 48 + 24 + 2 * t + 16)
                + w["island_contacts"] * (8 + 4 + 8 * t + 2 * t + 32 + 32))
    if stage == "k_islands":
        return w["live_contacts"] * (8 + 4) + c["bodies"] * 8 + w["island_contacts"] * 8 + w["island_bodies"] * 8
    if stage == "k_sync":
        return (c["fixtures"] * (8 + 4 + 8) + w["island_bodies"] * (72 + 32 + geo_per_fix)
                + w["moved_proxies"] * 32)
    if stage == "k_pairs":
        return (w["live_contacts"] * 12 + c["fixtures"] * 4 + w["moved_proxies"] * 32
                + w["new_contacts"] * (8 + 4 + 2 * t + man))
    return 0.0


# steps launched so far in this process, over all engines (the step ordinals of a profile)
STEP_COUNT = [0]


def settle(batch: L.BatchArrays, steps: int) -> L.BatchArrays:
    """The untimed prefix: `steps` graph-captured steps of the FP32 pipeline from the initial state,
    so that the timed window measures piles that have formed and settled."""
    from pypybox2d_amd._native import Engine
    eng = Engine("fp32")
    eng.upload(batch)
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        n = min(200, steps - done)
        eng.step_n(n, graph=True)
        eng.sync()
        done += n
        log("[settle] %d/%d steps, %.1f s" % (done, steps, time.perf_counter() - t0))
    state = eng.download()
    eng.close()
    STEP_COUNT[0] += steps
    state.world_i[:, L.WI_FLAGS] &= ~L.WF_NEW_FIXTURE
    return state


def run_gpu(start: L.BatchArrays, precision: str, steps: int, warmup: int, tune: bool = True,
            solver_times: bool = True):
    """Times `steps` graph-captured steps from `start` (after one tuning pass and `warmup` graph steps).

    value: wall clock of the timed steps, with the state resident in HBM.  Per-stage averages: the
    same window stepped again eagerly with HIP events between the kernels (b2g_profile_step),
    from the state saved at the window's start; the pipeline is deterministic, so this re-run does
    exactly the work the timed steps did (checked: its end state equals theirs bit for bit).
    Work counters (b2g_stats) are summed over the timed steps.
    """
    from pypybox2d_amd._native import Engine
    from pypybox2d_amd.autotune import LaunchTuner
    eng = Engine(precision)
    eng.upload(start)
    tuned = None
    if tune:
        tuner = LaunchTuner(eng)
        tuned = dict(tuner.tune(1.0 / 60, 8, 3))
        STEP_COUNT[0] += tuner.steps_per_tune
        log("[%s] launch shape %s (solve ms per candidate %s)" % (precision, tuned,
                                                                  np.round(tuner.last_times, 2).tolist()))
    if warmup > 0:
        eng.step_n(warmup, graph=True)
    eng.sync()
    STEP_COUNT[0] += warmup
    first = STEP_COUNT[0]
    profile_window = {"graph": [first, first + steps - 1], "eager": [first + steps, first + 2 * steps - 1]}
    STEP_COUNT[0] += 2 * steps
    window = eng.download()
    eng.stats(reset=True)
    log("[%s] timing %d graph-captured steps" % (precision, steps))
    t0 = time.perf_counter()
    eng.step_n(steps, graph=True)
    eng.sync()
    elapsed = time.perf_counter() - t0
    stats = eng.stats(reset=True)
    end = eng.download()
    # the same window again, eager, with events between the kernels
    eng.upload(window)
    ms = eng.profile(steps)
    eng.sync()
    again = eng.download()
    identical = bool(np.array_equal(again.body_d, end.body_d) and np.array_equal(again.con_d, end.con_d)
                     and np.array_equal(again.con_i, end.con_i))
    log("[%s] %.3f ms/step; stages %s ms; eager re-run identical: %s" % (
        precision, 1e3 * elapsed / steps, np.round(ms, 3).tolist(), identical))
    # one more (untimed) step with per-island timing, for the solver's latency report
    solver = None
    if solver_times:
        eng.set_param("island_times", 1)
        eng.profile(1)
        STEP_COUNT[0] += 1
        solver = solver_report(eng.island_times(), eng.get_param("solve_waves_per_cu"))
    eng.close()
    return {"elapsed": elapsed, "stage_ms": ms, "state": end, "launch_shape": tuned, "solver": solver,
            "stats": stats, "eager_rerun_identical": identical, "profile_window": profile_window}


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
                                                  "max": float(lat[m].max())},
                                   # share of the latency spent in the velocity iterations (the rest:
                                   # impulse store, integration, position iterations, write-back)
                                   "velocity_share": float(((rows[m, 6] - start[m]) / np.maximum(end[m] - start[m], 1)).mean()),
                                   "position_iterations_mean": float(rows[m, 7].mean())}
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
    """Worlds p, p + P, ... of the settled state, each stepped `steps` times (the GPU's timed window),
    until `seconds` pass."""
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


def host_cores() -> dict:
    """The cores this process may use, and the host's count."""
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"usable": usable, "host_cpu_count": os.cpu_count()}


def cpu_baseline(state: L.BatchArrays, steps: int, seconds: float, procs: int):
    """The FP64 oracle (oracle/b2oracle.c) on the host cores: one world per process at a time, each
    world of the settled state stepped for the GPU's timed window."""
    import multiprocessing as mp
    global _CPU_BATCH
    _CPU_BATCH = state
    procs = max(1, min(procs, state.n_worlds))  # one world per process
    jobs = [(p, procs, steps, seconds) for p in range(procs)]
    if procs > 1:
        with mp.get_context("fork").Pool(procs) as pool:
            res = pool.map(_cpu_worker, jobs)
    else:
        res = [_cpu_worker(j) for j in jobs]
    rate = sum(bs / el for bs, el, _ in res)
    nw = sum(r[2] for r in res)
    hc = host_cores()
    return {"value": rate, "unit": "body-steps/s", "cores": procs, "kind": "port",
            "host_cpu_count": hc["host_cpu_count"], "usable_cores": hc["usable"],
            "sample": "FP64 oracle (oracle/b2oracle.c, a C restatement of the reference's step) in %d processes, "
                      "one world per process at a time: worlds p, p+%d, ... of the same settled state, each "
                      "stepped for the GPU's %d timed steps, until %.0f s had passed (%d worlds). The host has %s "
                      "cores, %d usable by this process; the pool is held to 16, this GPU's share of the host"
                      % (procs, procs, steps, seconds, nw, hc["host_cpu_count"], hc["usable"])}


# the kernels of each timed stage (the stage names are those of the stage's main kernel)
STAGE_KERNELS = {"k_collide": ("k_reset_queue", "k_collide"), "k_islands": ("k_islands", "k_order_islands", "k_order_team_islands"),
                 "k_solve": ("k_island_prefix", "k_prep_bodies", "k_prep_contacts", "k_solve"), "k_sync": ("k_sync",),
                 "k_pairs": ("k_pairs",)}


def measured_traffic(stage: str, args) -> dict:
    """HBM bytes per step of a stage (FETCH_SIZE + WRITE_SIZE over its kernels' dispatches, scaled
    by the copy-kernel calibration) from the committed rocprofv3 --pmc summary of this same command
    (same workload, settle prefix, warmup and timed steps), or None when there is none."""
    path = os.path.join(HERE, "profiles", "traffic_%s.md" % workload_tag(args))
    if not os.path.exists(path):
        return None
    per, cmd_ok = {}, False
    for line in open(path):
        if line.startswith("window: ") and line.strip() == "window: " + window_tag(args):
            cmd_ok = True
        if line.startswith("traffic_bytes_per_step "):
            name, val = line[len("traffic_bytes_per_step "):].split(":", 1)
            per[name.strip()] = float(val.split()[0])
    parts = STAGE_KERNELS.get(stage, (stage,))
    if not cmd_ok or not any(k in per for k in parts):
        return None
    return sum(per.get(k, 0.0) for k in parts)


def workload_tag(args) -> str:
    return {"mixed": "mixed%d" % args.worlds, "lognormal": "lognormal%d" % args.worlds,
            "pyramids": "pyramids200"}[args.workload] + "_fp32"


def window_tag(args) -> str:
    return "workload=%s seed=%d settle=%d warmup=%d steps=%d" % (args.workload, args.seed, args.settle, args.warmup,
                                                                 args.steps)


def roofline(stage: str, ms: float, c: dict, w: dict, fp64: bool, args) -> dict:
    ab = stage_bytes(stage, c, w, fp64)
    achieved = ab / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    traffic = measured_traffic(stage, args) if not fp64 else None
    out = {"bound": "hbm", "kernel": stage, "kernels": list(STAGE_KERNELS.get(stage, (stage,))),
           "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
           "traffic": traffic, "algorithmic_bytes": ab, "avg_ms": ms,
           "traffic_over_algorithmic": (traffic / ab) if traffic and ab else None}
    if stage in LATENCY_LIMITS:
        out["limit"] = LATENCY_LIMITS[stage]
    return out


# Stages whose time is set by serial dependency chains rather than by bytes or flops; their HBM
# fraction is reported for completeness and the "solver" report gives the solve's own measure.
LATENCY_LIMITS = {
    "k_solve": "dependency chains inside the largest islands: passes (about the Gauss-Seidel dependency depth "
               "of all sweeps) x pass latency; see the solver report (occupancy and per-island latency)",
    "k_islands": "one lane walks each world's (or island's) depth-first search",
}


def copy_calibration(mib: int = 256, reps: int = 5) -> None:
    """A plain device-to-device copy kernel (torch's) of known size inside a selected region: what
    rocprofv3's FETCH_SIZE and WRITE_SIZE read for it against the bytes it must move gives the
    scale of those counters."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read_per_dispatch": mib << 20, "bytes_written_per_dispatch":
                                           mib << 20, "dispatches": reps}}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed steps (graph-captured)")
    ap.add_argument("--warmup", type=int, default=10, help="untimed graph steps before the timed window")
    ap.add_argument("--settle", type=int, default=None,
                    help="untimed graph steps of the FP32 pipeline from the initial state before the window "
                         "(default 1000; 300 for pyramids)")
    ap.add_argument("--workload", choices=("mixed", "lognormal", "pyramids"), default="mixed",
                    help="mixed: 4096 worlds, body counts U[32,512]; lognormal: 16384 worlds, median 64, sigma 0.75, "
                         "clipped to [16,1024]; pyramids: one world of 200 ten-row box pyramids")
    ap.add_argument("--worlds", type=int, default=None, help="default 4096 (mixed) or 16384 (lognormal)")
    ap.add_argument("--seed", type=int, default=None, help="default 1 (mixed) or 2 (lognormal)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--error-worlds", type=int, default=64)
    ap.add_argument("--no-solver-report", action="store_true",
                    help="skip the extra step with per-island solver timing (used for profiler runs)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-procs", type=int, default=None, help="default: min(16, usable cores)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB five times with a plain copy kernel (calibrates rocprofv3 "
                         "FETCH_SIZE/WRITE_SIZE against known bytes)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    world_size = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world_size != 1:
        print("bench.py runs on one GPU only: every GPU box here has one GPU, and the batch is one GPU's share "
              "(worlds would shard across GPUs with no communication); got --gpus %d, WORLD_SIZE=%d"
              % (args.gpus, world_size))
        sys.exit(2)

    if args.worlds is None:
        args.worlds = {"mixed": 4096, "lognormal": 16384, "pyramids": 1}[args.workload]
    if args.seed is None:
        args.seed = 1 if args.workload == "mixed" else 2
    if args.settle is None:
        args.settle = 300 if args.workload == "pyramids" else 1000
    if args.cpu_procs is None:
        args.cpu_procs = min(16, host_cores()["usable"])
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
    desc += ("; timed window: steps %d..%d after an untimed prefix of %d graph-captured FP32 steps from the "
             "initial state (settled piles), then one launch-tuning pass (7 steps) and %d warm-up steps"
             % (args.settle + 7 + args.warmup + 1, args.settle + 7 + args.warmup + args.steps, args.settle,
                args.warmup))

    settled = settle(batch, args.settle) if args.settle > 0 else batch
    c = workload_counts(settled)
    log("settled state: %d contacts, %d touching" % (c["contacts"], c["touching"]))
    r32 = run_gpu(settled, "fp32", args.steps, args.warmup, not args.no_tune, not args.no_solver_report)
    value = dyn * args.steps / r32["elapsed"]
    w32 = per_step(r32["stats"])
    ms32 = r32["stage_ms"]
    k = int(np.argmax(ms32))
    sb = step_bytes(c, w32)
    line = {
        "metric": "body-steps/sec, 4096 mixed worlds, FP32 vs FP64 on GPU; one-step err vs FP64",
        "value": value, "unit": "body-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * r32["elapsed"] / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 bulk (f64 positions/angles/velocities/impulses/lever arms/AABBs)",
        "data": "synthetic",
        "config": {"workload": desc, "worlds": args.worlds, "dynamic_bodies": dyn, "settle_steps": args.settle,
                   "contacts_at_window_start": c["contacts"], "touching_at_window_start": c["touching"]},
        "kernel_ms": {n: float(v) for n, v in zip(KERNELS, ms32)},
        "roofline": roofline(KERNELS[k], float(ms32[k]), c, w32, False, args),
        "roofline_other": {n: roofline(n, float(ms32[i]), c, w32, False, args) for i, n in enumerate(KERNELS)
                           if i != k},
        "step_bytes": {"algorithmic": sb, "achieved_GBps": sb / (r32["elapsed"] / args.steps) / 1e9,
                       "frac": sb / (r32["elapsed"] / args.steps) / 1e9 / HBM_PEAK_GBS,
                       "per_step_counters": w32,
                       "definition": "sum over dynamic bodies of (148 + G) + 128 C_persist + 8 P_new + 32 R_reinsert"},
        "solver": r32["solver"],
        "launch_shape": r32["launch_shape"],
        "eager_rerun_identical": r32["eager_rerun_identical"],
        "profile_window": {"fp32": r32["profile_window"]},
        "timing": "value: wall clock of the timed graph-captured steps (barrier = stream sync on both sides); "
                  "kernel_ms: HIP events around every kernel of the same window re-run eagerly from its saved "
                  "start state",
    }
    if not args.no_fp64:
        r64 = run_gpu(settled, "fp64", args.steps, args.warmup, not args.no_tune, not args.no_solver_report)
        v64 = dyn * args.steps / r64["elapsed"]
        w64 = per_step(r64["stats"])
        line["fp64"] = {"value": v64, "ms_per_step": 1e3 * r64["elapsed"] / args.steps,
                        "kernel_ms": {n: float(v) for n, v in zip(KERNELS, r64["stage_ms"])},
                        "launch_shape": r64["launch_shape"], "solver": r64["solver"],
                        "roofline": roofline(KERNELS[k], float(r64["stage_ms"][k]), c, w64, True, args),
                        "eager_rerun_identical": r64["eager_rerun_identical"]}
        line["profile_window"]["fp64"] = r64["profile_window"]
        line["fp32_over_fp64"] = value / v64
        log("fp64: %.3e body-steps/s; fp32/fp64 = %.2f" % (v64, value / v64))
        line["long_horizon"] = {"fp32": long_horizon(r32["state"]), "fp64": long_horizon(r64["state"]),
                                "steps": args.settle + 7 + args.warmup + args.steps}
    if args.workload == "pyramids" and args.error_worlds > 0:
        line["one_step_error_fp32_vs_fp64_oracle_every_50"] = sampled_errors(batch, 600, 50)
        log("sampled one-step errors:", line["one_step_error_fp32_vs_fp64_oracle_every_50"]["max"])
    rng = np.random.default_rng(0)
    worlds = sorted(rng.choice(args.worlds, size=min(args.error_worlds, args.worlds), replace=False).tolist())
    for prec in (("fp32",) + (() if args.no_fp64 else ("fp64",))) if worlds else ():
        key = "one_step_error_%s_vs_fp64_oracle" % prec
        try:
            line[key] = one_step_error(settled, prec, worlds)
            line[key]["from_step"] = args.settle
            log("%s one-step error:" % prec, line[key])
        except Exception as exc:  # the error check never hides the measurement
            line[key] = {"error": str(exc)}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(settled, args.steps, args.cpu_seconds, args.cpu_procs)
        log("cpu baseline: %.3e body-steps/s" % line["cpu_baseline"]["value"])
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
