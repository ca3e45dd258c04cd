#!/usr/bin/env python3
"""Contact-path benchmarks on one MI355X.

Default workload (the headline metric): contacts solved per second on the
random convex pile of 32,768 bodies (25 % spheres, boxes, capsules, convex
hulls of 16-32 points; dropped from y in [2, 60] into a 64 x 64 m walled bin;
density 1000; FP32; 8 velocity + 2 position iterations; dt = 1/60; sleeping
off).  A fixed settle phase of S untimed StepFixed calls (default 300; the
pile comes to rest around step 250) runs whatever --warmup says, then W
warm-up calls, then K timed calls.  A step is one StepFixed: host broadphase, then narrowphase,
constraint build, solve and integration on the GPU.  value = sum(contact
points) / sum(narrowphase kernels + host build compute + integrate
velocities + solve), with inputs in HBM; the rate including the PCIe
transfers at the boundary is reported beside it.  The learned warm start is
trained during the last settle / warm-up steps and compared with the cache
and cold warm starts on the final state (PGS residual after k sweeps).

Other workloads:
  --workload batched      4,096 independent worlds of 16-512 bodies (box
                          stacks and convex piles, each on its own ground
                          plane), one captured HIP graph per step, launch
                          configuration chosen by the autotuner
  --workload narrowphase  1,048,576 convex pairs (hull-hull 8-64 vertices,
                          box-hull, capsule-box), penetration U[0, 0.05] m,
                          10 % separated by U[0, 0.02] m

Every workload reports the reference CPU path (the reference engine built
from its own sources, oracle/_ref) on a bounded sample of the same workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--settle S] [--bodies N]
                    [--workload pile|batched|narrowphase|copy]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from kizuriphysics3d_amd import scenes as S  # noqa: E402

HBM_PEAK_GBS = 8000.0


def pmc_traffic(summary: str):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary in
    profiles/: FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each scaled by
    the copy-kernel calibration recorded in the same file.  None if absent."""
    path = os.path.join(ROOT, "profiles", summary)
    try:
        kv = {}
        for line in open(path):
            if "=" in line and line.split("=")[0].strip() in ("fetch_kib_per_dispatch", "write_kib_per_dispatch",
                                                              "fetch_scale", "write_scale"):
                k, v = line.split("=")
                kv[k.strip()] = float(v)
        return 1024.0 * (kv["fetch_scale"] * kv["fetch_kib_per_dispatch"] +
                         kv["write_scale"] * kv["write_kib_per_dispatch"])
    except (OSError, KeyError, ValueError):
        return None
# Algorithmic HBM bytes per contact point for one persistent solve launch:
# warm start 168 B + velocity iteration 312 B + bias iteration 260 B per point.
WARM_BYTES, VEL_BYTES, BIAS_BYTES = 168, 312, 260


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


# A one-GPU box grants the process a share of 16 host cores, whatever the
# machine has: the reference's JobSystem gets that many threads at most.
HOST_CORE_SHARE = 16


def host_threads():
    """Threads for the reference CPU path and what the host offers:
    (threads, {cpu_count, affinity, share, model})."""
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    threads = max(1, min(allowed, HOST_CORE_SHARE))
    return threads, {"cpu_count": os.cpu_count(), "affinity": allowed, "core_share": HOST_CORE_SHARE,
                     "model": cpu_model()}


def cpu_baseline(scene, states, budget_s: float, first_step: int):
    """The reference engine on the GPU run's pile: restart it from the body
    states at the start of the timed window and time its narrowphase + build
    + solve phases."""
    from oracle import refharness as R
    if not R.available():
        return None
    threads, host = host_threads()
    workers = R.init_jobs(max(threads - 1, 1))
    rw = R.RefWorld(scene)
    rw.set_states(states)
    t_begin = time.time()
    pts = 0
    ms = 0.0
    n = 0
    first = True
    while time.time() - t_begin < budget_s or n == 0:
        ph = rw.step_phases()
        hdr, _ = rw.constraints()
        if first:  # the restart step runs with an empty warm-start cache: not timed
            first = False
            if time.time() - t_begin > budget_s:
                break
            continue
        pts += int(hdr[:, 2].sum())
        ms += float(ph[1] + ph[2] + ph[3] + ph[4])
        n += 1
    rw.close()
    if ms <= 0:
        return None
    return {"value": pts / (ms / 1e3), "unit": "contacts/s", "cores": workers + 1, "kind": "reference",
            "host": host, "points_per_step": pts / n,
            "sample": f"{n} StepFixed calls of the reference (steps {first_step + 1}-{first_step + n}) on the same "
                      f"{len(states)}-body pile, restarted from the GPU run's body state after step {first_step - 1} "
                      f"(the restart step {first_step}, with an empty warm-start cache, is not timed); phases "
                      f"Collide + BuildContactConstraints + IntegrateVelocities + SolveContacts; JobSystem with "
                      f"{workers} workers + caller"}


def learned_report(res, train, state: str) -> dict:
    """The warm-start comparison as reported in the bench line: per mode the
    sweep-increment residual curve and the natural-map residual curve (k =
    1..8), penetration, and how many sweeps each takes to reach the cache
    warm start's 8-sweep value."""
    from kizuriphysics3d_amd import learned as Lw
    cache8 = float(res["cache"]["residual"][-1])
    nat8 = float(res["cache"]["natural"][-1]["total"]) if "natural" in res["cache"] else None
    out = {"definition": "increment: RMS over contact points of the impulse increments of sweep k; natural: RMS of "
                         "the natural-map residual of the iterate after k sweeps (normal complementarity "
                         "min(lambda_n, m_n (vRel.n - bias)), projected friction and cone terms); N s",
           "state": state, "points": res["cache"]["points"], "k": list(range(1, len(res["cache"]["residual"]) + 1)),
           "train": train}
    for name, r in res.items():
        out[name] = [float(x) for x in r["residual"]]
        out[name + "_penetration"] = r["penetration"]
        out[name + "_sweeps_to_cache8"] = Lw.iterations_to_reach(r["residual"], cache8)
        line = f"residual {name:13s} increment k=1..8 {np.array2string(r['residual'], precision=3)}"
        if "natural" in r:
            curve = np.array([x["total"] for x in r["natural"]])
            out[name + "_natural"] = curve.tolist()
            out[name + "_natural_terms"] = r["natural"][-1]
            out[name + "_natural_sweeps_to_cache8"] = Lw.iterations_to_reach(curve, nat8)
            line += f"; natural {np.array2string(curve, precision=3)}"
        log(line)
    return out


def bench_learned(args):
    """The learned warm start at its own workload: the random convex pile at
    65,536 bodies, settled, trained on the rows of its last steps, and the
    residual curves of every start on one state (learned.study)."""
    from kizuriphysics3d_amd import learned as Lw
    bodies = args.bodies if args.bodies != 32768 else 65536
    clips = tuple(float(x) for x in args.learn_clips.split(","))
    t0 = time.time()
    res = Lw.study(bodies=bodies, settle=args.settle, collect=args.learn_collect, train_steps=args.learn_steps,
                   batch=args.learn_batch, clips=clips, converged_iterations=64, targets=args.learn_targets, log=log)
    rows = Lw.study_summary(res)
    for name, r in rows.items():
        log(f"{name:22s} natural {np.array2string(np.array(r['natural']), precision=2)} "
            f"-> cache@8 in {r['sweeps_to_cache8_natural']} (increment: {r['sweeps_to_cache8_increment']})")
    best = min((r["sweeps_to_cache8_natural"] or 9, name) for name, r in rows.items() if name.startswith("learned_"))
    return {
        "metric": "sweeps for the learned warm start to reach the cache warm start's 8-sweep natural residual "
                  f"({bodies}-body random convex pile)",
        "value": best[0], "unit": "sweeps", "n_gpus": 1, "steps": 1, "warmup": args.settle,
        "ms_per_step": None, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 solve, bf16 MLP weights", "data": "synthetic",
        "config": {"workload": f"learned warm start, random convex pile {bodies} bodies, settled {args.settle} "
                               f"steps, rows from steps {res['collected_over_steps']} ({args.learn_targets} targets)",
                   "best": best[1],
                   "seconds": time.time() - t0},
        "targets": res["targets"], "variants": {k: {"loss_first": v["loss_first"], "loss_last": v["loss_last"]}
                                                for k, v in res["variants"].items()},
        "curves": rows,
    }


def bench_batched(args):
    """4,096 independent worlds, one captured graph per step."""
    import torch
    from kizuriphysics3d_amd.batch import BatchedWorlds
    t0 = time.time()
    scenes = S.batched_worlds(args.worlds)
    bw = BatchedWorlds(scenes)
    nb = sum(len(sc.bodies) for sc in scenes)
    log(f"batch: {len(scenes)} worlds, {nb} bodies, built in {time.time() - t0:.1f} s")
    tuned_steps = 0
    untimed = args.settle + args.warmup
    for i in range(untimed):
        bw.step()
        if (i + 1) % 50 == 0:
            st = bw.stats()
            log(f"settle/warmup {i + 1}/{untimed}: {st['num_points']} points, device {st['device_ms']:.2f} ms, "
                f"host broadphase {st['broadphase_ms']:.1f} ms, step {st['step_ms']:.1f} ms")
    cands = bw.candidates()
    best, times = bw.autotune(repeats=2)
    tuned_steps = 2 * len(cands)
    log(f"autotune over {len(cands)} configurations ({tuned_steps} steps): best {cands[best]}, "
        f"medians ms {np.round(times, 3).tolist()}")
    torch.cuda.synchronize()
    points = 0
    dev_ms = copies_ms = bp_ms = 0.0
    t_start = time.perf_counter()
    for _ in range(args.steps):
        bw.step()
        st = bw.stats()
        points += st["num_points"]
        dev_ms += st["device_ms"]
        copies_ms += st["device_with_copies_ms"]
        bp_ms += st["broadphase_ms"]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    st = bw.stats()
    log(f"timed {args.steps} steps: {points / args.steps:.0f} points/step, device {dev_ms / args.steps:.2f} ms "
        f"({copies_ms / args.steps:.2f} ms with the state copies), host broadphase {bp_ms / args.steps:.1f} ms, "
        f"levels vel/bias max {st['max_velocity_levels']}/{st['max_bias_levels']}")
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline_worlds(scenes, bw, args.cpu_budget)
        except Exception as e:
            log(f"cpu baseline unavailable: {e}")
    total_steps = untimed + tuned_steps + args.steps
    return {
        "metric": "contacts solved/s, 4096 batched worlds",
        "value": points / (dev_ms / 1e3),
        "unit": "contacts/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{len(scenes)} worlds of 16-512 bodies (log-uniform; box stacks and convex piles, "
                               f"each on its own ground plane), 8+2 iterations, dt 1/60, steps "
                               f"{total_steps - args.steps + 1}-{total_steps} after {args.settle} settle, "
                               f"{args.warmup} warm-up and {tuned_steps} autotuning steps",
                   "settle_steps": args.settle, "autotune_steps": tuned_steps,
                   "worlds": len(scenes), "bodies": nb, "contact_points_per_step": points / args.steps,
                   "device_ms_per_step": dev_ms / args.steps,
                   "device_ms_per_step_with_state_copies": copies_ms / args.steps,
                   "host_broadphase_ms_per_step": bp_ms / args.steps,
                   "world_steps_per_s_device": len(scenes) * args.steps / (dev_ms / 1e3),
                   "autotuned": cands[best]},
        "roofline": None,
        "cpu_baseline": cpu,
    }


def cpu_baseline_worlds(scenes, bw, budget_s: float):
    """The reference engine stepping worlds of the batch one after another,
    restarted from the GPU run's states (one untimed warm-start step each)."""
    from oracle import refharness as R
    if not R.available():
        return None
    threads, host = host_threads()
    workers = R.init_jobs(max(threads - 1, 1))
    t_begin = time.time()
    pts, ms, n, worlds = 0, 0.0, 0, 0
    for k, sc in enumerate(scenes):
        if time.time() - t_begin > budget_s:
            break
        rw = R.RefWorld(sc)
        rw.set_states(bw.states(k))
        rw.step_phases()
        for _ in range(3):
            ph = rw.step_phases()
            hdr, _ = rw.constraints()
            pts += int(hdr[:, 2].sum())
            ms += float(ph[1] + ph[2] + ph[3] + ph[4])
            n += 1
        rw.close()
        worlds += 1
    if ms <= 0:
        return None
    return {"value": pts / (ms / 1e3), "unit": "contacts/s", "cores": workers + 1, "kind": "reference",
            "host": host,
            "sample": f"{worlds} of the batch's worlds, each restarted from the GPU run's state and stepped 3 times "
                      f"after one untimed warm-start step ({n} StepFixed calls); phases Collide + "
                      f"BuildContactConstraints + IntegrateVelocities + SolveContacts; JobSystem with {workers} "
                      f"workers + caller"}


def bench_narrowphase(args):
    """1,048,576 convex pairs through the standalone narrowphase."""
    import torch
    from kizuriphysics3d_amd import device as D
    from kizuriphysics3d_amd.world import build_hulls
    t0 = time.time()
    pb = S.pair_batch(args.pairs)
    hulls = build_hulls(pb.shapes, pb.hull_points)
    ctx = D.DeviceContext(S.world_settings())
    ctx.upload_shapes(pb.shapes, hulls)
    log(f"narrowphase: {args.pairs} pairs over {len(pb.shapes)} shapes, prepared in {time.time() - t0:.1f} s")
    for _ in range(max(args.warmup, 1)):
        counts, pts, ms = ctx.collide_batch(pb.shape_a, pb.shape_b, pb.xf_a, pb.xf_b)
    torch.cuda.synchronize()
    total_ms = 0.0
    t_start = time.perf_counter()
    for _ in range(args.steps):
        counts, pts, ms = ctx.collide_batch(pb.shape_a, pb.shape_b, pb.xf_a, pb.xf_b)
        total_ms += ms
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    per = total_ms / args.steps
    # Algorithmic bytes per pair: 2 transforms (2 x 32 B) + shape records
    # (2 x 28 B) + hull vertices (16 B each) and, for box/hull pairs, their
    # triangles with face normals (20 B each), + count (4 B) + points (48 B each).
    nverts = np.zeros(len(pb.shapes), np.int64)
    ntris = np.zeros(len(pb.shapes), np.int64)
    for i, (v, t) in hulls.items():
        nverts[i], ntris[i] = len(v), len(t)
    poly = (pb.shapes["type"][pb.shape_a] != S.SHAPE_CAPSULE)
    bytes_pair = (64 + 56 + 16 * (nverts[pb.shape_a] + nverts[pb.shape_b]) +
                  20 * np.where(poly, ntris[pb.shape_a] + ntris[pb.shape_b], 0) + 4 + 48 * counts.astype(np.int64))
    achieved = float(bytes_pair.sum()) / (per / 1e3) / 1e9
    log(f"timed {args.steps} batches: {per:.2f} ms per {args.pairs} pairs, {int(counts.sum())} points, "
        f"points/pair histogram {np.bincount(counts, minlength=5).tolist()}")
    cpu = None
    if not args.no_cpu:
        try:
            from oracle import refharness as R
            if R.available():
                threads, host = host_threads()
                workers = R.init_jobs(max(threads - 1, 1))
                n = min(args.pairs, 1 << 16)
                sample = S.PairBatch(pb.shapes, pb.hull_points, pb.shape_a[:n], pb.shape_b[:n], pb.xf_a[:n],
                                     pb.xf_b[:n])
                R.collide_batch(sample)
                t1 = time.perf_counter()
                reps = 0
                while reps == 0 or time.perf_counter() - t1 < args.cpu_budget * 0.5:
                    R.collide_batch(sample)
                    reps += 1
                cpu_s = (time.perf_counter() - t1) / reps
                cpu = {"value": n / cpu_s, "unit": "pairs/s", "cores": workers + 1, "kind": "reference",
                       "host": host,
                       "sample": f"the first {n} pairs of the batch through NarrowPhase::Collide, {reps} passes, "
                                 f"ParallelFor with {workers} workers + caller"}
        except Exception as e:
            log(f"cpu baseline unavailable: {e}")
    return {
        "metric": "narrowphase pairs/s, 1M convex pairs",
        "value": args.pairs / (per / 1e3),
        "unit": "pairs/s",
        "n_gpus": 1, "steps": args.steps, "warmup": max(args.warmup, 1),
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.pairs} convex pairs: hull-hull (8-64 vertices), box-hull, capsule-box; "
                               "penetration U[0, 0.05] m, 10 % separated by U[0, 0.02] m",
                   "pairs": args.pairs, "contact_points": int(counts.sum()), "kernel_ms_per_batch": per},
        "roofline": {"bound": "hbm", "kernel": "k_collide_simple + k_collide_lanes + k_collide_wave", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": pmc_traffic("narrowphase1m_pmc.md") if args.pairs == 1 << 20 else None,
                     "algorithmic_bytes_per_launch": float(bytes_pair.sum())},
        "cpu_baseline": cpu,
    }


def bench_copy(args):
    """A plain device-to-device copy of known size: the reference point that
    scales the HBM byte counters (FETCH_SIZE, WRITE_SIZE) of other kernels."""
    import torch
    n = args.copy_mib << 18  # float32 elements
    a = torch.rand(n, device="cuda")
    b = torch.empty_like(a)
    for _ in range(max(args.warmup, 1)):
        b.copy_(a)
    torch.cuda.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(args.steps):
        b.copy_(a)
    t1.record()
    torch.cuda.synchronize()
    per = t0.elapsed_time(t1) / args.steps
    moved = 2 * 4 * n
    return {"metric": "device copy bandwidth", "value": moved / (per / 1e3) / 1e9, "unit": "GB/s", "n_gpus": 1,
            "steps": args.steps, "warmup": max(args.warmup, 1), "ms_per_step": per, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"copy of {args.copy_mib} MiB (read {args.copy_mib} MiB, write {args.copy_mib} MiB)",
                       "bytes_read": 4 * n, "bytes_written": 4 * n},
            "roofline": None, "cpu_baseline": None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--settle", type=int, default=300,
                    help="pile and batched: untimed steps before the warm-up, whatever --warmup is (the 32k pile "
                         "comes to rest around step 250)")
    ap.add_argument("--bodies", type=int, default=32768)
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of reference CPU work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--dump", default="", help="save the last step's contact table (npz) for host-side tuning")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned warm-start leg")
    ap.add_argument("--learn-collect", type=int, default=100, help="last warm-up steps that collect training rows")
    ap.add_argument("--learn-steps", type=int, default=2000, help="Adam steps on the device")
    ap.add_argument("--learn-batch", type=int, default=16384)
    ap.add_argument("--learn-clips", default="0,8,2", help="learned workload: target clips to train with")
    ap.add_argument("--learn-targets", choices=("solver", "converged"), default="solver",
                    help="learned workload: train on the step's own impulses or on a 64-sweep side solve's")
    ap.add_argument("--cooperative", type=int, default=0,
                    help="workgroups sharing the largest island's levels (0 auto, 1 one per group)")
    ap.add_argument("--device-bucketing", type=int, choices=(0, 1), default=0,
                    help="0: the early warm start's velocity schedule is bucketed by level on the device; 1: on the host")
    ap.add_argument("--placement", choices=("one-xcd", "spread"), default="one-xcd",
                    help="where the workgroups sharing the largest island's levels run")
    ap.add_argument("--np-waves", type=int, default=0, help="narrowphase EPA / clipping waves (0: auto)")
    ap.add_argument("--units-vel", type=int, default=0, help="most points per velocity sweep unit (0: default)")
    ap.add_argument("--units-ws", type=int, default=0, help="most points per warm-start/bias sweep unit (0: default)")
    ap.add_argument("--workload", choices=("pile", "batched", "narrowphase", "copy", "learned"), default="pile")
    ap.add_argument("--copy-mib", type=int, default=1024, help="copy workload: buffer size")
    ap.add_argument("--worlds", type=int, default=4096, help="batched workload: number of worlds")
    ap.add_argument("--pairs", type=int, default=1 << 20, help="narrowphase workload: number of pairs")
    ap.add_argument("--maps", default="", help="write the process's library map (/proc/self/maps) to this file at "
                                             "the end, to resolve the addresses of a crash report")
    args = ap.parse_args()
    if args.maps:
        import atexit

        def write_maps(path=args.maps):
            with open("/proc/self/maps") as src, open(path, "w") as dst:
                dst.write(src.read())
        atexit.register(write_maps)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the contact path of one world runs on one MI355X "
              "(multi-GPU sharding is a design note, not part of this build); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        print(json.dumps(bench_batched(args)), flush=True)
        return
    if args.workload == "narrowphase":
        print(json.dumps(bench_narrowphase(args)), flush=True)
        return
    if args.workload == "copy":
        print(json.dumps(bench_copy(args)), flush=True)
        return
    if args.workload == "learned":
        print(json.dumps(bench_learned(args)), flush=True)
        return

    import torch  # device synchronisation around the timed region
    from kizuriphysics3d_amd.world import PhysicsWorld, device_timings

    scene = S.pile(args.bodies)
    t0 = time.time()
    w = PhysicsWorld(scene)
    log(f"world: {args.bodies} bodies + 5 static, built in {time.time() - t0:.1f} s")
    learn = not args.no_learned and args.learn_collect > 0
    ctx = w.device_context()
    ctx.set_cooperative(args.cooperative)
    ctx.set_cooperative_placement(0 if args.placement == "one-xcd" else 1)
    ctx.set_device_bucketing(args.device_bucketing)
    ctx.set_sweep_units(args.units_vel, args.units_ws)
    ctx.set_narrowphase_waves(args.np_waves)
    if learn:
        ctx.mlp_init(0x4B5A0005)
    # Settle, then warm up: neither is timed.  The learned warm start
    # collects its rows over the last learn_collect of these steps.
    untimed = args.settle + args.warmup
    for i in range(untimed):
        if learn and i == max(untimed - args.learn_collect, 0):
            ctx.mlp_collect(1 << 24)  # rows of (features, impulse corrections) from these steps
        w.step_fixed()
        if (i + 1) % 50 == 0 or i + 1 == untimed:
            st = w.stats()
            log(f"{'settle' if i < args.settle else 'warmup'} {i + 1}/{untimed}: {st['num_contact_points']} points, "
                f"step {st['step_ms']:.1f} ms (contact path {st['contact_path_ms']:.1f} ms, broadphase "
                f"{st['broadphase_ms']:.1f} ms)")
    train = None
    if learn:
        t1 = time.perf_counter()
        written, held = ctx.mlp_num_samples()
        ctx.mlp_fit_normalization()
        losses = ctx.mlp_train(args.learn_steps, batch=args.learn_batch, lr=1e-3, seed=0x4B5A0005)
        train_s = time.perf_counter() - t1
        ctx.mlp_collect(0)  # the timed steps run the plain path
        train = {"rows": int(held), "adam_steps": args.learn_steps, "batch": args.learn_batch,
                 "seconds": train_s, "loss_first": float(losses[:10].mean()), "loss_last": float(losses[-100:].mean()),
                 "collected_over_steps": f"{max(untimed - args.learn_collect, 0) + 1}-{untimed}"}
        log(f"learned warm start: {held} rows, {args.learn_steps} Adam steps of {args.learn_batch} in {train_s:.2f} s, "
            f"loss {train['loss_first']:.3f} -> {train['loss_last']:.3f}")
    start_states = w.states()  # the CPU baseline restarts from here
    torch.cuda.synchronize()
    points = 0
    path_ms = 0.0
    resident_ms = 0.0
    solve_kernel_ms = 0.0
    solve_bytes = 0.0
    coop_ms = 0.0
    coop_bytes = 0.0
    coop_blocks = set()
    early_steps = 0
    levels = []
    phases = np.zeros(13)
    bpm = bpq = 0.0
    t_start = time.perf_counter()
    for i in range(args.steps):
        w.step_fixed()
        st = w.stats()
        tm = device_timings(w)
        points += st["num_contact_points"]
        path_ms += st["contact_path_ms"]
        bpm += st["broadphase_move_ms"]
        bpq += st["broadphase_query_ms"]
        # Device work that ran while the host finished the velocity schedule
        # (the early warm start) counts once.
        resident_ms += (tm["narrowphase_ms"] + tm["build_device_ms"] + tm["build_ms"] + tm["integrate_velocities_ms"] +
                        tm["solve_ms"] - tm["overlap_ms"])
        solve_kernel_ms += tm["solve_kernel_ms"]
        solve_bytes += st["num_contact_points"] * (WARM_BYTES + 8 * VEL_BYTES + 2 * BIAS_BYTES)
        coop_blocks.add(tm["cooperative_blocks"])
        early_steps += tm["early_warm_start_ms"] > 0
        if tm["cooperative_blocks"] > 1:
            coop_ms += tm["cooperative_kernel_ms"]
            coop_bytes += (tm["cooperative_ws_points"] * (WARM_BYTES + 2 * BIAS_BYTES)
                           + tm["cooperative_vel_points"] * 8 * VEL_BYTES)
        levels.append((tm["velocity_levels"], tm["bias_levels"], tm["groups"], tm["islands"]))
        phases += [tm["narrowphase_ms"], tm["build_ms"], tm["integrate_velocities_ms"], tm["solve_ms"],
                   tm["solve_kernel_ms"], st["broadphase_ms"], tm["schedule_ms"], tm["build_device_ms"],
                   tm["schedule_ws_ms"], tm["schedule_wait_ms"], tm["early_warm_start_ms"],
                   tm["schedule_velocity_after_ms"], tm["overlap_ms"]]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    # value: the path's phases with inputs in HBM (kernel times, host build
    # compute); the PCIe-inclusive rate over the host wall time of the same
    # phases (pair list up, manifold counts down, schedule up) is reported
    # beside it.
    value = points / (resident_ms / 1e3)
    pcie_value = points / (path_ms / 1e3)
    # Roofline of the dominant kernel: the cooperative sweep of the largest
    # island's group when it ran, else the per-group sweep kernel (the span
    # then also holds the body pack/unpack kernels).
    if coop_ms > 0:
        # One-XCD placement launches the cooperative sweep as k_solve_xcd.
        rl_kernel = "k_solve_xcd" if args.placement == "one-xcd" else "k_solve_coop"
        achieved = coop_bytes / (coop_ms / 1e3) / 1e9
    else:
        rl_kernel, achieved = "k_solve_groups", solve_bytes / (solve_kernel_ms / 1e3) / 1e9
    lv = np.array(levels)
    log(f"timed {args.steps} steps: {points / args.steps:.0f} points/step, contact path {resident_ms / args.steps:.2f} ms "
        f"resident, {path_ms / args.steps:.2f} ms with transfers, "
        f"solve kernels {solve_kernel_ms / args.steps:.2f} ms (cooperative {coop_ms / args.steps:.2f} ms, "
        f"workgroups {sorted(coop_blocks)}), levels vel/bias max {lv[:, 0].max()}/{lv[:, 1].max()}, "
        f"groups {lv[:, 2].mean():.0f}, islands {lv[:, 3].mean():.0f}")
    ph = phases / args.steps
    log(f"per step (ms): narrowphase {ph[0]:.2f}, build on the device {ph[7]:.2f} (compaction, islands), "
        f"build on the host {ph[1]:.2f} (schedule {ph[6]:.2f}: warm-start order {ph[8]:.2f}, "
        f"velocity order {ph[11]:.2f} after it, waited {ph[9]:.2f}), integrate velocities "
        f"{ph[2]:.2f}, solve {ph[3]:.2f} (sweep kernel {ph[4]:.2f}; early warm start {ph[10]:.2f}), "
        f"overlapped {ph[12]:.2f}, broadphase {ph[5]:.2f} (proxy moves {bpm / args.steps:.2f}, pair query "
        f"{bpq / args.steps:.2f})")
    if args.dump:
        import ctypes as C
        from kizuriphysics3d_amd import device as D
        h = w.device_context_handle()
        a, b, c3 = C.c_uint32(), C.c_uint32(), C.c_uint32()
        D.library().kzg_query_sizes(h, C.byref(a), C.byref(b), C.byref(c3))
        hdr = np.zeros((b.value, 4), np.uint32)
        pts = np.zeros((b.value, 8, 44), np.float32)
        D.library().kzg_download_constraints(h, hdr.ctypes.data_as(C.c_void_p), pts.ctypes.data_as(C.c_void_p))
        dyn = (scene.body_array()["motion_type"] == S.MOTION_DYNAMIC).astype(np.uint8)
        np.savez_compressed(args.dump, hdr=hdr, dynamic=dyn)
    residual = None
    if learn:
        from kizuriphysics3d_amd import learned as Lw
        res = Lw.compare_warm_starts(w, penetration_iterations=(4, 8))
        residual = learned_report(res, train, f"after step {untimed + args.steps}")
    # The world is closed before the CPU baseline and before interpreter
    # teardown: its helper thread is joined and its device memory, events and
    # stream are released while the runtime is whole.
    w.close()
    torch.cuda.synchronize()
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(scene, start_states, args.cpu_budget, untimed + 1)
        except Exception as e:  # the baseline is reported, never required
            log(f"cpu baseline unavailable: {e}")
    line = {
        "metric": "contacts solved/s at 32k bodies; PGS residual at fixed iters, learned vs plain",
        "value": value,
        "unit": "contacts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"random convex pile, {args.bodies} bodies, 8+2 iterations, dt 1/60, "
                               f"settled: steps {untimed + 1}-{untimed + args.steps} after {args.settle} settle "
                               f"and {args.warmup} warm-up steps",
                   "bodies": args.bodies, "settle_steps": args.settle,
                   "timed_steps": [untimed + 1, untimed + args.steps],
                   "contact_points_per_step": points / args.steps,
                   "contact_path_ms_per_step": resident_ms / args.steps,
                   "contact_path_ms_per_step_with_pcie": path_ms / args.steps,
                   "contacts_per_s_with_pcie": pcie_value},
        "roofline": {"bound": "hbm", "kernel": rl_kernel, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": pmc_traffic(f"pile32k_{rl_kernel}_pmc.md") if coop_ms > 0 else None,
                     # Per step: with the early warm start the sweep is two cooperative launches (the
                     # warm start, then the velocity iterations with the bias alongside), and
                     # `achieved` and `traffic` cover both.
                     "algorithmic_bytes_per_step": (coop_bytes / max(len(levels), 1)) if coop_ms > 0 else
                     solve_bytes / max(len(levels), 1),
                     "launches_per_step": "warm start + velocity" if early_steps else "one"},
        "cpu_baseline": cpu,
        "residual": residual,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
