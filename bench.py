#!/usr/bin/env python3
"""Contact-path benchmark: contacts solved per second on the 32k-body pile.

Workload: the random convex pile of 32,768 bodies (25 % spheres, boxes,
capsules, convex hulls of 16-32 points; dropped from y in [2, 60] into a
64 x 64 m walled bin; density 1000; FP32; 8 velocity + 2 position
iterations; dt = 1/60; sleeping off).  W warm-up StepFixed calls let the pile
form, then K StepFixed calls are timed.  A "step" is one StepFixed: host
broadphase, then narrowphase + constraint build + solve + integration on the
GPU.  The metric is sum(contact points) / sum(time of narrowphase + build +
solve), measured on the host clock with the device synchronised (the host
part of the build - compaction and the solve schedule - is inside it).

Also reported: the roofline of the dominant kernel (HIP events on the
library's stream) and the reference CPU path (the reference engine built from
its own sources, oracle/_ref) on a bounded sample of the same pile.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--bodies N]
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


def cpu_baseline(scene, states, budget_s: float, threads: int):
    """The reference engine on the GPU run's pile: restart it from the same
    body states and time its narrowphase + build + solve phases."""
    from oracle import refharness as R
    if not R.available():
        return None
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
            "sample": f"{n} StepFixed calls of the reference on the same {len(states)}-body pile restarted from "
                      f"the GPU run's body state (after one untimed warm-start step); phases Collide + "
                      f"BuildContactConstraints + IntegrateVelocities + SolveContacts; JobSystem with {workers} "
                      f"workers + caller on {cpu_model()}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--bodies", type=int, default=32768)
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of reference CPU work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--dump", default="", help="save the last step's contact table (npz) for host-side tuning")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned warm-start leg")
    ap.add_argument("--learn-collect", type=int, default=100, help="last warm-up steps that collect training rows")
    ap.add_argument("--learn-steps", type=int, default=2000, help="Adam steps on the device")
    ap.add_argument("--learn-batch", type=int, default=16384)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the contact path of one world runs on one MI355X "
              "(multi-GPU sharding is a design note, not part of this build); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch  # device synchronisation around the timed region
    from kizuriphysics3d_amd.world import PhysicsWorld, device_timings

    scene = S.pile(args.bodies)
    t0 = time.time()
    w = PhysicsWorld(scene)
    log(f"world: {args.bodies} bodies + 5 static, built in {time.time() - t0:.1f} s")
    learn = not args.no_learned and args.learn_collect > 0
    ctx = w.device_context()
    if learn:
        ctx.mlp_init(0x4B5A0005)
    for i in range(args.warmup):
        if learn and i == max(args.warmup - args.learn_collect, 0):
            ctx.mlp_collect(1 << 24)  # rows of (features, impulse corrections) from these steps
        w.step_fixed()
        if (i + 1) % 50 == 0:
            st = w.stats()
            log(f"warmup {i + 1}/{args.warmup}: {st['num_contact_points']} points, step {st['step_ms']:.1f} ms "
                f"(contact path {st['contact_path_ms']:.1f} ms, broadphase {st['broadphase_ms']:.1f} ms)")
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
                 "collected_over_steps": f"{args.warmup - args.learn_collect + 1}-{args.warmup}"}
        log(f"learned warm start: {held} rows, {args.learn_steps} Adam steps of {args.learn_batch} in {train_s:.2f} s, "
            f"loss {train['loss_first']:.3f} -> {train['loss_last']:.3f}")
    torch.cuda.synchronize()
    points = 0
    path_ms = 0.0
    resident_ms = 0.0
    solve_kernel_ms = 0.0
    solve_bytes = 0.0
    levels = []
    phases = np.zeros(6)
    t_start = time.perf_counter()
    for i in range(args.steps):
        w.step_fixed()
        st = w.stats()
        tm = device_timings(w)
        points += st["num_contact_points"]
        path_ms += st["contact_path_ms"]
        resident_ms += tm["narrowphase_ms"] + tm["build_ms"] + tm["integrate_velocities_ms"] + tm["solve_ms"]
        solve_kernel_ms += tm["solve_kernel_ms"]
        solve_bytes += st["num_contact_points"] * (WARM_BYTES + 8 * VEL_BYTES + 2 * BIAS_BYTES)
        levels.append((tm["velocity_levels"], tm["bias_levels"], tm["groups"], tm["islands"]))
        phases += [tm["narrowphase_ms"], tm["build_ms"], tm["integrate_velocities_ms"], tm["solve_ms"],
                   tm["solve_kernel_ms"], st["broadphase_ms"]]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    # value: the path's phases with inputs in HBM (kernel times, host build
    # compute); the PCIe-inclusive rate over the host wall time of the same
    # phases (pair list up, manifold counts down, schedule up) is reported
    # beside it.
    value = points / (resident_ms / 1e3)
    pcie_value = points / (path_ms / 1e3)
    achieved = solve_bytes / (solve_kernel_ms / 1e3) / 1e9
    lv = np.array(levels)
    log(f"timed {args.steps} steps: {points / args.steps:.0f} points/step, contact path {resident_ms / args.steps:.2f} ms "
        f"resident, {path_ms / args.steps:.2f} ms with transfers, "
        f"solve kernel {solve_kernel_ms / args.steps:.2f} ms, levels vel/bias max {lv[:, 0].max()}/{lv[:, 1].max()}, "
        f"groups {lv[:, 2].mean():.0f}, islands {lv[:, 3].mean():.0f}")
    ph = phases / args.steps
    log(f"per step (ms): narrowphase {ph[0]:.2f}, build (incl. host schedule) {ph[1]:.2f}, integrate velocities "
        f"{ph[2]:.2f}, solve {ph[3]:.2f} (sweep kernel {ph[4]:.2f}), host broadphase {ph[5]:.2f}")
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
        cache8 = float(res["cache"]["residual"][-1])
        residual = {"definition": "RMS over contact points of the impulse increments of sweep k (N s)",
                    "state": f"after step {args.warmup + args.steps}", "points": res["cache"]["points"],
                    "k": list(range(1, len(res["cache"]["residual"]) + 1)), "train": train}
        for name, r in res.items():
            residual[name] = [float(x) for x in r["residual"]]
            residual[name + "_penetration"] = r["penetration"]
            residual[name + "_sweeps_to_cache8"] = Lw.iterations_to_reach(r["residual"], cache8)
            log(f"residual {name:13s} k=1..8 {np.array2string(r['residual'], precision=3)} "
                f"penetration k=4 {r['penetration'][4]}")
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(scene, w.states(), args.cpu_budget, min(os.cpu_count() or 1, 16))
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
                               f"steps {args.warmup + 1}-{args.warmup + args.steps}",
                   "bodies": args.bodies, "contact_points_per_step": points / args.steps,
                   "contact_path_ms_per_step": resident_ms / args.steps,
                   "contact_path_ms_per_step_with_pcie": path_ms / args.steps,
                   "contacts_per_s_with_pcie": pcie_value},
        "roofline": {"bound": "hbm", "kernel": "k_solve_groups", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "cpu_baseline": cpu,
        "residual": residual,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
