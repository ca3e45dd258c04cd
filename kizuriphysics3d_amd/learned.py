"""Learned warm start: training from a running world and the comparison of
warm-start modes by PGS residual and penetration.

The reference initialises every contact point's accumulated impulses from
its manifold cache (src/World/PhysicsWorld.cpp:492-530) or, with KZ_NOWS,
with zeros (:493).  The network (kzg_mlp_* in include/kzg.h) predicts them
from contact-frame features instead, for every point (WARM_LEARNED) or only
where the cache has no match (WARM_LEARNED_MISS).  The residual after k
velocity iterations is measured two ways: the RMS over contact points of the
impulse increments of sweep k (kzg_get_residuals: one 8-iteration solve gives
the curve for k = 1..8), and the natural-map residual of the iterate after k
sweeps (kzg_get_natural_residual: normal complementarity, projected friction
and cone terms, one solve per k).
"""
from __future__ import annotations

import numpy as np

from . import device as D
from . import scenes as S

MODE_NAMES = {D.WARM_CACHE: "cache", D.WARM_COLD: "cold", D.WARM_LEARNED: "learned",
              D.WARM_LEARNED_MISS: "learned_miss"}


def penetration(ctx: D.DeviceContext, pairs: np.ndarray) -> dict:
    """Penetration depth over the manifold points of `pairs` at the
    context's current state: maximum, 99th percentile and mean of
    max(0, -separation).  The narrowphase reproduces the reference's
    separated-pair clipping, whose GJK normal can point the wrong way for
    near-touching boxes and hulls and then reports separations of several
    metres; the percentile and the mean are the figures such points do not
    dominate."""
    ctx.narrowphase(pairs)
    counts, pts = ctx.manifolds()
    live = np.arange(4)[None, :] < counts[:, None]
    depth = np.maximum(0.0, -pts["separation"][live].astype(np.float64))
    if not len(depth):
        return {"max": 0.0, "p99": 0.0, "mean": 0.0}
    return {"max": float(depth.max()), "p99": float(np.quantile(depth, 0.99)), "mean": float(depth.mean())}


def compare_warm_starts(world, modes=(D.WARM_COLD, D.WARM_CACHE, D.WARM_LEARNED, D.WARM_LEARNED_MISS),
                        iterations: int = 8, position_iterations: int = 2, penetration_iterations=(),
                        natural: bool = True, dt: float = S.DT, converged_iterations: int = 0) -> dict:
    """From the world's current state, one contact-path step per warm-start
    mode (the cache is left as it is), each restarting from the same body
    states.  Returns per mode the sweep-increment residual after k =
    1..iterations sweeps ("residual"), with natural=True the natural-map
    residual of the iterate after k sweeps ("natural": a solve with k
    velocity iterations each, kzg_get_natural_residual), and, for each k in
    penetration_iterations, the penetration statistics after a step solved
    with k velocity iterations (measured over the step's broadphase pairs).
    With converged_iterations > 0 a last entry, "converged", starts from the
    impulses a solve of the same step with that many sweeps reaches: the
    bound of any warm start that predicts each point's impulses (the cache
    is swapped for one holding those impulses and restored after).
    The world's state is restored afterwards."""
    ctx = world.device_context()
    snap = world.states()
    world.update_broadphase()
    pairs = world.pairs()
    out = {}
    try:
        for mode in modes:
            res = {}
            ctx.set_iterations(iterations, position_iterations)
            ctx.upload_states(snap)
            ctx.narrowphase(pairs)
            ctx.build_constraints(mode | D.WARM_NO_STORE)
            ctx.integrate_velocities(dt)
            ctx.solve(dt)
            res["residual"] = ctx.residuals()[:iterations]
            res["points"] = ctx.sizes()[2]
            _side_solve_done(ctx, snap, dt)
            if natural:
                nat = []
                for k in range(1, iterations + 1):
                    ctx.set_iterations(k, position_iterations)
                    ctx.upload_states(snap)
                    ctx.narrowphase(pairs)
                    ctx.build_constraints(mode | D.WARM_NO_STORE)
                    ctx.integrate_velocities(dt)
                    ctx.solve(dt)
                    nat.append(ctx.natural_residual())
                    _side_solve_done(ctx, snap, dt)
                res["natural"] = nat
            pen = {}
            for k in penetration_iterations:
                ctx.set_iterations(k, position_iterations)
                ctx.upload_states(snap)
                ctx.narrowphase(pairs)
                ctx.build_constraints(mode | D.WARM_NO_STORE)
                ctx.integrate_velocities(dt)
                ctx.solve(dt)
                ctx.integrate_positions(dt)
                pen[int(k)] = penetration(ctx, pairs)
                ctx.upload_states(snap)
            res["penetration"] = pen
            out[MODE_NAMES.get(mode, str(mode))] = res
        if converged_iterations:
            image = ctx.cache_image()
            try:
                ctx.set_iterations(converged_iterations, position_iterations)
                ctx.upload_states(snap)
                ctx.narrowphase(pairs)
                ctx.build_constraints(D.WARM_CACHE)  # stores this step's converged impulses
                ctx.integrate_velocities(dt)
                ctx.solve(dt)
                _side_solve_done(ctx, snap, dt)
                res = {}
                ctx.set_iterations(iterations, position_iterations)
                ctx.upload_states(snap)
                ctx.narrowphase(pairs)
                ctx.build_constraints(D.WARM_CACHE | D.WARM_NO_STORE)
                ctx.integrate_velocities(dt)
                ctx.solve(dt)
                res["residual"] = ctx.residuals()[:iterations]
                res["points"] = ctx.sizes()[2]
                _side_solve_done(ctx, snap, dt)
                if natural:
                    nat = []
                    for k in range(1, iterations + 1):
                        ctx.set_iterations(k, position_iterations)
                        ctx.upload_states(snap)
                        ctx.narrowphase(pairs)
                        ctx.build_constraints(D.WARM_CACHE | D.WARM_NO_STORE)
                        ctx.integrate_velocities(dt)
                        ctx.solve(dt)
                        nat.append(ctx.natural_residual())
                        _side_solve_done(ctx, snap, dt)
                    res["natural"] = nat
                res["penetration"] = {}
                out["converged"] = res
            finally:
                ctx.set_cache_image(image)
    finally:
        ctx.set_iterations(world_iterations(world)[0], world_iterations(world)[1])
        ctx.upload_states(snap)
    return out


def _side_solve_done(ctx, snap, dt: float = S.DT):
    """After a solve that is not the world's step: integrate the positions,
    which clears the bias sweeps' pseudo velocities as the end of a step
    does, and put the states back.  (A state upload leaves pseudo
    velocities alone: between steps they are zero.)"""
    ctx.integrate_positions(dt)
    ctx.upload_states(snap)


def world_iterations(world):
    w = world._world
    return int(w["velocity_iterations"]), int(w["position_iterations"])


def iterations_to_reach(curve: np.ndarray, target: float):
    """Smallest k (1-based) whose residual is at or below target, else None."""
    hit = np.nonzero(np.asarray(curve) <= target)[0]
    return int(hit[0]) + 1 if len(hit) else None


def train_on_world(world, collect_steps: int, train_steps: int, batch: int = 4096, lr: float = 1e-3,
                   capacity: int = 1 << 21, seed: int = 0x4B5A0005, init: bool = True, clip: float = 0.0) -> np.ndarray:
    """Step the world `collect_steps` times with the cache warm start while
    collecting (features, final impulses) per contact point, fit the feature
    normalisation, then run `train_steps` Adam steps on the device.  Returns
    the per-step training loss."""
    ctx = world.device_context()
    if init:
        ctx.mlp_init(seed)
    ctx.mlp_collect(capacity)
    for _ in range(collect_steps):
        world.step_fixed()
    ctx.mlp_fit_normalization()
    ctx.mlp_set_target_clip(clip)
    losses = ctx.mlp_train(train_steps, batch=batch, lr=lr, seed=seed)
    ctx.mlp_collect(0)
    return losses


def _step_with_converged_rows(w, ctx, iterations: int, dt: float = S.DT):
    """One world step whose training rows come from a side solve of the same
    contacts with `iterations` sweeps (the cache untouched); the step itself
    then runs as kzw_step runs it and appends no rows."""
    vel_it, pos_it = world_iterations(w)
    snap = w.states()
    w.update_broadphase()
    pairs = w.pairs()
    ctx.set_iterations(iterations, pos_it)
    ctx.upload_states(snap)
    ctx.narrowphase(pairs)
    ctx.build_constraints(D.WARM_CACHE | D.WARM_NO_STORE)
    ctx.integrate_velocities(dt)
    ctx.solve(dt)
    _side_solve_done(ctx, snap, dt)
    ctx.set_iterations(vel_it, pos_it)
    ctx.narrowphase(pairs)
    ctx.build_constraints(D.WARM_CACHE | D.WARM_NO_COLLECT)
    ctx.integrate_velocities(dt)
    ctx.solve(dt)
    ctx.integrate_positions(dt)
    ctx.update_sleeping(dt)
    w.set_states(ctx.states())


def study(bodies: int = 65536, settle: int = 600, collect: int = 100, train_steps: int = 3000,
          batch: int = 16384, lr: float = 1e-3, clips=(0.0, 8.0, 2.0), converged_iterations: int = 64,
          seed: int = 0x4B5A0005, targets: str = "solver", load_features: bool = True, log=print) -> dict:
    """The learned warm start on the random convex pile of `bodies` bodies:
    settle it (the last `collect` steps collect training rows), then, on the
    same state, the residual curves of the cold start, the cache, a start from
    the impulses of a `converged_iterations`-sweep solve (the bound of a
    per-point predictor), and the network trained from one initialisation
    with each target clip in `clips` (learned everywhere, and on cache misses
    only).  targets: "solver" trains on the impulses the step's own solve
    stores; "converged" on those of a `converged_iterations`-sweep side solve
    of each collected step.  load_features: the rows carry the body-load
    features (14-15), and every clip is trained twice, without them (their
    inverse deviations set to 0, which is the 14-feature network) and with
    them (variant names ending in "_load")."""
    from .world import PhysicsWorld
    w = PhysicsWorld(S.pile(bodies))
    ctx = w.device_context()
    ctx.mlp_init(seed)
    ctx.set_learned_features(D.FEATURES_LOAD if load_features else 0)
    for i in range(settle):
        if i == max(settle - collect, 0):
            ctx.mlp_collect(1 << 24)
        if targets == "converged" and i >= settle - collect:
            _step_with_converged_rows(w, ctx, converged_iterations)
        else:
            w.step_fixed()
        if (i + 1) % 100 == 0:
            st = w.stats()
            log(f"settle {i + 1}/{settle}: {st['num_contact_points']} points, step {st['step_ms']:.1f} ms")
    _, held = ctx.mlp_num_samples()
    rows = ctx.mlp_samples(0, min(int(held), 1 << 20))
    tg = rows[:, 16:19]
    target_stats = {"rows": int(held), "abs_p50": float(np.percentile(np.abs(tg), 50)),
                    "abs_p99": float(np.percentile(np.abs(tg), 99)), "abs_max": float(np.abs(tg).max()),
                    "hit_fraction": float(rows[:, 13].mean())}
    log(f"training rows: {target_stats}")
    ctx.mlp_fit_normalization()
    norm = ctx.mlp_normalization()
    out = {"bodies": bodies, "settle": settle, "collected_over_steps": f"{settle - collect + 1}-{settle}",
           "target_kind": targets, "targets": target_stats}
    base = compare_warm_starts(w, modes=(D.WARM_COLD, D.WARM_CACHE), converged_iterations=converged_iterations)
    out["cold"], out["cache"], out["converged"] = base["cold"], base["cache"], base["converged"]
    out["variants"] = {}
    mean, istd = norm
    no_load = np.array(istd, np.float32)
    no_load[14:16] = 0.0
    for clip in clips:
        for load in ((False, True) if load_features else (False,)):
            ctx.mlp_init(seed)
            ctx.mlp_set_normalization(mean, istd if load else no_load)
            ctx.mlp_set_target_clip(clip)
            losses = ctx.mlp_train(train_steps, batch=batch, lr=lr, seed=seed)
            r = compare_warm_starts(w, modes=(D.WARM_LEARNED, D.WARM_LEARNED_MISS))
            name = f"clip{clip:g}" + ("_load" if load else "")
            out["variants"][name] = {"clip": clip, "load_features": load, "loss_first": float(losses[:10].mean()),
                                     "loss_last": float(losses[-100:].mean()), **r}
            log(f"{name}: loss {losses[:10].mean():.3f} -> {losses[-100:].mean():.3f}")
    ctx.mlp_collect(0)
    w.close()
    return out


def study_summary(res: dict) -> dict:
    """Per start: increment and natural curves, and the sweeps each takes to
    reach the cache's 8-sweep value of both."""
    cache_inc = float(res["cache"]["residual"][-1])
    cache_nat = float(res["cache"]["natural"][-1]["total"])
    rows = {}

    def add(name, r):
        nat = np.array([x["total"] for x in r["natural"]])
        rows[name] = {"increment": [float(x) for x in r["residual"]], "natural": nat.tolist(),
                      "sweeps_to_cache8_increment": iterations_to_reach(r["residual"], cache_inc),
                      "sweeps_to_cache8_natural": iterations_to_reach(nat, cache_nat),
                      "natural_terms_k8": r["natural"][-1]}
    for name in ("cold", "cache", "converged"):
        add(name, res[name])
    for vname, v in res["variants"].items():
        add(f"learned_{vname}", v["learned"])
        add(f"learned_miss_{vname}", v["learned_miss"])
    return rows
