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
                        natural: bool = True, dt: float = S.DT) -> dict:
    """From the world's current state, one contact-path step per warm-start
    mode (the cache is left as it is), each restarting from the same body
    states.  Returns per mode the sweep-increment residual after k =
    1..iterations sweeps ("residual"), with natural=True the natural-map
    residual of the iterate after k sweeps ("natural": a solve with k
    velocity iterations each, kzg_get_natural_residual), and, for each k in
    penetration_iterations, the penetration statistics after a step solved
    with k velocity iterations (measured over the step's broadphase pairs).
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
            res["penetration"] = pen
            out[MODE_NAMES.get(mode, str(mode))] = res
    finally:
        ctx.set_iterations(world_iterations(world)[0], world_iterations(world)[1])
        ctx.upload_states(snap)
    return out


def world_iterations(world):
    w = world._world
    return int(w["velocity_iterations"]), int(w["position_iterations"])


def iterations_to_reach(curve: np.ndarray, target: float):
    """Smallest k (1-based) whose residual is at or below target, else None."""
    hit = np.nonzero(np.asarray(curve) <= target)[0]
    return int(hit[0]) + 1 if len(hit) else None


def train_on_world(world, collect_steps: int, train_steps: int, batch: int = 4096, lr: float = 1e-3,
                   capacity: int = 1 << 21, seed: int = 0x4B5A0005, init: bool = True) -> np.ndarray:
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
    losses = ctx.mlp_train(train_steps, batch=batch, lr=lr, seed=seed)
    ctx.mlp_collect(0)
    return losses
