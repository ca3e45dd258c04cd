"""GPU parity: the HIP contact path against the reference engine (built from
its sources into oracle/_ref) on identical seeded inputs, bit for bit.

Every test calls the product through its C-ABI (include/kzg.h) and compares
with the reference run in the same process on the host CPU."""
import numpy as np
import pytest

from kizuriphysics3d_amd import scenes as S
from kizuriphysics3d_amd import device as D
from oracle import oracle as O
from oracle import refharness as R
from tests.helpers import first_mismatch, reference_hulls

pytestmark = [pytest.mark.gpu, pytest.mark.reference]


def _ctx_for(scene, rw):
    ctx = D.DeviceContext(scene.world)
    shapes = scene.shape_array()
    ctx.upload_shapes(shapes, reference_hulls(shapes, scene.hull_array()))
    ctx.upload_bodies(D.device_bodies(scene.body_array(), rw.mass()[:, :10]))
    return ctx


@pytest.mark.parametrize("n", [4096])
def test_narrowphase_batch_bit_exact(n):
    R.init_jobs(0)
    pb = S.pair_batch(n, seed=0x4B5A0103, num_shapes=512)
    rc, rp, _ = R.collide_batch(pb)
    assert rc.max() <= 4
    ctx = D.DeviceContext(S.world_settings())
    ctx.upload_shapes(pb.shapes, reference_hulls(pb.shapes, pb.hull_points))
    gc, gp, ms = ctx.collide_batch(pb.shape_a, pb.shape_b, pb.xf_a, pb.xf_b)
    print(f"narrowphase {n} pairs: {ms:.3f} ms, points/pair hist {np.bincount(rc)}", flush=True)
    bad = np.nonzero(gc != rc)[0]
    assert len(bad) == 0, f"point counts differ at pairs {bad[:10]} ref {rc[bad[:10]]} gpu {gc[bad[:10]]}"
    for k in range(4):
        live = rc > k
        i = first_mismatch(rp[live, k], gp[live, k])
        assert i is None, f"point {k} differs at pair {np.nonzero(live)[0][i]}"


def test_plane_pairs_bit_exact():
    """CollidePlaneConvex and the plane dispatch (CollideShape.cpp:370-430,
    511-520) on the device: plane against sphere, box, capsule and hulls of
    more than 16 vertices, both orders, faces resting on the plane (more
    than 4 candidates reduced to 4), bit for bit against the reference run
    here and against its stored outputs."""
    from tests.golden_data import PLANE_N, PLANE_SEED, PLANE_SHAPES, load_fixture
    R.init_jobs(0)
    pb = S.plane_pair_batch(PLANE_N, seed=PLANE_SEED, num_shapes=PLANE_SHAPES)
    rc, rp, _ = R.collide_batch(pb)
    z = load_fixture("planes2048")
    assert np.array_equal(rc, z["counts"]) and rp[:, :4].tobytes() == z["points"].tobytes()
    ctx = D.DeviceContext(S.world_settings())
    ctx.upload_shapes(pb.shapes, reference_hulls(pb.shapes, pb.hull_points))
    gc, gp, ms = ctx.collide_batch(pb.shape_a, pb.shape_b, pb.xf_a, pb.xf_b)
    print(f"plane pairs {PLANE_N}: {ms:.3f} ms, points/pair hist {np.bincount(rc, minlength=5)}", flush=True)
    bad = np.nonzero(gc != rc)[0]
    assert len(bad) == 0, f"point counts differ at pairs {bad[:10]} ref {rc[bad[:10]]} gpu {gc[bad[:10]]}"
    for k in range(4):
        live = rc > k
        i = first_mismatch(rp[live, k], gp[live, k])
        assert i is None, f"point {k} differs at pair {np.nonzero(live)[0][i]}"


def _trajectory(scene, steps, cooperative=0, units=0, early=0, island_lds=0, np_waves=0, log_every=0,
                placement=D.COOP_ONE_XCD):
    R.init_jobs(0)
    rw = R.RefWorld(scene)
    ctx = _ctx_for(scene, rw)
    ctx.set_cooperative(cooperative)
    ctx.set_cooperative_placement(placement)
    ctx.set_early_warm_start(early)
    ctx.set_island_lds_limit(island_lds)
    ctx.set_narrowphase_waves(np_waves)
    ctx.set_sweep_units(*(units if isinstance(units, tuple) else (units, units)))
    for i in range(steps):
        if log_every and i % log_every == 0:
            print(f"{scene.name}: step {i}", flush=True)
        rw.step_phases()
        pairs = rw.pairs()
        ctx.step(pairs)
        rs, gs = rw.states(), ctx.states()
        j = first_mismatch(rs, gs)
        if j is not None:
            rh, rpt = rw.constraints()
            gh, gpt = ctx.constraints()
            msg = f"step {i}: body {j} differs: ref {rs[j]} gpu {gs[j]}; constraints {len(rh)} vs {len(gh)}"
            if len(rh) == len(gh):
                d = np.argwhere(rpt.view(np.uint32) != gpt.view(np.uint32))
                msg += f"; first constraint field diffs (c, point, field): {d[:8].tolist()}"
            pytest.fail(msg)
    hdr, _ = rw.constraints()
    print(f"{scene.name}: {steps} steps bit-exact, last step {len(hdr)} constraints "
          f"{int(hdr[:, 2].sum())} points; timings {ctx.timings()}", flush=True)


def test_pyramid_trajectory_bit_exact():
    _trajectory(S.pyramid(), 120)


def test_pile_trajectory_bit_exact():
    _trajectory(S.pile(512, seed=0x4B5A0102), 240)


def test_pile_trajectory_global_union_find_bit_exact():
    """Islands from the global-memory union-find (k_uf_init / k_uf_union, the
    path of worlds above 36,864 bodies) forced on the 512-body pile."""
    _trajectory(S.pile(512, seed=0x4B5A0102), 240, island_lds=1)


def test_pile_trajectory_narrowphase_waves_bit_exact():
    """The single-world narrowphase with a non-default wave count (64 and
    4096 waves of the EPA / clipping kernel)."""
    _trajectory(S.pile(512, seed=0x4B5A0102), 120, np_waves=64)
    _trajectory(S.pile(512, seed=0x4B5A0102), 120, np_waves=4096)


def test_pile65536_lockstep_bit_exact():
    """The 65,536-body pile of the learned warm start's workload: 75 steps in
    lockstep with the reference (the pile forming: up to ~45k contact points
    a step), islands from the global-memory union-find (above the LDS
    variant's 36,864 bodies) checked against the host union-find's count."""
    scene = S.pile(65536)
    R.init_jobs(0)
    rw = R.RefWorld(scene)
    ctx = _ctx_for(scene, rw)
    dyn = (scene.body_array()["motion_type"] == S.MOTION_DYNAMIC).astype(np.uint8)
    for i in range(75):
        rw.step_phases()
        ctx.step(rw.pairs())
        j = first_mismatch(rw.states(), ctx.states())
        assert j is None, f"step {i}: body {j} differs"
        if i % 15 == 14:
            hdr, _ = ctx.constraints()
            first = np.concatenate([[0], np.cumsum(hdr[:, 2])[:-1]]).astype(np.uint32)
            info = np.stack([hdr[:, 0], hdr[:, 1], first, hdr[:, 2] | ((hdr[:, 3] & 1) << 8)], 1).astype(np.uint32)
            _, _, _, st = D.schedule(dyn, info, True, 1 << 20)
            k, _ = rw.islands()
            assert ctx.timings()["islands"] == st[0] == k, (i, ctx.timings()["islands"], st[0], k)
            print(f"pile65536: step {i} bit-exact, {int(hdr[:, 2].sum())} points, {k} islands", flush=True)


def test_device_islands_match_host_union_find():
    """The constraint table and islands built on the device behind the
    narrowphase: the island count equals the host union-find's over the same
    table (BuildIslands numbering), along a 512-body pile."""
    scene = S.pile(512, seed=0x4B5A0102)
    R.init_jobs(0)
    rw = R.RefWorld(scene)
    ctx = _ctx_for(scene, rw)
    dyn = (scene.body_array()["motion_type"] == S.MOTION_DYNAMIC).astype(np.uint8)
    checked = 0
    for i in range(200):
        rw.step_phases()
        ctx.step(rw.pairs())
        if i % 25 == 24:
            hdr, _ = ctx.constraints()
            first = np.concatenate([[0], np.cumsum(hdr[:, 2])[:-1]]).astype(np.uint32)
            info = np.stack([hdr[:, 0], hdr[:, 1], first, hdr[:, 2] | ((hdr[:, 3] & 1) << 8)], 1).astype(np.uint32)
            _, _, _, st = D.schedule(dyn, info, True, 1 << 20)
            assert ctx.timings()["islands"] == st[0], (i, ctx.timings()["islands"], st)
            checked += 1
    assert checked == 8


@pytest.mark.parametrize("blocks", [2, 7, 64])
def test_pile_trajectory_cooperative_bit_exact(blocks):
    """The largest group's levels split over `blocks` workgroups of one XCD
    (64 requested: the XCD's 32) with a barrier through that XCD's L2
    between levels (forced even where levels are narrow, so some workgroups
    get empty slices)."""
    _trajectory(S.pile(512, seed=0x4B5A0102), 240, cooperative=blocks)


@pytest.mark.parametrize("blocks", [7, 64])
def test_pile_trajectory_cooperative_spread_bit_exact(blocks):
    """The same split over workgroups anywhere on the device, with
    agent-scope fences between levels."""
    _trajectory(S.pile(512, seed=0x4B5A0102), 240, cooperative=blocks, placement=D.COOP_SPREAD)


@pytest.mark.parametrize("placement", [D.COOP_ONE_XCD, D.COOP_SPREAD])
@pytest.mark.parametrize("simd,vbias", [(0, 0), (1, 1), (0, 1)])
def test_pile_trajectory_solver_variants_bit_exact(placement, simd, vbias):
    """SolverSettings::useSIMDSolver off (the scalar sweep in constraint
    order, ContactSolver.cpp:137-181) and useVelocityBias on (no split
    impulse: the position bias rides in the velocity bias, no bias sweeps),
    split over cooperating workgroups both ways: the trajectory stays the
    reference's."""
    sc = S.pile(512, seed=0x4B5A0102)
    sc.world["use_simd_solver"] = simd
    sc.world["use_velocity_bias"] = vbias
    _trajectory(sc, 200, cooperative=6, placement=placement)


def test_cooperative_sweep_gives_up_and_reports():
    """A cooperative sweep whose waits are bounded by one poll gives up on
    every workgroup (each kernel still ends), and the next call that waits
    for the device reports it; with the bound restored the world steps
    again."""
    from kizuriphysics3d_amd.world import PhysicsWorld
    w = PhysicsWorld(S.pile(4096, seed=0x4B5A0104))
    ctx = w.device_context()
    ctx.set_cooperative(8)
    for _ in range(30):
        w.step_fixed()
    ctx.set_spin_limit(1)
    with pytest.raises(D.KzgError, match="gave up"):
        for _ in range(3):  # the step that gives up, then the call that reports it
            w.step_fixed()
            w.states()
    ctx.set_spin_limit(0)
    for _ in range(3):
        w.step_fixed()
    assert ctx.timings()["cooperative_blocks"] == 8
    w.close()


@pytest.mark.parametrize("units,blocks", [((1, 1), 0), ((2, 2), 0), ((3, 1), 1), ((4, 4), 5), ((1, 4), 0)])
def test_pile_trajectory_unit_caps_bit_exact(units, blocks):
    """Velocity / warm-start sweep units capped at other sizes than the
    default (2, 4), split and on one workgroup: the trajectory stays the
    reference's."""
    _trajectory(S.pile(512, seed=0x4B5A0102), 240, cooperative=blocks, units=units)


@pytest.mark.parametrize("blocks", [0, 1, 7])
def test_pile_trajectory_early_warm_start_bit_exact(blocks):
    """The warm-start sweep launched before the velocity schedule is ready
    (forced at this size), with the velocity records written afterwards from
    the prepared table: the trajectory stays the reference's."""
    _trajectory(S.pile(512, seed=0x4B5A0102), 240, cooperative=blocks, early=2)


def test_early_warm_start_off_equals_on_large_pile():
    """On a 4096-body pile, with the largest group split over cooperating
    workgroups, the state after every step is the same with the early warm
    start off and forced on, the latter with its velocity schedule bucketed
    on the device (default) and on the host."""
    from kizuriphysics3d_amd.world import PhysicsWorld
    scene = S.pile(4096, seed=0x4B5A0104)
    worlds = []
    for mode, bucketing in ((1, 0), (2, 0), (2, 1)):
        w = PhysicsWorld(scene)
        w.device_context().set_early_warm_start(mode)
        w.device_context().set_device_bucketing(bucketing)
        worlds.append(w)
    early = 0
    for i in range(300):
        for w in worlds:
            w.step_fixed()
        base = worlds[0].states()
        for k, w in enumerate(worlds[1:]):
            j = first_mismatch(base, w.states())
            assert j is None, f"step {i}: body {j} differs with the early warm start (world {k + 1})"
        early += worlds[1].device_context().timings()["early_warm_start_ms"] > 0
    assert early > 0, "the warm start never started early"
    print(f"early warm start on {early} of 300 steps", flush=True)


def test_cooperative_split_invariant_large_pile():
    """On a 4096-body pile the state after every step is the same with one
    workgroup, the automatic split, 4 and 32 workgroups on one XCD and 16
    spread over the device."""
    from kizuriphysics3d_amd.world import PhysicsWorld
    scene = S.pile(4096, seed=0x4B5A0104)
    worlds = []
    for blocks, placement in ((1, 0), (0, D.COOP_ONE_XCD), (4, D.COOP_ONE_XCD), (16, D.COOP_SPREAD),
                              (32, D.COOP_ONE_XCD)):
        w = PhysicsWorld(scene)
        w.device_context().set_cooperative(blocks)
        w.device_context().set_cooperative_placement(placement)
        worlds.append(w)
    used = set()
    for i in range(300):
        for w in worlds:
            w.step_fixed()
        base = worlds[0].states()
        for w in worlds[1:]:
            j = first_mismatch(base, w.states())
            assert j is None, f"step {i}: body {j} differs between cooperative splits"
        used.add(int(worlds[1].device_context().timings()["cooperative_blocks"]))
    print(f"automatic split used {sorted(used)} workgroups", flush=True)


def test_residuals_match_oracle():
    """The PGS residual (RMS over points of each velocity sweep's impulse
    increments) of the device solve against the CPU oracle on the same
    trajectory.  The increments are bit-identical; only the order of the
    FP32 sums differs."""
    scene = S.pile(512, seed=0x4B5A0102)
    R.init_jobs(0)
    rw = R.RefWorld(scene)
    ctx = _ctx_for(scene, rw)
    shapes = scene.shape_array()
    ow = O.OracleWorld(scene, O.Derived.pack(shapes, reference_hulls(shapes, scene.hull_array()), rw.mass()[:, :10]))
    checked = 0
    for i in range(160):
        rw.step_phases()
        pairs = rw.pairs()
        ctx.step(pairs)
        ow.step(pairs)
        if i % 20 == 19:
            g, o = ctx.residuals(), ow.residuals()
            assert len(g) == len(o) == 8
            assert np.allclose(g, o, rtol=2e-5, atol=1e-9), (i, g, o)
            checked += int(o[0] > 0)
    assert checked >= 4


def test_natural_residual_matches_oracle():
    """The natural-map residual of the solved iterate (kzg_get_natural_residual,
    computed on the device in FP32 with FP64 sums) against the oracle's FP64
    restatement over the same constraints and velocities, along a 512-body
    pile: within 1e-3 relative (the per-point terms are FP32 on the
    device)."""
    scene = S.pile(512, seed=0x4B5A0102)
    R.init_jobs(0)
    rw = R.RefWorld(scene)
    ctx = _ctx_for(scene, rw)
    checked = 0
    for i in range(160):
        rw.step_phases()
        pairs = rw.pairs()
        ctx.step(pairs, stop_after=4)
        if i % 20 == 19:
            g = ctx.natural_residual()
            hdr, pts = ctx.constraints()
            o = O.natural_residual(hdr, pts, ctx.states())
            for k in ("total", "normal", "friction", "cone"):
                assert abs(g[k] - o[k]) <= 1e-3 * max(o[k], 1e-6) + 1e-7, (i, k, g, o)
            checked += o["total"] > 0
        ctx.integrate_positions()
    assert checked >= 4
