"""Sleeping and the active list (PhysicsWorld::UpdateSleeping,
src/World/PhysicsWorld.cpp:866-951; the wake-ups of BuildContactConstraints,
:473-490; BuildIslands over the active list, :587-664; BodyManager's
activation, src/Body/BodyManager.cpp:138-160).

CPU tests pin the two reformulations the device path rests on against
literal restatements of the reference's loops on random inputs:
  * the wake-up walk in constraint order equals the fixed point of the
    order-free relaxation the device runs;
  * UpdateSleeping's deactivation loop, which reads a view of the live
    active list while swap-removing from it, deactivates exactly the bodies
    that sat at the decided positions when it started, in the same order.
GPU tests step the device path beside the reference engine on worlds where
bodies fall asleep and are woken, bit for bit, active list included."""
import random

import numpy as np
import pytest

from kizuriphysics3d_amd import scenes as S

NONE = 0xFFFFFFFF


# ---- wake-ups ---------------------------------------------------------------
def _wake_walk(cons, motion, active, moving_kin):
    """BuildContactConstraints' loop (:473-490): returns the woken bodies in
    activation order."""
    act = list(active)
    woken = []

    def moving(x):
        if motion[x] == 2:
            return act[x]
        if motion[x] == 1:
            return moving_kin[x]
        return False

    for a, b in cons:
        if motion[a] == 2 and not act[a] and moving(b):
            act[a] = True
            woken.append(a)
        if motion[b] == 2 and not act[b] and moving(a):
            act[b] = True
            woken.append(b)
    return woken


def _wake_relax(cons, motion, active, moving_kin, rng):
    """The device formulation (k_wake_init / k_wake_relax / k_wake_append):
    t = 0 for bodies moving at the start, k + 1 for a sleeping body first
    reached by constraint k; relaxed in a shuffled order until stable."""
    n = len(motion)
    t = [NONE] * n
    for x in range(n):
        if motion[x] == 2 and active[x]:
            t[x] = 0
        elif motion[x] == 1 and moving_kin[x]:
            t[x] = 0
    sleeper = [motion[x] == 2 and not active[x] for x in range(n)]
    order = list(range(len(cons)))
    while True:
        rng.shuffle(order)
        changed = False
        for k in order:
            a, b = cons[k]
            if sleeper[a] and t[b] < k + 1 and t[a] > k + 1:
                t[a] = k + 1
                changed = True
            if sleeper[b] and t[a] < k + 1 and t[b] > k + 1:
                t[b] = k + 1
                changed = True
        if not changed:
            break
    out = []
    for k, (a, b) in enumerate(cons):
        if sleeper[a] and t[a] == k + 1:
            out.append(a)
        elif sleeper[b] and t[b] == k + 1:
            out.append(b)
    return out


@pytest.mark.parametrize("seed", range(40))
def test_wake_relaxation_equals_the_walk(seed):
    rng = random.Random(seed)
    n = rng.randint(2, 40)
    motion = [rng.choice([2, 2, 2, 2, 1, 0]) for _ in range(n)]
    active = [rng.random() < 0.3 for _ in range(n)]
    moving_kin = [rng.random() < 0.5 for _ in range(n)]
    cons = []
    for _ in range(rng.randint(0, 80)):
        a, b = rng.sample(range(n), 2)
        if motion[a] != 2 and motion[b] != 2:
            continue
        cons.append((min(a, b), max(a, b)) if rng.random() < 0.5 else (max(a, b), min(a, b)))
    assert _wake_relax(cons, motion, active, moving_kin, rng) == _wake_walk(cons, motion, active, moving_kin)


def test_wake_chain_runs_forward_only():
    """A sleeping chain woken from its end: each constraint only sees bodies
    woken by earlier constraints, so a chain listed backwards wakes one link."""
    motion = [2] * 5
    active = [False, False, False, False, True]
    fwd = [(4, 3), (3, 2), (2, 1), (1, 0)]
    back = list(reversed(fwd))
    assert _wake_walk(fwd, motion, active, [False] * 5) == [3, 2, 1, 0]
    assert _wake_walk(back, motion, active, [False] * 5) == [3]
    rng = random.Random(1)
    assert _wake_relax(back, motion, active, [False] * 5, rng) == [3]


# ---- the deactivation loop --------------------------------------------------
def _deactivate_literal(lst, cap_tail, decide_root):
    """UpdateSleeping's loop (:932-941) as the reference runs it: `view`
    keeps the data pointer and the size n it had when the loop started,
    RemoveSwap (include/Kizuri/Core/Memory.h:242-250) moves the last element
    into the hole and leaves the old last slot as it was.  decide_root[i]:
    the island of list position i qualifies.  Returns the final list and the
    deactivated bodies in order."""
    data = list(lst) + list(cap_tail)
    size = len(lst)
    n = size
    active = {x: True for x in lst}
    out = []
    for i in range(n):
        x = data[i]
        if decide_root[i]:
            if active.get(x, False):
                active[x] = False
                for p in range(size):
                    if data[p] == x:
                        data[p] = data[size - 1]
                        size -= 1
                        break
                out.append(x)
    return data[:size], out


def _deactivate_model(lst, decide_root):
    """The device/host split: body lst[i] sleeps iff position i qualifies;
    the swap-removes are replayed in position order with a position map."""
    data = list(lst)
    pos = {x: i for i, x in enumerate(lst)}
    size = len(lst)
    out = []
    for i, x in enumerate(lst):
        if not decide_root[i]:
            continue
        p = pos.pop(x)
        last = data[size - 1]
        data[p] = last
        if last != x:
            pos[last] = p
        size -= 1
        out.append(x)
    return data[:size], out


@pytest.mark.parametrize("seed", range(60))
def test_deactivation_loop_reads_the_starting_occupants(seed):
    rng = random.Random(seed)
    n = rng.randint(1, 60)
    lst = rng.sample(range(1000), n)
    tail = rng.sample(range(1000, 1100), rng.randint(0, 5))  # capacity past the size
    decide = [rng.random() < rng.choice([0.1, 0.5, 0.9]) for _ in range(n)]
    assert _deactivate_literal(lst, tail, decide) == _deactivate_model(lst, decide)


def test_sleeping_scene_settings():
    sc = S.sleeping_pile(64, late=3)
    assert int(sc.world["allow_sleeping"]) == 1
    b = sc.body_array()
    assert len(b) == 5 + 64 + 3 and np.all(b["allow_sleeping"] == 1)


# ---- GPU: the device path beside the reference ------------------------------
def _pyramid_with_late_boxes():
    sc = S.pyramid(10)
    sc.name = "sleeping_pyramid10"
    sc.world["allow_sleeping"] = 1
    box = sc.bodies[1]["shape"]
    for x, y in [(0.3, 25.0), (-2.2, 40.0), (1.7, 55.0)]:
        b = S.default_body(box)
        b["position"] = (x, y, 0.2)
        b["mass_override"] = 1.0
        sc.add_body(b)
    return sc


def _shared_bodies(rw) -> int:
    """Dynamic bodies the reference's last BuildIslands put in two islands."""
    hdr, _ = rw.constraints()
    _, isl = rw.islands()
    dyn = rw.mass()[:, 0] > 0
    owner, shared = {}, set()
    for c in range(len(isl)):
        if isl[c] == NONE:
            continue
        for x in (int(hdr[c, 0]), int(hdr[c, 1])):
            if dyn[x] and owner.setdefault(x, int(isl[c])) != isl[c]:
                shared.add(x)
    return len(shared)


def _sleeping_trajectory(scene, steps, **opts):
    from kizuriphysics3d_amd import device as D
    from oracle import refharness as R
    from tests.helpers import first_mismatch, reference_hulls
    R.init_jobs(0)
    rw = R.RefWorld(scene)
    ctx = D.DeviceContext(scene.world)
    shapes = scene.shape_array()
    ctx.upload_shapes(shapes, reference_hulls(shapes, scene.hull_array()))
    ctx.upload_bodies(D.device_bodies(scene.body_array(), rw.mass()[:, :10]))
    if "cooperative" in opts:
        ctx.set_cooperative(opts["cooperative"])
    slept = woke = shared_steps = 0
    prev = len(scene.bodies)
    for i in range(steps):
        rw.step_phases()
        ctx.step(rw.pairs())
        if opts.get("count_shared"):
            shared_steps += _shared_bodies(rw) > 0
        ra, ga = rw.active(), ctx.active_bodies()
        assert np.array_equal(ra, ga), f"step {i}: active lists differ ({len(ra)} vs {len(ga)})"
        rs, gs = rw.states(), ctx.states()
        j = first_mismatch(rs, gs)
        assert j is None, f"step {i}: body {j} differs: ref {rs[j]} gpu {gs[j]}"
        slept += max(0, prev - len(ra))
        woke += max(0, len(ra) - prev)
        prev = len(ra)
    print(f"{scene.name}: {steps} steps bit-exact, {slept} sleep / {woke} wake events, "
          f"{len(scene.bodies) - prev} asleep at the end, {shared_steps} steps with shared bodies", flush=True)
    return (slept, woke, shared_steps) if opts.get("count_shared") else (slept, woke)


@pytest.mark.gpu
@pytest.mark.reference
def test_sleeping_pile_trajectory_bit_exact():
    slept, woke = _sleeping_trajectory(S.sleeping_pile(256), 400)
    assert slept > 0 and woke > 0


@pytest.mark.gpu
@pytest.mark.reference
def test_sleeping_pyramid_trajectory_bit_exact():
    slept, woke = _sleeping_trajectory(_pyramid_with_late_boxes(), 400)
    assert slept > 0 and woke > 0


@pytest.mark.gpu
@pytest.mark.reference
def test_sleeping_field_islands_sharing_a_body_bit_exact():
    """The box field: sleeping bodies routed through stale indices put one
    dynamic body into two islands on some steps (the reference solves those
    islands one after another); the device solves them in generations."""
    slept, woke, shared = _sleeping_trajectory(S.sleeping_field(10), 300, count_shared=True)
    assert slept > 0 and woke > 0 and shared > 0


def _world_vs_reference(ours, ref, steps, label):
    from tests.helpers import first_mismatch
    ctx = ours.device_context()
    asleep = 0
    for i in range(steps):
        ref.step()
        ours.step_fixed()
        ra, ga = ref.active(), ctx.active_bodies()
        assert np.array_equal(ra, ga), f"{label} step {i}: active lists differ ({len(ra)} vs {len(ga)})"
        j = first_mismatch(ref.states(), ours.states())
        assert j is None, f"{label} step {i}: body {j} differs"
        asleep = max(asleep, ours.num_bodies - len(ga))
        assert ours.stats()["num_active_bodies"] == len(ga)
    return asleep


@pytest.mark.gpu
@pytest.mark.reference
def test_world_with_sleeping_matches_reference():
    """The whole step through the world layer (its own broadphase tree, the
    pair query on the device, proxies moved in active-list order as
    UpdateBroadPhase moves them, PhysicsWorld.cpp:333-345) beside the
    reference on the sleeping box field."""
    from kizuriphysics3d_amd.world import PhysicsWorld
    from oracle import refharness as R
    scene = S.sleeping_field(8)
    R.init_jobs(0)
    asleep = _world_vs_reference(PhysicsWorld(scene), R.RefWorld(scene), 240, "field")
    assert asleep > 0


@pytest.mark.gpu
@pytest.mark.reference
def test_kzp1_resume_with_sleeping_bodies():
    """A KZP1 record saved by the reference while part of the world sleeps:
    both load it (inactive bodies removed from the active list as the loader
    creates them, velocities zeroed, Serializer.cpp:230-234) and step on
    bit-identically, active lists included; the record this library writes
    back carries the same activity bytes."""
    from kizuriphysics3d_amd.world import PhysicsWorld
    from oracle import refharness as R
    scene = S.sleeping_field(8)
    R.init_jobs(0)
    rw = R.RefWorld(scene)
    for _ in range(90):
        rw.step()
    assert len(rw.active()) < len(scene.bodies)
    rec = rw.save_kzp1()
    ref2 = R.RefWorld.load_kzp1(rec, scene.world)
    ours = PhysicsWorld.load(rec, world=scene.world)
    assert np.array_equal(ref2.active(), ours.device_context().active_bodies())
    _world_vs_reference(ours, ref2, 120, "resume")
    # Reloading a record written after the run gives both the same list.
    rec_ref = ref2.save_kzp1()
    again = PhysicsWorld.load(rec_ref, world=scene.world)
    ref3 = R.RefWorld.load_kzp1(rec_ref, scene.world)
    assert np.array_equal(again.device_context().active_bodies(), ref3.active())
    assert np.array_equal(np.sort(ref3.active()), np.sort(ref2.active()))
