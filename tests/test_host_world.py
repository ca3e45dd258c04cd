"""The product's host code (convex hull building, mass properties and the
dynamic AABB tree broadphase) against the reference fixtures, bit for bit.
These run on the CPU through the C-ABI of libkizuri_amd.so (host-only world,
no GPU)."""
import numpy as np

from kizuriphysics3d_amd import scenes as S
from kizuriphysics3d_amd.world import PhysicsWorld
from tests.golden_data import PILE_N, PILE_SEED, fixture_hulls, fixture_pairs, load_fixture


def _shapes_scene():
    rng = S.SplitMix64(0x4B5A0404)
    sc = S.Scene("shapes", S.world_settings())
    for i in range(64):
        b = S.random_convex_body(sc, rng, i % 4)
        b["position"] = (3.0 * i, 5.0, 0.0)
        sc.add_body(b)
    return sc


def test_hulls_match_reference_create():
    z = load_fixture("shapes64")
    sc = _shapes_scene()
    w = PhysicsWorld(sc, device=-1)
    ref = fixture_hulls(z)
    assert ref
    for i, (rv, rt) in ref.items():
        v, t = w.hull(i)
        assert v.tobytes() == rv.tobytes(), f"hull {i}: vertex order/values differ"
        assert np.array_equal(t, rt), f"hull {i}: triangles differ"


def test_mass_properties_match_reference():
    z = load_fixture("shapes64")
    w = PhysicsWorld(_shapes_scene(), device=-1)
    assert w.mass().tobytes() == z["mass"][:, :10].tobytes()


def test_pile_mass_and_hulls_match_reference():
    z = load_fixture("pile256")
    sc = S.pile(PILE_N, seed=PILE_SEED)
    w = PhysicsWorld(sc, device=-1)
    assert w.mass().tobytes() == z["mass"][:, :10].tobytes()
    for i, (rv, rt) in fixture_hulls(z).items():
        v, t = w.hull(i)
        assert v.tobytes() == rv.tobytes() and np.array_equal(t, rt)


def _broadphase_replay(name, scene, steps):
    """Feed the reference's body states of step i-1 and compare the pair list
    of step i (the tree evolves with every MoveProxy, so every step is
    replayed in order)."""
    z = load_fixture(name)
    w = PhysicsWorld(scene, device=-1)
    checkpoints = sorted(int(c) for c in z["checkpoints"])
    prev = None
    for i in range(steps):
        if prev is not None:
            w.set_states(prev)
        w.update_broadphase()
        p = w.pairs()
        ref = fixture_pairs(z, i)
        assert np.array_equal(p, ref), f"pair list differs at step {i}: {len(p)} vs {len(ref)}"
        if i in checkpoints:
            prev = z[f"state_{i}"]
        else:
            # states between checkpoints are not stored: replay only while
            # consecutive checkpoints exist
            break


def test_broadphase_initial_pairs_match_reference():
    _broadphase_replay("pyramid20", S.pyramid(), 3)
    _broadphase_replay("pile256", S.pile(PILE_N, seed=PILE_SEED), 3)


import pytest  # noqa: E402


@pytest.mark.reference
@pytest.mark.parametrize("scene_fn,steps", [(lambda: S.pyramid(), 300), (lambda: S.pile(2048, seed=77), 300)])
def test_broadphase_pair_order_matches_reference_lockstep(scene_fn, steps):
    """Reference and product broadphase in lockstep on the reference's states:
    the pair list (the narrowphase's input order) must be identical at every
    step while the tree rebalances."""
    from oracle import refharness as R
    sc = scene_fn()
    rw = R.RefWorld(sc)
    w = PhysicsWorld(sc, device=-1)
    for i in range(steps):
        w.set_states(rw.states())
        rw.step()
        w.update_broadphase()
        assert np.array_equal(w.pairs(), rw.pairs()), f"pair list differs at step {i}"


@pytest.mark.reference
def test_large_broadphase_pair_order_matches_the_reference():
    """A tree beyond the threaded-walk threshold (16,384 bodies): the pairs,
    walked in parallel subtrees and concatenated, come out in the
    reference's order."""
    from oracle import refharness as R
    from kizuriphysics3d_amd.world import PhysicsWorld
    scene = S.pile(16384, seed=0x4B5A0701, width=40.0, y_range=(1.0, 20.0))
    R.init_jobs(0)
    rw = R.RefWorld(scene)
    ours = PhysicsWorld(scene, device=-1)
    rw.step_phases(stop_after=0)
    ours.update_broadphase()
    rp, op = rw.pairs(), ours.pairs()
    assert len(rp) > 10000
    assert np.array_equal(rp, op)


def test_broadphase_update_is_the_same_on_any_number_of_host_threads():
    """kzw_set_host_threads: the proxies' tight boxes and fat-box tests run
    on host threads, the re-insertions in list order, so the tree and the
    pair list are the same as on one thread, step after step."""
    sc = S.pile(6000, seed=0x4B5A0111, width=24.0, y_range=(0.5, 12.0))
    worlds = [PhysicsWorld(sc, device=-1) for _ in range(2)]
    worlds[1].set_host_threads(4)
    st = worlds[0].states()
    rng = np.random.default_rng(3)
    for step in range(8):
        st["position"] += rng.normal(0.0, 0.05, st["position"].shape).astype(np.float32)
        for w in worlds:
            w.set_states(st)
            w.update_broadphase()
        a, b = worlds[0].pairs(), worlds[1].pairs()
        assert len(a) > 1000 and np.array_equal(a, b), f"update {step}: pair lists differ"
