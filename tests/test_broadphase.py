"""The broadphase pair query on the device (kzg_broadphase_pairs) against the
reference's DynamicAABBTree::QueryPairs / BroadPhase::UpdatePairs
(include/Kizuri/BroadPhase/BroadPhase.h:61-104, src/BroadPhase/
BroadPhase.cpp:368-381): the same pairs in the same order at every step, the
pairs the reference's query loses (met as n1 > n2) included."""
import ctypes as C

import numpy as np
import pytest

from kizuriphysics3d_amd import device as D
from kizuriphysics3d_amd import scenes as S
from tests.helpers import first_mismatch

pytestmark = pytest.mark.gpu


@pytest.mark.reference
@pytest.mark.parametrize("scene_fn,steps", [(lambda: S.pyramid(), 200), (lambda: S.pile(2048, seed=77), 300)])
def test_device_broadphase_lockstep_with_reference(scene_fn, steps):
    """A world whose pair query runs on the device, stepped beside the
    reference: its pair list and its body states equal the reference's at
    every step."""
    from oracle import refharness as R
    from kizuriphysics3d_amd.world import PhysicsWorld
    sc = scene_fn()
    R.init_jobs(0)
    rw = R.RefWorld(sc)
    w = PhysicsWorld(sc)
    total = 0
    for i in range(steps):
        rw.step()
        w.step_fixed()
        rp, gp = rw.pairs(), w.pairs()
        assert len(rp) == len(gp) and np.array_equal(rp, gp), f"step {i}: pair lists differ ({len(rp)} vs {len(gp)})"
        j = first_mismatch(rw.states(), w.states())
        assert j is None, f"step {i}: body {j} differs"
        total += len(rp)
    print(f"{sc.name}: {steps} steps, {total} pairs in the reference's order, states bit-exact", flush=True)


def test_device_broadphase_equals_host_walk_large_pile():
    """16,384 bodies (the host walks its tree on threads): two worlds, one
    querying pairs on the device and one on the host, step the same pile;
    pairs and states are identical at every step."""
    from kizuriphysics3d_amd.world import PhysicsWorld
    scene = S.pile(16384, seed=0x4B5A0701, width=40.0, y_range=(1.0, 20.0))
    dev, host = PhysicsWorld(scene), PhysicsWorld(scene)
    host.set_device_broadphase(False)
    for i in range(120):
        dev.step_fixed()
        host.step_fixed()
        if i % 10 == 9:
            a, b = dev.pairs(), host.pairs()
            assert np.array_equal(a, b), f"step {i}: pair lists differ ({len(a)} vs {len(b)})"
            assert first_mismatch(dev.states(), host.states()) is None, f"step {i}: states differ"
    st = dev.stats()
    print(f"16384 bodies: {st['num_broadphase_pairs']} pairs at the last step, broadphase {st['broadphase_ms']:.2f} ms "
          f"on the device path, {host.stats()['broadphase_ms']:.2f} ms on host threads", flush=True)
    assert st["num_broadphase_pairs"] > 20000


def test_device_broadphase_rejects_a_bad_child_id():
    """A node whose child id is outside the image is an argument error (the
    kernels check every id they read)."""
    ctx = D.DeviceContext(S.world_settings())
    sc = S.pile(8, seed=0x4B5A0702)
    shapes = sc.shape_array()
    from kizuriphysics3d_amd.world import build_hulls
    ctx.upload_shapes(shapes, build_hulls(shapes, sc.hull_array()))
    # three nodes: a root over two leaves, the second child id out of range
    node = np.zeros((3, 16), np.float32)
    u = node.view(np.uint32)
    for k in range(3):
        node[k, 0:3], node[k, 4:7] = -1.0, 1.0
        node[k, 8:11], node[k, 12:15] = -1.0, 1.0
    u[0, 3], u[0, 7], u[0, 15] = 1, 7, 1          # root: children 1 and 7 (bad), height 1
    for k in (1, 2):
        u[k, 3] = u[k, 7] = 0xFFFFFFFF            # leaves
        u[k, 11], u[k, 15] = 0, 0
    lay = np.ones(3, np.uint32)
    msk = np.full(3, 0xFFFFFFFF, np.uint32)
    n = C.c_uint32()
    rc = D.library().kzg_broadphase_pairs(ctx.h, node.ctypes.data_as(C.c_void_p), 3, 0,
                                          lay.ctypes.data_as(C.c_void_p), msk.ctypes.data_as(C.c_void_p), 3,
                                          C.byref(n))
    assert rc == -1, rc
