"""Snapshots (kzw_save / kzw_load): the reference's KZP1 world record plus
an extension that makes a resumed world bit-identical.

CPU tests use host-only worlds (shapes, bodies, mass and broadphase, no
device) and the reference built from its own sources; GPU tests resume a
stepping world and compare every later step bit for bit."""
import numpy as np
import pytest

from kizuriphysics3d_amd import scenes as S
from kizuriphysics3d_amd.world import PhysicsWorld
from oracle import refharness as R
from tests.helpers import first_mismatch


def _drift(states, k):
    """A deterministic state change, as if the bodies had moved."""
    st = states.copy()
    st["position"][:, 0] += np.float32(0.013 * k)
    st["position"][:, 1] -= np.float32(0.021 * k)
    return st


def test_snapshot_round_trip_host_world():
    """save -> load gives the same bodies, mass, hulls and broadphase: the
    pair lists of later broadphase updates are identical."""
    scene = S.pile(256, seed=0x4B5A0601, width=12.0, y_range=(1.0, 8.0))
    a = PhysicsWorld(scene, device=-1)
    for k in range(3):
        a.set_states(_drift(a.states(), 1))
        a.update_broadphase()
    b = PhysicsWorld.load(a.save(), device=-1)
    assert first_mismatch(a.states(), b.states()) is None
    assert np.array_equal(a.mass(), b.mass())
    shapes = scene.shape_array()
    for i, (v, t) in a.hulls(shapes).items():
        vb, tb = b.hull(i)
        assert np.array_equal(v, vb) and np.array_equal(t, tb)
    for k in range(4):
        st = _drift(a.states(), 1)
        a.set_states(st)
        b.set_states(st)
        a.update_broadphase()
        b.update_broadphase()
        assert np.array_equal(a.pairs(), b.pairs()), k
        assert len(a.pairs()) > 0


def test_snapshot_rejects_malformed_input():
    scene = S.pile(32, seed=0x4B5A0602, width=6.0, y_range=(1.0, 4.0))
    data = PhysicsWorld(scene, device=-1).save()
    with pytest.raises(Exception):
        PhysicsWorld.load(data[:20], device=-1)
    bad = bytearray(data)
    bad[0] ^= 0xFF  # magic
    with pytest.raises(Exception):
        PhysicsWorld.load(bytes(bad), device=-1)


def _blocks(data: bytes):
    """Offsets of the extension blocks: {tag: (payload offset, length)}."""
    kzx = data.find(b"KZX1")
    assert kzx > 0
    off, out = kzx + 8, {}
    while True:
        tag = data[off:off + 4].decode()
        n = int(np.frombuffer(data[off + 4:off + 12], np.uint64)[0])
        out[tag] = (off + 12, n)
        if tag == "END ":
            return out
        off += 12 + n


def _must_fail(data):
    with pytest.raises(Exception):
        PhysicsWorld.load(bytes(data), device=-1)


def test_snapshot_rejects_truncated_and_corrupted_records():
    """kzw_load checks every length and index it reads: truncated records,
    counts and lengths beyond the data, a shape the device path lacks, a
    hull of fewer than 4 vertices and broadphase proxies that do not name
    the body's leaf are errors, never reads out of bounds."""
    scene = S.pile(48, seed=0x4B5A0603, width=6.0, y_range=(1.0, 4.0))
    data = PhysicsWorld(scene, device=-1).save()
    PhysicsWorld.load(data, device=-1).close()  # the intact record loads
    for cut in list(range(0, len(data) - 16, max(len(data) // 97, 1))):
        _must_fail(data[:cut])
    blk = _blocks(data)
    # a body count far beyond the record
    bad = bytearray(data)
    bad[8:12] = np.uint32(0x00FFFFFF).tobytes()
    _must_fail(bad)
    # a block length that wraps the bounds check
    o, _ = blk["SHAP"]
    bad = bytearray(data)
    bad[o - 8:o] = np.uint64(0xFFFFFFFFFFFFFFF0).tobytes()
    _must_fail(bad)
    # SHAP: first record's type (kz_shape_desc starts with the type) -> cylinder-like unknown code
    bad = bytearray(data)
    bad[o + 4:o + 8] = np.uint32(7).tobytes()
    _must_fail(bad)
    # BPHT: proxy of body 0 pointing at another node
    o, n = blk["BPHT"]
    words = int(np.frombuffer(data[o:o + 8], np.uint64)[0])
    proxies = o + 8 + 4 * words
    bad = bytearray(data)
    p0 = int(np.frombuffer(data[proxies:proxies + 4], np.uint32)[0])
    bad[proxies:proxies + 4] = np.uint32(p0 ^ 1).tobytes()
    _must_fail(bad)
    bad[proxies:proxies + 4] = np.uint32(0xFFFFFFF0).tobytes()
    _must_fail(bad)
    # BPHT: a child link of the root pointing back at the root (a cycle)
    root = int(np.frombuffer(data[o + 8 + 4:o + 8 + 8], np.uint32)[0])
    node = o + 8 + 16 + 18 * 4 * root
    bad = bytearray(data)
    bad[node + 13 * 4:node + 14 * 4] = np.uint32(root).tobytes()
    _must_fail(bad)


def test_snapshot_rejects_degenerate_hull_record():
    """A SHAP record that calls a shape a hull but carries no hull vertices
    (ConvexHullShape::Create would have returned null) is rejected, as is a
    hull record claiming 3 vertices."""
    scene = S.pile(16, seed=0x4B5A0604, width=6.0, y_range=(1.0, 4.0))
    data = PhysicsWorld(scene, device=-1).save()
    o, n = _blocks(data)["SHAP"]
    shapes = scene.shape_array()
    assert shapes["type"][0] != S.SHAPE_HULL
    desc = shapes.dtype.itemsize
    bad = bytearray(data)
    bad[o + 4:o + 8] = np.uint32(S.SHAPE_HULL).tobytes()  # shape 0 retyped, its record holds 0 vertices
    _must_fail(bad)
    off = o + 4
    for i in range(len(shapes)):
        nv, nt = (int(x) for x in np.frombuffer(data[off + desc:off + desc + 8], np.uint32))
        if shapes["type"][i] == S.SHAPE_HULL:
            bad = bytearray(data)
            bad[off + desc:off + desc + 4] = np.uint32(3).tobytes()
            _must_fail(bad)
            return
        off += desc + 8 + 12 * nv + 4 * nt + 24 + 4 + 24
    pytest.fail("no hull shape in the scene")


@pytest.mark.reference
def test_kzp1_records_interoperate_with_the_reference():
    """The reference's SerializeWorld record loads here (hulls rebuilt as
    ConvexHullShape::Create builds them, broadphase rebuilt by CreateBody
    order) and gives the reference's own pairs after loading its record;
    the KZP1 part of this library's snapshot loads into the reference."""
    scene = S.pile(256, seed=0x4B5A0603, width=12.0, y_range=(1.0, 8.0))
    R.init_jobs(0)
    rw = R.RefWorld(scene)
    for _ in range(30):
        rw.step()
    rec = rw.save_kzp1()
    ours = PhysicsWorld.load(rec, device=-1, world=scene.world)
    ref2 = R.RefWorld.load_kzp1(rec, scene.world)
    assert first_mismatch(ref2.states(), ours.states()) is None
    assert np.array_equal(ref2.mass()[:, :10], ours.mass())
    # the first broadphase after loading: same pairs in the same order
    ref2.step_phases(stop_after=0)
    ours.update_broadphase()
    assert np.array_equal(ref2.pairs(), ours.pairs())
    # and the other way round: the reference reads this library's snapshot
    back = R.RefWorld.load_kzp1(ours.save(), scene.world)
    assert first_mismatch(back.states(), ours.states()) is None


@pytest.mark.gpu
def test_resumed_world_is_bit_identical():
    """A world saved after 150 steps and loaded into a new world steps on
    exactly like the original: states after each of the next 60 steps."""
    scene = S.pile(512, seed=0x4B5A0102)
    a = PhysicsWorld(scene)
    for _ in range(150):
        a.step_fixed()
    b = PhysicsWorld.load(a.save())
    for i in range(60):
        a.step_fixed()
        b.step_fixed()
        j = first_mismatch(a.states(), b.states())
        assert j is None, f"step {150 + i + 1}: body {j} differs after resuming"


@pytest.mark.gpu
@pytest.mark.reference
def test_kzp1_resume_matches_the_reference():
    """The reference saves a KZP1 record after 100 steps; the reference and
    this library both load it (fresh broadphase, empty warm-start cache) and
    step on bit-identically."""
    scene = S.pile(512, seed=0x4B5A0102)
    R.init_jobs(0)
    rw = R.RefWorld(scene)
    for _ in range(100):
        rw.step()
    rec = rw.save_kzp1()
    ref2 = R.RefWorld.load_kzp1(rec, scene.world)
    ours = PhysicsWorld.load(rec, world=scene.world)
    for i in range(60):
        ref2.step()
        ours.step_fixed()
        j = first_mismatch(ref2.states(), ours.states())
        assert j is None, f"step {i}: body {j} differs"
