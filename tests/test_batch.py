"""Batched worlds (include/kzb.h): every world of a batch must step exactly
as it steps alone through the single-world path (itself bit-exact against
the reference), whatever the launch configuration, and the device-built
solve schedule must equal the host schedule of the single-world path."""
import numpy as np
import pytest

from kizuriphysics3d_amd import device as D
from kizuriphysics3d_amd import scenes as S


def test_batched_world_generator_is_deterministic_and_in_range():
    a = S.batched_worlds(40, seed=0x4B5A0404, min_bodies=16, max_bodies=512)
    b = S.batched_worlds(40, seed=0x4B5A0404, min_bodies=16, max_bodies=512)
    assert [w.name for w in a] == [w.name for w in b]
    for x, y in zip(a, b):
        assert x.body_array().tobytes() == y.body_array().tobytes()
        assert x.shape_array().tobytes() == y.shape_array().tobytes()
    kinds = {w.name.rstrip("0123456789") for w in a}
    assert kinds == {"stack", "pile"}
    for w in a:
        dyn = int((w.body_array()["motion_type"] == S.MOTION_DYNAMIC).sum())
        assert 16 <= dyn <= 512
        assert w.shape_array()["type"][w.body_array()["shape"][0]] == S.SHAPE_PLANE  # own ground plane first


def test_kzb_exports_every_declared_function():
    import os
    import re
    from kizuriphysics3d_amd import batch as Bt
    text = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "kzb.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = sorted(set(re.findall(r"\b(kzb_[a-z0-9_]+)\s*\(", text)))
    lib = D.library()
    assert set(names) == set(Bt.KZB_EXPORTS)
    assert all(hasattr(lib, n) for n in names)


def _worlds(n=24, seed=0x4B5A0404):
    return S.batched_worlds(n, seed=seed, min_bodies=16, max_bodies=160)


@pytest.mark.gpu
def test_batched_worlds_match_single_worlds():
    from kizuriphysics3d_amd.batch import BatchedWorlds
    from kizuriphysics3d_amd.world import PhysicsWorld
    from tests.helpers import first_mismatch
    scenes = _worlds()
    batch = BatchedWorlds(scenes)
    singles = [PhysicsWorld(sc) for sc in scenes]
    steps = 90
    for i in range(steps):
        batch.step()
        for w in singles:
            w.step_fixed()
        if i % 30 == 29 or i == steps - 1:
            for k, w in enumerate(singles):
                j = first_mismatch(w.states(), batch.states(k))
                assert j is None, f"step {i}: world {k} ({scenes[k].name}) body {j} differs"
    st = batch.stats()
    print(f"{len(scenes)} worlds, {st['num_bodies']} bodies, {steps} steps bit-identical to single-world runs; "
          f"last step {st['num_points']} points, device {st['device_ms']:.2f} ms", flush=True)
    assert st["num_points"] > 1000


@pytest.mark.gpu
@pytest.mark.reference
def test_batched_worlds_match_the_reference():
    """Worlds of a batch against the reference stepping each of them alone,
    bit for bit: box stacks and convex piles of 16-160 bodies, each on its
    own ground plane (so the plane-convex path of CollideShape.cpp:370-430
    carries every ground contact)."""
    from oracle import refharness as R
    from kizuriphysics3d_amd.batch import BatchedWorlds
    from tests.helpers import first_mismatch
    R.init_jobs(0)
    scenes = _worlds(16, seed=0x4B5A0407)
    batch = BatchedWorlds(scenes)
    refs = [R.RefWorld(sc) for sc in scenes]
    kinds = {sc.name.rstrip("0123456789") for sc in scenes}
    assert kinds == {"stack", "pile"}
    steps = 150
    for i in range(steps):
        batch.step()
        for rw in refs:
            rw.step()
        if i % 10 == 9 or i == steps - 1:
            for k, rw in enumerate(refs):
                j = first_mismatch(rw.states(), batch.states(k))
                assert j is None, f"step {i}: world {k} ({scenes[k].name}) body {j} differs from the reference"
    st = batch.stats()
    plane_points = 0
    for rw, sc in zip(refs, scenes):
        hdr, _ = rw.constraints()
        plane_points += int(hdr[(hdr[:, 0] == 0) | (hdr[:, 1] == 0), 2].sum())  # body 0: the ground plane
    print(f"{len(scenes)} worlds ({sorted(kinds)}), {st['num_bodies']} bodies, {steps} steps bit-identical to the "
          f"reference; last step {st['num_points']} points, {plane_points} on the ground planes", flush=True)
    assert plane_points > 100


@pytest.mark.gpu
def test_batched_schedule_matches_host_schedule():
    from kizuriphysics3d_amd.batch import BatchedWorlds
    from kizuriphysics3d_amd.world import PhysicsWorld
    scenes = _worlds(12, seed=0x4B5A0405)
    batch = BatchedWorlds(scenes)
    singles = [PhysicsWorld(sc) for sc in scenes]
    for _ in range(60):
        batch.step()
        for w in singles:
            w.step_fixed()
    checked = 0
    for k, (sc, w) in enumerate(zip(scenes, singles)):
        ctx = w.device_context()
        hdr, _ = ctx.constraints()
        npts = hdr[:, 2].astype(np.int64)
        info = np.zeros((len(hdr), 4), np.uint32)
        info[:, 0], info[:, 1] = hdr[:, 0], hdr[:, 1]
        info[:, 2] = np.concatenate([[0], np.cumsum(npts)[:-1]]) if len(hdr) else 0
        info[:, 3] = hdr[:, 2] | ((hdr[:, 3] & 1) << 8)
        dyn = (sc.body_array()["motion_type"] == S.MOTION_DYNAMIC).astype(np.uint8)
        vl, wl, _, _ = D.schedule(dyn, info, use_simd=True)
        bvl, bwl = batch.world_schedule(k)
        assert np.array_equal(bwl, wl), f"world {k}: warm-start levels differ"
        assert np.array_equal(bvl, vl), f"world {k}: velocity levels differ"
        checked += len(vl)
    assert checked > 500


@pytest.mark.gpu
def test_launch_configuration_does_not_change_results():
    from kizuriphysics3d_amd.batch import DIRECT_LAUNCH, BatchedWorlds
    from tests.helpers import first_mismatch
    scenes = _worlds(16, seed=0x4B5A0406)
    a, b = BatchedWorlds(scenes), BatchedWorlds(scenes)
    a.set_config(narrowphase_waves=256, solve_threads=64)
    b.set_config(narrowphase_waves=8192, solve_threads=512, flags=DIRECT_LAUNCH)
    for _ in range(40):
        a.step()
        b.step()
    for k in range(len(scenes)):
        assert first_mismatch(a.states(k), b.states(k)) is None, k
    # direct launches time the phases with events; the graph does not
    ph = np.array(b.stats()["phase_ms"])
    assert np.all(ph > 0) and np.all(np.array(a.stats()["phase_ms"]) == 0), ph
    b.set_config()
    # the autotuner steps the batch with every candidate; results stay identical
    n = len(b.candidates())
    bi, times = b.autotune(repeats=1)
    for _ in range(n):
        a.step()
    assert 0 <= bi < n and np.all(times > 0)
    for k in range(len(scenes)):
        assert first_mismatch(a.states(k), b.states(k)) is None, k
    print(f"autotune: best {b.config()} of {n}, medians {np.round(times, 3)}", flush=True)
