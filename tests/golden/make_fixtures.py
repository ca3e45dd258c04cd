"""Write the golden fixtures under tests/golden/ from the REFERENCE engine.

Run in a container where the reference checkout exists (it is compiled from
its own sources by ``make -C oracle ref``):

    python tests/golden/make_fixtures.py          # every fixture
    python tests/golden/make_fixtures.py planes   # planes2048.npz only

The fixtures are data only: seeded scene ids, and the reference's outputs on
them (broadphase pairs per step, body states, raw manifolds, contact
constraints with impulses, built convex hulls, mass data).  The scenes
themselves are regenerated from kizuriphysics3d_amd.scenes with the stored
seeds.
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from kizuriphysics3d_amd import scenes as S  # noqa: E402
from oracle import refharness as R  # noqa: E402

CHECKPOINTS = (0, 1, 2, 9, 59, 149, 299)     # full body states
DETAIL_STEPS = (0, 1, 9, 59)                 # manifolds and constraints as well
PILE_SEED = 0x4B5A0202
PILE_N = 256
BATCH_SEED = 0x4B5A0303
BATCH_N = 4096
BATCH_SHAPES = 512
PLANE_SEED = 0x4B5A0107
PLANE_N = 2048
PLANE_SHAPES = 256


def state_digest(st: np.ndarray) -> np.ndarray:
    return np.frombuffer(hashlib.sha256(st.tobytes()).digest()[:8], np.uint64)[0]


def reference_hulls(shapes, hull):
    out = {}
    for i, s in enumerate(shapes):
        if s["type"] == S.SHAPE_HULL:
            out[i] = R.hull_create(hull[s["hull_offset"]: s["hull_offset"] + s["hull_count"]])
    return out


def pack_hulls(shapes, hulls):
    refs = np.zeros((len(shapes), 4), np.uint32)
    vs, ts, nv, nt = [], [], 0, 0
    for i, (v, t) in sorted(hulls.items()):
        refs[i] = (nv, len(v), nt, len(t))
        vs.append(v)
        ts.append(t)
        nv += len(v)
        nt += len(t)
    verts = np.concatenate(vs).astype(np.float32) if vs else np.zeros((0, 3), np.float32)
    tris = np.concatenate(ts).astype(np.uint32) if ts else np.zeros((0, 3), np.uint32)
    return refs, verts, tris


def trajectory(scene: S.Scene, steps: int, path: str):
    rw = R.RefWorld(scene)
    pairs, offs, digests = [], [0], []
    ck = {}
    for i in range(steps):
        # One whole StepFixed; the pair list, raw manifolds and constraints of
        # that step stay readable until the next step.
        rw.step_phases(stop_after=5)
        p = rw.pairs()
        pairs.append(p)
        offs.append(offs[-1] + len(p))
        st = rw.states()
        digests.append(state_digest(st))
        if i in CHECKPOINTS or i in DETAIL_STEPS:
            hdr, pts = rw.constraints()
            cnt, mpts, keys = rw.manifolds()
            ck[f"state_{i}"] = st
        if i in DETAIL_STEPS:
            ck[f"chdr_{i}"] = hdr
            live = np.arange(8)[None, :] < hdr[:, 2:3]
            ck[f"cpts_{i}"] = pts[live]  # live points only, constraint order
            ck[f"mcount_{i}"] = cnt
            ck[f"mpts_{i}"] = mpts[:, :4][np.arange(4)[None, :] < cnt[:, None]]  # live points only
            ck[f"mkeys_{i}"] = keys
    shapes = scene.shape_array()
    refs, verts, tris = pack_hulls(shapes, reference_hulls(shapes, scene.hull_array()))
    allp = np.concatenate(pairs)
    assert allp.max() < 65536
    np.savez_compressed(path, pairs=allp.astype(np.uint16), pair_offsets=np.asarray(offs, np.uint32),
                        digests=np.asarray(digests, np.uint64), mass=rw.mass(), hull_refs=refs,
                        hull_verts=verts, hull_tris=tris, checkpoints=np.asarray(CHECKPOINTS),
                        detail_steps=np.asarray(DETAIL_STEPS), **ck)
    print(path, os.path.getsize(path) // 1024, "KiB")


def plane_pairs():
    """CollidePlaneConvex over plane_pair_batch: every convex kind, both
    orders, hulls of more than 16 vertices, faces resting on the plane."""
    pb = S.plane_pair_batch(PLANE_N, seed=PLANE_SEED, num_shapes=PLANE_SHAPES)
    counts, pts, _ = R.collide_batch(pb)
    refs, verts, tris = pack_hulls(pb.shapes, reference_hulls(pb.shapes, pb.hull_points))
    path = os.path.join(HERE, "planes2048.npz")
    np.savez_compressed(path, counts=counts, points=pts[:, :4].copy(), hull_refs=refs, hull_verts=verts,
                        hull_tris=tris)
    print(path, os.path.getsize(path) // 1024, "KiB")


def main():
    R.init_jobs(0)
    if sys.argv[1:] == ["planes"]:
        plane_pairs()
        return
    trajectory(S.pyramid(), 300, os.path.join(HERE, "pyramid20.npz"))
    trajectory(S.pile(PILE_N, seed=PILE_SEED), 300, os.path.join(HERE, "pile256.npz"))
    pb = S.pair_batch(BATCH_N, seed=BATCH_SEED, num_shapes=BATCH_SHAPES)
    counts, pts, _ = R.collide_batch(pb)
    refs, verts, tris = pack_hulls(pb.shapes, reference_hulls(pb.shapes, pb.hull_points))
    np.savez_compressed(os.path.join(HERE, "narrowphase4096.npz"), counts=counts, points=pts[:, :4].copy(),
                        hull_refs=refs, hull_verts=verts, hull_tris=tris)
    # Standalone hull builds and mass data for a spread of inputs.
    rng = S.SplitMix64(0x4B5A0404)
    sc = S.Scene("shapes", S.world_settings())
    for i in range(64):
        k = i % 4
        b = S.random_convex_body(sc, rng, k)
        b["position"] = (3.0 * i, 5.0, 0.0)
        sc.add_body(b)
    rw = R.RefWorld(sc)
    shapes = sc.shape_array()
    refs, verts, tris = pack_hulls(shapes, reference_hulls(shapes, sc.hull_array()))
    np.savez_compressed(os.path.join(HERE, "shapes64.npz"), mass=rw.mass(), hull_refs=refs, hull_verts=verts,
                        hull_tris=tris)
    plane_pairs()
    print("done")


if __name__ == "__main__":
    main()
