"""The solve schedule (islands, first-fit SIMD batches, levels) against a
brute-force restatement of the reference's loops on random contact graphs.

The device sweeps points level by level; it reproduces the reference exactly
iff, for every dynamic body, its points appear in the same order as in the
reference's sequence, and no level holds two points of one dynamic body."""
import numpy as np
import pytest

from kizuriphysics3d_amd import device as D


def reference_sequences(dyn, info, use_simd=True):
    """Brute force, as PhysicsWorld::BuildIslands + SolveIsland do it
    (src/World/PhysicsWorld.cpp:587-717): per island, the velocity sequence of
    first-fit batches (lane order), and the constraint-order sequence."""
    nb = len(dyn)
    idx = {}
    for i in range(nb):
        if dyn[i]:
            idx[i] = len(idx)
    parent = list(range(len(idx)))
    rank = [0] * len(idx)

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x
    for a, b, first, w in info:
        if dyn[a] and dyn[b]:
            ra, rb = find(idx[a]), find(idx[b])
            if ra != rb:
                if rank[ra] < rank[rb]:
                    ra, rb = rb, ra
                parent[rb] = ra
                if rank[ra] == rank[rb]:
                    rank[ra] += 1
    root_island = {}
    for i in range(len(idx)):
        r = find(i)
        root_island.setdefault(r, len(root_island))
    islands = [[] for _ in range(len(root_island))]
    for ci, (a, b, first, w) in enumerate(info):
        body = a if dyn[a] else b
        islands[root_island[find(idx[body])]].append(ci)
    vel, ws = [], []
    for isl in islands:
        for ci in isl:
            a, b, first, w = info[ci]
            ws += [first + k for k in range(w & 0xFF)]
        if not use_simd:
            continue
        batches = []
        for ci in isl:
            a, b, first, w = info[ci]
            if (w >> 8) & 1:
                continue
            for k in range(w & 0xFF):
                placed = False
                for bt in batches:
                    if len(bt) >= 4:
                        continue
                    if any(x in (a, b) or y in (a, b) for (x, y, _) in bt):
                        continue
                    bt.append((a, b, first + k))
                    placed = True
                    break
                if not placed:
                    batches.append([(a, b, first + k)])
        for bt in batches:
            vel += [p for (_, _, p) in bt]
    if not use_simd:
        vel = list(ws)
    return vel, ws


def random_graph(rng, nb, nc, nstatic):
    dyn = np.ones(nb, np.uint8)
    dyn[:nstatic] = 0
    info = []
    first = 0
    for _ in range(nc):
        while True:
            a, b = rng.integers(0, nb, 2)
            if a != b and (dyn[a] or dyn[b]):
                break
        a, b = min(a, b), max(a, b)
        k = int(rng.integers(1, 5))
        sensor = 1 if rng.random() < 0.03 else 0
        info.append((a, b, first, k | (sensor << 8)))
        first += k
    return dyn, np.array(info, np.uint32)


def body_of(info):
    pb = []
    for a, b, first, w in info:
        pb += [(a, b)] * (w & 0xFF)
    return pb


@pytest.mark.parametrize("seed,nb,nc,nstatic,groups", [(1, 12, 30, 1, 1024), (2, 40, 150, 2, 3), (3, 200, 900, 3, 8),
                                                      (4, 30, 120, 0, 1), (5, 400, 1500, 5, 64)])
@pytest.mark.parametrize("use_simd", [True, False])
def test_schedule_preserves_every_dynamic_body_order(seed, nb, nc, nstatic, groups, use_simd):
    rng = np.random.default_rng(seed)
    dyn, info = random_graph(rng, nb, nc, nstatic)
    vl, wl, gr, st = D.schedule(dyn, info, use_simd=use_simd, max_groups=groups)
    vel, ws = reference_sequences(dyn, info, use_simd)
    pb = body_of(info)
    for seq, lvl in ((vel, vl), (ws, wl)):
        assert sorted(seq) == sorted(np.nonzero(lvl)[0].tolist())
        for body in np.nonzero(dyn)[0]:
            ref = [p for p in seq if body in pb[p]]
            got = sorted(ref, key=lambda p: lvl[p])
            assert got == ref, f"body {body}: order differs"
            assert len({lvl[p] for p in ref}) == len(ref), f"body {body}: two points share a level"
        # points of one group and level touch each dynamic body once
        seen = set()
        for p in seq:
            for body in set(pb[p]):
                if dyn[body]:
                    key = (gr[p], lvl[p], body)
                    assert key not in seen
                    seen.add(key)
    assert st[1] == min(groups, st[0]) or st[0] == 0


def test_schedule_levels_are_as_deep_as_the_reference_chain():
    # A chain of bodies: each constraint shares a body with the next, so the
    # sequence is fully serial.
    nb = 10
    dyn = np.ones(nb, np.uint8)
    info = np.array([(i, i + 1, i, 1) for i in range(nb - 1)], np.uint32)
    vl, wl, gr, st = D.schedule(dyn, info)
    assert list(wl) == list(range(1, nb))
    assert st[0] == 1 and st[3] == nb - 1
