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


@pytest.mark.parametrize("seed,nb,nc,nstatic,groups", [(11, 12, 30, 1, 1024), (12, 40, 150, 2, 3),
                                                      (13, 200, 900, 3, 8), (14, 30, 120, 0, 1),
                                                      (15, 400, 1500, 5, 64), (16, 6, 40, 1, 2)])
@pytest.mark.parametrize("use_simd", [True, False])
def test_units_preserve_every_dynamic_body_order(seed, nb, nc, nstatic, groups, use_simd):
    """Multi-point units: one thread applies a unit's points in point order,
    so the sweep equals the reference sequence iff sorting a body's points by
    (level, point) gives its reference order, a unit holds one body pair at
    one level, and the units of a level share no dynamic body."""
    rng = np.random.default_rng(seed)
    dyn, info = random_graph(rng, nb, nc, nstatic)
    vl, wl, gr, st, vu, wu = D.schedule(dyn, info, use_simd=use_simd, max_groups=groups, units=True,
                                        with_units=True)
    vl1, wl1, _, st1 = D.schedule(dyn, info, use_simd=use_simd, max_groups=groups)
    vel, ws = reference_sequences(dyn, info, use_simd)
    pb = body_of(info)
    for seq, lvl, unit, lvl1, nunits in ((vel, vl, vu, vl1, st[4]), (ws, wl, wu, wl1, st[5])):
        assert sorted(seq) == sorted(np.nonzero(lvl)[0].tolist())
        assert np.array_equal(lvl > 0, lvl1 > 0)
        members = {}
        for p in seq:
            members.setdefault(int(unit[p]), []).append(p)
        assert len(members) == nunits
        for u, pts in members.items():
            assert 1 <= len(pts) <= 4
            assert len({pb[p] for p in pts}) == 1, f"unit {u} mixes body pairs"
            assert len({(gr[p], lvl[p]) for p in pts}) == 1, f"unit {u} spans levels"
        for body in np.nonzero(dyn)[0]:
            ref = [p for p in seq if body in pb[p]]
            got = sorted(ref, key=lambda p: (lvl[p], p))
            assert got == ref, f"body {body}: order differs"
        seen = set()
        for u, pts in members.items():
            for body in set(pb[pts[0]]):
                if dyn[body]:
                    key = (gr[pts[0]], lvl[pts[0]], body)
                    assert key not in seen, f"two units of one level share body {body}"
                    seen.add(key)
        # joining never deepens a group
        for g in np.unique(gr[lvl > 0]):
            sel = (gr == g) & (lvl > 0)
            assert lvl[sel].max() <= lvl1[sel].max()


def test_units_join_a_manifold_chain():
    # Three bodies in a row, two four-point manifolds: a chain of 8 point
    # levels becomes 4 unit levels (units of up to 2 points) in both orders.
    dyn = np.ones(3, np.uint8)
    info = np.array([(0, 1, 0, 4), (1, 2, 4, 4)], np.uint32)
    vl, wl, gr, st = D.schedule(dyn, info, units=True)
    assert list(wl) == [1, 1, 2, 2, 3, 3, 4, 4]
    assert list(vl) == [1, 1, 2, 2, 3, 3, 4, 4]
    assert st[4] == 4 and st[5] == 4 and st[2] == 4 and st[3] == 4
    vl1, wl1, _, st1 = D.schedule(dyn, info)
    assert st1[3] == 8 and st1[4] == 8
    # a point of another pair between two points of one manifold breaks the unit
    dyn = np.array([0, 1, 1], np.uint8)
    info = np.array([(0, 1, 0, 2), (1, 2, 2, 1), (0, 1, 3, 2)], np.uint32)
    _, wl, _, st, _, wu = D.schedule(dyn, info, use_simd=False, units=True, with_units=True)
    assert list(wl) == [1, 1, 2, 3, 3]
    assert st[5] == 3
    # a manifold of 4 points is cut into units of 2
    dyn = np.array([0, 1], np.uint8)
    info = np.array([(0, 1, 0, 4)], np.uint32)
    _, wl, _, st, _, wu = D.schedule(dyn, info, units=True, with_units=True)
    assert list(wl) == [1, 1, 2, 2] and wu[0] == wu[1] != wu[2] == wu[3]


@pytest.mark.parametrize("use_simd", [True, False])
@pytest.mark.parametrize("groups", [1, 64])
@pytest.mark.parametrize("nb,nc", [(3000, 12000), (40, 150)])
def test_deferred_build_equals_plain(use_simd, groups, nb, nc):
    """The early-warm-start build (warm-start order first, velocity order
    finished on the helper thread) gives the same schedule as the plain one,
    above the size where the plain build uses two threads (~30k points) and
    below it."""
    rng = np.random.default_rng(21)
    dyn, info = random_graph(rng, nb, nc, 20 if nb > 100 else 2)
    plain = D.schedule(dyn, info, use_simd=use_simd, max_groups=groups, units=True, with_units=True)
    deferred = D.schedule(dyn, info, use_simd=use_simd, max_groups=groups, units=True, with_units=True, defer=True)
    for a, b in zip(plain, deferred):
        assert np.array_equal(a, b)
