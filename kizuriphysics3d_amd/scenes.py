"""Seeded scene generators for the contact-path workloads.

A scene is plain data in the layout of ``include/kz_scene.h``: a shape table,
a float3 pool of convex-hull input points and a body table, plus the world
settings.  The same scene drives the product (``kizuriphysics3d_amd.world``),
the CPU oracle and the reference harness, so every implementation steps
identical inputs.

Random numbers come from splitmix64 (Steele, Lea, Flood 2014); a float in
[0, 1) is ``(u64 >> 40) * 2**-24``; random rotations are Shoemake's uniform
quaternions.  Workloads:

* ``pyramid``      20-level box pyramid (210 unit cubes, 1 kg, friction 0.5)
                   on a static plane y=0, 10 velocity + 2 position iterations.
* ``pile``         random convex pile: 25 % spheres, boxes, capsules and
                   convex hulls of 16-32 points on ellipsoids, dropped from
                   y in [2, 60] into a 64 x 64 m bin (static box floor and
                   four static box walls), density 1000, 8 + 2 iterations.
* ``pair_batch``   synthetic convex-convex pairs (hull-hull 8-64 vertices,
                   box-hull, capsule-box) with penetration ~U[0, 0.05] m and
                   10 % separated by ~U[0, 0.02] m along a random axis.
* ``world_batch``  many independent small worlds (box stacks or piles).

The reference's own creation defaults are those of ``BodySettings``
(include/Kizuri/Body/Body.h:62-90) and ``WorldSettings`` / ``SolverSettings``
(include/Kizuri/World/PhysicsWorld.h:20-43, include/Kizuri/Solver/
ContactSolver.h:67-82).  Sleeping is disabled in every workload.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

SHAPE_SPHERE, SHAPE_BOX, SHAPE_CAPSULE, SHAPE_HULL, SHAPE_PLANE = 0, 1, 2, 4, 8
MOTION_STATIC, MOTION_KINEMATIC, MOTION_DYNAMIC = 0, 1, 2

SHAPE_DTYPE = np.dtype([
    ("type", "<u4"), ("params", "<f4", (4,)), ("hull_offset", "<u4"), ("hull_count", "<u4"),
])
BODY_DTYPE = np.dtype([
    ("shape", "<u4"), ("motion_type", "<u4"), ("is_sensor", "<u4"), ("allow_sleeping", "<u4"),
    ("position", "<f4", (3,)), ("rotation", "<f4", (4,)),
    ("linear_velocity", "<f4", (3,)), ("angular_velocity", "<f4", (3,)),
    ("friction", "<f4"), ("restitution", "<f4"), ("linear_damping", "<f4"),
    ("angular_damping", "<f4"), ("gravity_factor", "<f4"), ("max_linear_velocity", "<f4"),
    ("max_angular_velocity", "<f4"), ("mass_override", "<f4"), ("density", "<f4"),
])
WORLD_DTYPE = np.dtype([
    ("gravity", "<f4", (3,)), ("velocity_iterations", "<u4"), ("position_iterations", "<u4"),
    ("baumgarte", "<f4"), ("penetration_slop", "<f4"), ("restitution_threshold", "<f4"),
    ("max_bias_velocity", "<f4"), ("warm_start_factor", "<f4"), ("use_simd_solver", "<u4"),
    ("use_velocity_bias", "<u4"), ("allow_sleeping", "<u4"), ("use_multithreading", "<u4"),
    ("fixed_timestep", "<f4"),
])
STATE_DTYPE = np.dtype([
    ("position", "<f4", (3,)), ("rotation", "<f4", (4,)),
    ("linear_velocity", "<f4", (3,)), ("angular_velocity", "<f4", (3,)),
])
CONTACT_DTYPE = np.dtype([
    ("point_on_a", "<f4", (3,)), ("point_on_b", "<f4", (3,)), ("normal", "<f4", (3,)),
    ("separation", "<f4"), ("feature_id", "<u4"),
])
assert SHAPE_DTYPE.itemsize == 28 and BODY_DTYPE.itemsize == 104
assert WORLD_DTYPE.itemsize == 60 and STATE_DTYPE.itemsize == 52 and CONTACT_DTYPE.itemsize == 44

DT = np.float32(1.0 / 60.0)


class SplitMix64:
    """splitmix64 stream; ``uniform`` gives float32-exact values in [0, 1)."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.state = seed & self.MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        u = (self.next_u64() >> 40) * (1.0 / (1 << 24))
        return lo + (hi - lo) * u

    def integer(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + self.next_u64() % (hi - lo + 1)

    def quaternion(self) -> tuple[float, float, float, float]:
        """Shoemake's uniform random rotation (x, y, z, w)."""
        u1, u2, u3 = self.uniform(), self.uniform(), self.uniform()
        a, b = math.sqrt(1.0 - u1), math.sqrt(u1)
        t2, t3 = 2.0 * math.pi * u2, 2.0 * math.pi * u3
        return (a * math.sin(t2), a * math.cos(t2), b * math.sin(t3), b * math.cos(t3))

    def unit_vector(self) -> tuple[float, float, float]:
        while True:
            x, y, z = self.uniform(-1, 1), self.uniform(-1, 1), self.uniform(-1, 1)
            r2 = x * x + y * y + z * z
            if 1e-6 < r2 <= 1.0:
                r = math.sqrt(r2)
                return (x / r, y / r, z / r)


def world_settings(velocity_iterations=8, position_iterations=2, **over) -> np.ndarray:
    w = np.zeros((), WORLD_DTYPE)
    w["gravity"] = (0.0, -9.81, 0.0)
    w["velocity_iterations"] = velocity_iterations
    w["position_iterations"] = position_iterations
    w["baumgarte"] = 0.2
    w["penetration_slop"] = 0.005
    w["restitution_threshold"] = 1.0
    w["max_bias_velocity"] = 3.0
    w["warm_start_factor"] = 1.0
    w["use_simd_solver"] = 1
    w["use_velocity_bias"] = 0
    w["allow_sleeping"] = 0
    w["use_multithreading"] = 1
    w["fixed_timestep"] = DT
    for k, v in over.items():
        w[k] = v
    return w


def default_body(shape: int, motion: int = MOTION_DYNAMIC) -> np.ndarray:
    b = np.zeros((), BODY_DTYPE)
    b["shape"] = shape
    b["motion_type"] = motion
    b["allow_sleeping"] = 1
    b["rotation"] = (0, 0, 0, 1)
    b["friction"] = 0.5
    b["restitution"] = 0.0
    b["linear_damping"] = 0.05
    b["angular_damping"] = 0.05
    b["gravity_factor"] = 1.0
    b["max_linear_velocity"] = 500.0
    b["max_angular_velocity"] = 50.0
    b["mass_override"] = 0.0
    b["density"] = 1000.0
    return b


@dataclass
class Scene:
    """Plain-data scene: world settings, shapes, hull point pool, bodies."""

    name: str
    world: np.ndarray
    shapes: list = field(default_factory=list)
    hull_points: list = field(default_factory=list)
    bodies: list = field(default_factory=list)

    def add_shape(self, kind: int, params=(0, 0, 0, 0), hull=None) -> int:
        s = np.zeros((), SHAPE_DTYPE)
        s["type"] = kind
        p = list(params) + [0.0] * (4 - len(params))
        s["params"] = p
        if hull is not None:
            s["hull_offset"] = len(self.hull_points)
            s["hull_count"] = len(hull)
            self.hull_points.extend(hull)
        self.shapes.append(s)
        return len(self.shapes) - 1

    def add_body(self, body: np.ndarray) -> int:
        self.bodies.append(body)
        return len(self.bodies) - 1

    # Frozen arrays in the C layout.
    def shape_array(self) -> np.ndarray:
        return np.array(self.shapes, dtype=SHAPE_DTYPE)

    def hull_array(self) -> np.ndarray:
        if not self.hull_points:
            return np.zeros((1, 3), np.float32)
        return np.asarray(self.hull_points, dtype=np.float32).reshape(-1, 3)

    def body_array(self) -> np.ndarray:
        return np.array(self.bodies, dtype=BODY_DTYPE)


def pyramid(levels: int = 20, half: float = 0.5, mass: float = 1.0) -> Scene:
    """Box pyramid: row i (from the ground) holds ``levels - i`` boxes, each row
    shifted by half a box; boxes start exactly touching.  Static plane first,
    so it has the lowest packed BodyID."""
    sc = Scene(f"pyramid{levels}", world_settings(10, 2))
    plane = sc.add_shape(SHAPE_PLANE, (0.0, 1.0, 0.0, 0.0))
    g = default_body(plane, MOTION_STATIC)
    sc.add_body(g)
    box = sc.add_shape(SHAPE_BOX, (half, half, half))
    side = 2.0 * half
    for row in range(levels):
        n = levels - row
        x0 = -0.5 * (n - 1) * side
        for i in range(n):
            b = default_body(box)
            b["position"] = (x0 + i * side, half + row * side, 0.0)
            b["mass_override"] = mass
            b["friction"] = 0.5
            b["restitution"] = 0.0
            sc.add_body(b)
    return sc


def ellipsoid_hull_points(rng: SplitMix64, n: int, axes) -> list:
    pts = []
    for _ in range(n):
        u = rng.unit_vector()
        pts.append((u[0] * axes[0], u[1] * axes[1], u[2] * axes[2]))
    return pts


def _add_bin(sc: Scene, half_width: float, wall_height: float = 62.0) -> None:
    """Static box floor (top face at y=0) and four static box walls, the
    container style of the reference's own benchmark (examples/Benchmark.cpp:
    28-40).  Plane walls are avoided on purpose: a plane's 1e6-wide AABB
    degrades the reference's SAH tree (8 ms instead of 1.9 ms of broadphase
    per step at 2k bodies, measured)."""
    h, hy = half_width, 0.5 * wall_height
    boxes = [
        ((0.0, -0.5, 0.0), (h, 0.5, h)),
        ((-h - 0.5, hy, 0.0), (0.5, hy, h)),
        ((h + 0.5, hy, 0.0), (0.5, hy, h)),
        ((0.0, hy, -h - 0.5), (h, hy, 0.5)),
        ((0.0, hy, h + 0.5), (h, hy, 0.5)),
    ]
    for pos, he in boxes:
        s = sc.add_shape(SHAPE_BOX, he)
        b = default_body(s, MOTION_STATIC)
        b["position"] = pos
        sc.add_body(b)


def random_convex_body(sc: Scene, rng: SplitMix64, kind: int) -> np.ndarray:
    """One body of the pile mix; ``kind`` 0..3 = sphere, box, capsule, hull."""
    if kind == 0:
        s = sc.add_shape(SHAPE_SPHERE, (rng.uniform(0.25, 0.75),))
    elif kind == 1:
        s = sc.add_shape(SHAPE_BOX, (rng.uniform(0.2, 0.6), rng.uniform(0.2, 0.6), rng.uniform(0.2, 0.6)))
    elif kind == 2:
        r = rng.uniform(0.2, 0.4)
        hh = rng.uniform(0.2, 0.8)
        s = sc.add_shape(SHAPE_CAPSULE, (hh, r))
    else:
        n = rng.integer(16, 32)
        axes = (rng.uniform(0.3, 0.7), rng.uniform(0.3, 0.7), rng.uniform(0.3, 0.7))
        s = sc.add_shape(SHAPE_HULL, hull=ellipsoid_hull_points(rng, n, axes))
    b = default_body(s)
    b["rotation"] = rng.quaternion()
    return b


def pile(num_bodies: int = 32768, seed: int = 0x4B5A0002, width: float = 64.0,
         y_range=(2.0, 60.0)) -> Scene:
    """Random convex pile dropped into a walled bin (sleeping disabled)."""
    sc = Scene(f"pile{num_bodies}", world_settings(8, 2))
    rng = SplitMix64(seed)
    half = 0.5 * width
    _add_bin(sc, half)
    margin = 1.0
    for i in range(num_bodies):
        b = random_convex_body(sc, rng, i % 4)
        b["position"] = (rng.uniform(-half + margin, half - margin), rng.uniform(*y_range),
                         rng.uniform(-half + margin, half - margin))
        sc.add_body(b)
    return sc


def _ground_plane(sc: Scene) -> None:
    plane = sc.add_shape(SHAPE_PLANE, (0.0, 1.0, 0.0, 0.0))
    sc.add_body(default_body(plane, MOTION_STATIC))


def box_stack(num_bodies: int, rng: SplitMix64, name: str = "stack") -> Scene:
    """A random box stack on its own ground plane: ``num_bodies`` boxes of one
    random half extent h ~ U[0.3, 0.6] in pyramid rows (the widest row at the
    bottom, each row centred on the one below, a gap of 1% of a box between
    neighbours), rows starting exactly touching."""
    sc = Scene(name, world_settings(8, 2))
    _ground_plane(sc)
    h = rng.uniform(0.3, 0.6)
    box = sc.add_shape(SHAPE_BOX, (h, h, h))
    base = 1
    while base * (base + 1) // 2 < num_bodies:
        base += 1
    pitch = 2.0 * h * 1.01
    placed, row = 0, 0
    while placed < num_bodies:
        n = min(base - row, num_bodies - placed)
        x0 = -0.5 * (n - 1) * pitch
        for i in range(n):
            b = default_body(box)
            b["position"] = (x0 + i * pitch, h + row * 2.0 * h, 0.0)
            sc.add_body(b)
        placed += n
        row += 1
    return sc


def small_pile(num_bodies: int, rng: SplitMix64, name: str = "pile") -> Scene:
    """The random convex pile's body mix on its own ground plane, inside four
    static box walls; the bin width grows with the body count so that the
    drop column is about 4 bodies deep."""
    sc = Scene(name, world_settings(8, 2))
    _ground_plane(sc)
    width = max(4.0, math.sqrt(num_bodies / 4.0) * 1.6)
    half = 0.5 * width
    hy = 8.0
    for pos, he in (((-half - 0.5, hy, 0.0), (0.5, hy, half)), ((half + 0.5, hy, 0.0), (0.5, hy, half)),
                    ((0.0, hy, -half - 0.5), (half, hy, 0.5)), ((0.0, hy, half + 0.5), (half, hy, 0.5))):
        s = sc.add_shape(SHAPE_BOX, he)
        b = default_body(s, MOTION_STATIC)
        b["position"] = pos
        sc.add_body(b)
    top = 1.0 + 1.6 * num_bodies / (width * width) * 4.0
    for i in range(num_bodies):
        b = random_convex_body(sc, rng, i % 4)
        b["position"] = (rng.uniform(-half + 0.8, half - 0.8), rng.uniform(1.0, top), rng.uniform(-half + 0.8, half - 0.8))
        sc.add_body(b)
    return sc


def sleeping_pile(num_bodies: int, seed: int = 0x4B5A0108, late: int = 4, late_height: float = 30.0) -> Scene:
    """small_pile with sleeping on (the reference's default settings: 0.5 s
    below 0.05 m/s and 0.1 rad/s) and ``late`` more bodies released
    ``late_height`` above the pile's top, so that they land on a pile that
    has fallen asleep and wake it."""
    rng = SplitMix64(seed)
    sc = small_pile(num_bodies, rng, f"sleeping_pile{num_bodies}")
    sc.world["allow_sleeping"] = 1
    width = max(4.0, math.sqrt(num_bodies / 4.0) * 1.6)
    half = 0.5 * width
    top = 1.0 + 1.6 * num_bodies / (width * width) * 4.0
    for i in range(late):
        b = random_convex_body(sc, rng, i % 4)
        b["position"] = (rng.uniform(-half + 0.8, half - 0.8), top + late_height + 2.0 * i,
                         rng.uniform(-half + 0.8, half - 0.8))
        sc.add_body(b)
    return sc


def sleeping_field(grid: int = 10, seed: int = 7, half: float = 0.5) -> Scene:
    """Sleeping on: a grid x grid field of box stacks 1-3 high on a ground
    plane (1.6 apart), and grid*grid/3 more boxes released one above the
    other at random spots of the field, 0.9 apart from 6 m up.  The stacks
    fall asleep within a second and the late boxes wake them one by one,
    so most of the world sleeps while the rest moves: the case where the
    reference routes contacts of sleeping bodies through stale island
    indices."""
    sc = Scene(f"sleeping_field{grid}", world_settings(8, 2))
    sc.world["allow_sleeping"] = 1
    _ground_plane(sc)
    box = sc.add_shape(SHAPE_BOX, (half, half, half))
    rng = SplitMix64(seed)
    for gx in range(grid):
        for gz in range(grid):
            for level in range(1 + rng.integer(0, 3)):
                b = default_body(box)
                b["position"] = (gx * 1.6 + rng.uniform(-0.02, 0.02), half + 2.0 * half * level,
                                 gz * 1.6 + rng.uniform(-0.02, 0.02))
                sc.add_body(b)
    for k in range(grid * grid // 3):
        b = default_body(box)
        b["position"] = (rng.uniform(0.0, (grid - 1) * 1.6), 6.0 + 0.9 * k, rng.uniform(0.0, (grid - 1) * 1.6))
        sc.add_body(b)
    return sc


def batched_worlds(num_worlds: int = 4096, seed: int = 0x4B5A0004, min_bodies: int = 16,
                   max_bodies: int = 512) -> list:
    """Independent worlds of uneven size: body counts log-uniform in
    [min_bodies, max_bodies], each world a random box stack or a convex pile
    with its own ground plane (chosen with equal odds)."""
    rng = SplitMix64(seed)
    out = []
    for w in range(num_worlds):
        n = int(round(min_bodies * (max_bodies / min_bodies) ** rng.uniform(0.0, 1.0)))
        n = max(min_bodies, min(max_bodies, n))
        if rng.uniform(0.0, 1.0) < 0.5:
            out.append(box_stack(n, rng, f"stack{w}"))
        else:
            out.append(small_pile(n, rng, f"pile{w}"))
    return out


def _quat_rotate(q: np.ndarray, v: np.ndarray, inverse: bool = False) -> np.ndarray:
    """Rotate each row of v by the matching quaternion row of q (or its inverse)."""
    qv = -q[:, :3] if inverse else q[:, :3]
    t = 2.0 * np.cross(qv, v)
    return v + t * q[:, 3:4] + np.cross(qv, t)


def _support_points(shapes: np.ndarray, hull: np.ndarray, idx: np.ndarray, q: np.ndarray,
                    d: np.ndarray, chunk: int = 65536) -> np.ndarray:
    """Support point of each shape (rotated by q) in direction d, relative to
    the shape origin, float64 (rows: pairs)."""
    q = q.astype(np.float64)
    ld = _quat_rotate(q, d.astype(np.float64), inverse=True)
    kind = shapes["type"][idx]
    prm = shapes["params"][idx].astype(np.float64)
    loc = np.zeros_like(ld)
    sph = kind == SHAPE_SPHERE
    loc[sph] = ld[sph] / np.linalg.norm(ld[sph], axis=1, keepdims=True) * prm[sph, :1]
    box = kind == SHAPE_BOX
    loc[box] = np.where(ld[box] >= 0, prm[box, :3], -prm[box, :3])
    cap = kind == SHAPE_CAPSULE
    loc[cap] = ld[cap] / np.linalg.norm(ld[cap], axis=1, keepdims=True) * prm[cap, 1:2]
    loc[cap, 1] += np.where(ld[cap, 1] >= 0, prm[cap, 0], -prm[cap, 0])
    hl = np.nonzero(kind == SHAPE_HULL)[0]
    if len(hl):
        offs, cnts = shapes["hull_offset"], shapes["hull_count"]
        maxv = int(cnts[shapes["type"] == SHAPE_HULL].max())
        padded = np.zeros((len(shapes), maxv, 3))
        for s in np.nonzero(shapes["type"] == SHAPE_HULL)[0]:
            pts = hull[offs[s]: offs[s] + cnts[s]].astype(np.float64)
            padded[s, : len(pts)] = pts
            padded[s, len(pts):] = pts[0]
        for c0 in range(0, len(hl), chunk):
            sel = hl[c0: c0 + chunk]
            arg = np.einsum("nvk,nk->nv", padded[idx[sel]], ld[sel]).argmax(1)
            loc[sel] = padded[idx[sel], arg]
    return _quat_rotate(q, loc)


@dataclass
class PairBatch:
    """Pairs of (shape, transform) for the standalone narrowphase."""

    shapes: np.ndarray
    hull_points: np.ndarray
    shape_a: np.ndarray
    shape_b: np.ndarray
    xf_a: np.ndarray  # (n, 7): quaternion xyzw, translation xyz
    xf_b: np.ndarray


def pair_batch(num_pairs: int = 1 << 20, seed: int = 0x4B5A0003, num_shapes: int = 4096) -> PairBatch:
    """Synthetic convex pairs: one third hull-hull (8-64 vertices), one third
    box-hull, one third capsule-box, random rotations.  Along a random axis
    the pair is brought into touch and then pushed together by ~U[0, 0.05] m,
    or (10 %) pulled apart by ~U[0, 0.02] m."""
    rng = SplitMix64(seed)
    sc = Scene("pairs", world_settings())
    hulls, boxes, caps = [], [], []
    for i in range(num_shapes):
        k = i % 4
        if k in (0, 1):
            n = rng.integer(8, 64)
            axes = (rng.uniform(0.3, 0.7), rng.uniform(0.3, 0.7), rng.uniform(0.3, 0.7))
            hulls.append(sc.add_shape(SHAPE_HULL, hull=ellipsoid_hull_points(rng, n, axes)))
        elif k == 2:
            boxes.append(sc.add_shape(SHAPE_BOX, (rng.uniform(0.2, 0.6), rng.uniform(0.2, 0.6), rng.uniform(0.2, 0.6))))
        else:
            caps.append(sc.add_shape(SHAPE_CAPSULE, (rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.4))))
    shapes = sc.shape_array()
    hull = sc.hull_array()
    # Vectorised pair placement with numpy (seeded from the same stream).
    g = np.random.Generator(np.random.PCG64(rng.next_u64()))
    cls = np.arange(num_pairs) % 3
    hulls, boxes, caps = np.array(hulls), np.array(boxes), np.array(caps)
    sa = np.where(cls == 0, g.choice(hulls, num_pairs), np.where(cls == 1, g.choice(boxes, num_pairs), g.choice(caps, num_pairs)))
    sb = np.where(cls == 2, g.choice(boxes, num_pairs), g.choice(hulls, num_pairs))

    def quats(n):
        u = g.random((n, 3))
        a, b = np.sqrt(1 - u[:, 0]), np.sqrt(u[:, 0])
        t2, t3 = 2 * np.pi * u[:, 1], 2 * np.pi * u[:, 2]
        return np.stack([a * np.sin(t2), a * np.cos(t2), b * np.sin(t3), b * np.cos(t3)], 1)

    qa, qb = quats(num_pairs), quats(num_pairs)
    axis = g.normal(size=(num_pairs, 3))
    axis /= np.linalg.norm(axis, axis=1, keepdims=True)
    pen = g.uniform(0.0, 0.05, num_pairs)
    sep = g.uniform(0.0, 0.02, num_pairs)
    separated = g.random(num_pairs) < 0.10
    offset = np.where(separated, sep, -pen)
    pa = g.uniform(-100.0, 100.0, (num_pairs, 3))
    # Make B's support point in -axis coincide with A's support point in
    # +axis, then move B along the axis: the pair touches, with penetration
    # at most `pen` or distance at most `sep`.
    touch = _support_points(shapes, hull, sa, qa, axis) - _support_points(shapes, hull, sb, qb, -axis)
    pb = pa + touch + axis * offset[:, None]
    xf_a = np.concatenate([qa, pa], 1).astype(np.float32)
    xf_b = np.concatenate([qb, pb], 1).astype(np.float32)
    return PairBatch(shapes, hull, sa.astype(np.uint32), sb.astype(np.uint32), xf_a, xf_b)


def _quat_from_to_y(n: np.ndarray) -> np.ndarray:
    """Quaternions (xyzw) rotating +Y onto each unit row of n (float64)."""
    y = np.array([0.0, 1.0, 0.0])
    c = np.clip(n @ y, -1.0, 1.0)
    axis = np.cross(np.broadcast_to(y, n.shape), n)
    ln = np.linalg.norm(axis, axis=1, keepdims=True)
    axis = np.where(ln > 1e-9, axis / np.maximum(ln, 1e-30), np.array([1.0, 0.0, 0.0]))
    half = 0.5 * np.arccos(c)
    return np.concatenate([axis * np.sin(half)[:, None], np.cos(half)[:, None]], 1)


def _quat_mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Row-wise Hamilton product a * b (xyzw)."""
    ax, ay, az, aw = a.T
    bx, by, bz, bw = b.T
    return np.stack([aw * bx + ax * bw + ay * bz - az * by, aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw, aw * bw - ax * bx - ay * by - az * bz], 1)


def prism_hull_points(rng: SplitMix64, sides: int, radius: float, half_height: float, extra: int) -> list:
    """A `sides`-gon prism (bottom ring at -half_height, then the top ring)
    followed by `extra` points on the lateral surface: hulls with flat faces
    that give plane pairs many touching vertices."""
    pts = []
    for y in (-half_height, half_height):
        for k in range(sides):
            a = 2.0 * np.pi * k / sides
            pts.append((radius * float(np.cos(a)), y, radius * float(np.sin(a))))
    for _ in range(extra):
        a = rng.uniform(0.0, 2.0 * np.pi)
        pts.append((radius * float(np.cos(a)), rng.uniform(-half_height, half_height), radius * float(np.sin(a))))
    return pts


def plane_pair_batch(num_pairs: int = 4096, seed: int = 0x4B5A0107, num_shapes: int = 256) -> PairBatch:
    """Plane-convex pairs for CollidePlaneConvex (CollideShape.cpp:370-430)
    and the plane dispatch of NarrowPhase::Collide (:511-520): planes with
    random normals and offsets against spheres, boxes, capsules, ellipsoid
    hulls of 8-64 vertices (more than the 16 the plane path reads) and prism
    hulls of 10-30 vertices; the plane is shape A in half of the pairs and B
    in the other half.  Half of the box and prism pairs rest a face on the
    plane within a tilt of ~1e-3 rad, so that more than 4 vertices lie within
    0.02 m and the reduction to 4 points runs; the rest take uniform random
    rotations.  Penetration U[0, 0.05] m; 10 % separated by U[0, 0.02] m and
    5 % by U[0.02, 0.1] m (no contact)."""
    rng = SplitMix64(seed)
    sc = Scene("plane-pairs", world_settings())
    planes, spheres, boxes, caps, ells, prisms = [], [], [], [], [], []
    for i in range(num_shapes):
        k = i % 6
        if k == 0:
            u = rng.unit_vector()
            planes.append(sc.add_shape(SHAPE_PLANE, (u[0], u[1], u[2], rng.uniform(-1.0, 1.0))))
        elif k == 1:
            spheres.append(sc.add_shape(SHAPE_SPHERE, (rng.uniform(0.25, 0.75),)))
        elif k == 2:
            boxes.append(sc.add_shape(SHAPE_BOX, (rng.uniform(0.2, 0.6), rng.uniform(0.2, 0.6), rng.uniform(0.2, 0.6))))
        elif k == 3:
            caps.append(sc.add_shape(SHAPE_CAPSULE, (rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.4))))
        elif k == 4:
            n = rng.integer(8, 64)
            axes = (rng.uniform(0.3, 0.7), rng.uniform(0.3, 0.7), rng.uniform(0.3, 0.7))
            ells.append(sc.add_shape(SHAPE_HULL, hull=ellipsoid_hull_points(rng, n, axes)))
        else:
            sides = rng.integer(5, 12)
            prisms.append(sc.add_shape(SHAPE_HULL, hull=prism_hull_points(
                rng, sides, rng.uniform(0.3, 0.6), rng.uniform(0.1, 0.4), rng.integer(0, 6))))
    shapes = sc.shape_array()
    hull = sc.hull_array()
    g = np.random.Generator(np.random.PCG64(rng.next_u64()))
    kinds = [np.array(spheres), np.array(boxes), np.array(caps), np.array(ells), np.array(prisms)]
    cls = np.arange(num_pairs) % len(kinds)
    conv = np.zeros(num_pairs, np.int64)
    for k, arr in enumerate(kinds):
        sel = cls == k
        conv[sel] = g.choice(arr, int(sel.sum()))
    pl = g.choice(np.array(planes), num_pairs)

    def quats(n):
        u = g.random((n, 3))
        a, b = np.sqrt(1 - u[:, 0]), np.sqrt(u[:, 0])
        t2, t3 = 2 * np.pi * u[:, 1], 2 * np.pi * u[:, 2]
        return np.stack([a * np.sin(t2), a * np.cos(t2), b * np.sin(t3), b * np.cos(t3)], 1)

    qp = quats(num_pairs)
    tp = g.uniform(-20.0, 20.0, (num_pairs, 3))
    prm = shapes["params"][pl].astype(np.float64)
    nw = _quat_rotate(qp, prm[:, :3])
    dw = prm[:, 3] + np.einsum("ij,ij->i", nw, tp)
    qc = quats(num_pairs)
    flat = ((cls == 1) | (cls == 4)) & (g.random(num_pairs) < 0.5)
    if flat.any():
        m = int(flat.sum())
        spin = np.zeros((m, 4))
        ang = g.uniform(0.0, 2.0 * np.pi, m)
        spin[:, 1], spin[:, 3] = np.sin(0.5 * ang), np.cos(0.5 * ang)
        tilt_axis = g.normal(size=(m, 3))
        tilt_axis /= np.linalg.norm(tilt_axis, axis=1, keepdims=True)
        tilt = g.uniform(0.0, 1e-3, m)
        tq = np.concatenate([tilt_axis * np.sin(0.5 * tilt)[:, None], np.cos(0.5 * tilt)[:, None]], 1)
        qc[flat] = _quat_mul(tq, _quat_mul(_quat_from_to_y(nw[flat]), spin))
    pen = g.uniform(0.0, 0.05, num_pairs)
    r = g.random(num_pairs)
    offset = np.where(r < 0.10, g.uniform(0.0, 0.02, num_pairs), np.where(r < 0.15, g.uniform(0.02, 0.1, num_pairs),
                                                                            -pen))
    lateral = g.uniform(-5.0, 5.0, (num_pairs, 3))
    sup = _support_points(shapes, hull, conv, qc, -nw)
    pc = tp + lateral
    pc += nw * (dw + offset - np.einsum("ij,ij->i", nw, pc + sup))[:, None]
    xf_p = np.concatenate([qp, tp], 1).astype(np.float32)
    xf_c = np.concatenate([qc, pc], 1).astype(np.float32)
    first = (np.arange(num_pairs) // len(kinds)) % 2 == 0  # plane as shape A
    sa = np.where(first, pl, conv).astype(np.uint32)
    sb = np.where(first, conv, pl).astype(np.uint32)
    xa = np.where(first[:, None], xf_p, xf_c)
    xb = np.where(first[:, None], xf_c, xf_p)
    return PairBatch(shapes, hull, sa, sb, np.ascontiguousarray(xa), np.ascontiguousarray(xb))
