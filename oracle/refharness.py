"""ctypes front end of oracle/_ref/libkzref.so, the REFERENCE engine built from
its own sources (oracle/Makefile target ``ref``) plus oracle/ref_harness.cpp.

TEST INFRASTRUCTURE ONLY: used by tests/, by tests/golden/make_fixtures.py and
by bench.py's cpu_baseline leg.  The product never loads it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from kizuriphysics3d_amd import scenes as S

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "libkzref.so")
POINT_FLOATS = 44  # floats per constraint point in kzr_get_constraints

_lib = None


def available() -> bool:
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise FileNotFoundError(f"{LIB_PATH} is not built (make -C oracle ref)")
        L = C.CDLL(LIB_PATH)
        vp, u32, f32, f64, i32 = C.c_void_p, C.c_uint32, C.c_float, C.c_double, C.c_int
        L.kzr_init_jobs.argtypes = [u32]; L.kzr_init_jobs.restype = u32
        L.kzr_world_create.argtypes = [vp, vp, u32, vp, vp, u32]; L.kzr_world_create.restype = vp
        L.kzr_world_destroy.argtypes = [vp]
        L.kzr_step.argtypes = [vp, f32]
        L.kzr_step_phases.argtypes = [vp, f32, vp, i32]
        L.kzr_num_bodies.argtypes = [vp]; L.kzr_num_bodies.restype = u32
        L.kzr_get_states.argtypes = [vp, vp]
        L.kzr_set_states.argtypes = [vp, vp]
        L.kzr_get_mass.argtypes = [vp, vp]
        L.kzr_num_pairs.argtypes = [vp]; L.kzr_num_pairs.restype = u32
        L.kzr_get_pairs.argtypes = [vp, vp]
        L.kzr_get_manifolds.argtypes = [vp, vp, vp, vp]
        L.kzr_num_constraints.argtypes = [vp]; L.kzr_num_constraints.restype = u32
        L.kzr_get_constraints.argtypes = [vp, vp, vp]
        L.kzr_get_islands.argtypes = [vp, vp]; L.kzr_get_islands.restype = u32
        L.kzr_get_active.argtypes = [vp, vp]; L.kzr_get_active.restype = u32
        L.kzr_state_hash.argtypes = [vp]; L.kzr_state_hash.restype = C.c_uint64
        L.kzr_hull_create.argtypes = [vp, u32, vp, vp, vp, vp]; L.kzr_hull_create.restype = i32
        L.kzr_collide_batch.argtypes = [vp, u32, vp, vp, vp, vp, vp, u32, f32, vp, vp, i32]
        L.kzr_collide_batch.restype = f64
        L.kzr_save_kzp1.argtypes = [vp, vp, C.c_uint64]; L.kzr_save_kzp1.restype = C.c_uint64
        L.kzr_load_kzp1.argtypes = [vp, vp, C.c_uint64]; L.kzr_load_kzp1.restype = vp
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def init_jobs(workers: int = 0) -> int:
    return lib().kzr_init_jobs(workers)


class RefWorld:
    """A reference PhysicsWorld built from a scenes.Scene."""

    def __init__(self, scene: S.Scene | None, world: np.ndarray | None = None):
        L = lib()
        if scene is None:  # filled in by load_kzp1()
            self.h = None
            return
        self._world = np.array(scene.world if world is None else world, dtype=S.WORLD_DTYPE)
        self._shapes = scene.shape_array()
        self._hull = np.ascontiguousarray(scene.hull_array(), dtype=np.float32)
        self._bodies = scene.body_array()
        self.h = L.kzr_world_create(_p(self._world), _p(self._shapes), len(self._shapes),
                                    _p(self._hull), _p(self._bodies), len(self._bodies))
        if not self.h:
            raise ValueError("reference world creation failed")
        self.n = len(self._bodies)

    def save_kzp1(self) -> bytes:
        """The reference's SerializeWorld record of this world."""
        n = lib().kzr_save_kzp1(self.h, None, 0)
        if n == 0:
            raise ValueError("SerializeWorld failed")
        buf = np.zeros(n, np.uint8)
        lib().kzr_save_kzp1(self.h, _p(buf), n)
        return buf.tobytes()

    @classmethod
    def load_kzp1(cls, data: bytes, world: np.ndarray) -> "RefWorld":
        """A fresh reference world with settings `world` filled by DeserializeWorld."""
        ws = np.array(world, dtype=S.WORLD_DTYPE)
        buf = np.frombuffer(data, np.uint8).copy()
        h = lib().kzr_load_kzp1(_p(ws), _p(buf), len(buf))
        if not h:
            raise ValueError("DeserializeWorld failed")
        rw = cls(None)
        rw.h = h
        rw._world = ws
        rw.n = int(lib().kzr_num_bodies(h))
        return rw

    def close(self):
        if self.h:
            lib().kzr_world_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, dt=S.DT):
        lib().kzr_step(self.h, float(dt))

    def step_phases(self, dt=S.DT, stop_after=5) -> np.ndarray:
        ms = np.zeros(6, np.float64)
        lib().kzr_step_phases(self.h, float(dt), _p(ms), stop_after)
        return ms

    def states(self) -> np.ndarray:
        out = np.zeros(self.n, S.STATE_DTYPE)
        lib().kzr_get_states(self.h, _p(out))
        return out

    def set_states(self, st: np.ndarray):
        st = np.ascontiguousarray(st, dtype=S.STATE_DTYPE)
        lib().kzr_set_states(self.h, _p(st))

    def mass(self) -> np.ndarray:
        out = np.zeros((self.n, 19), np.float32)
        lib().kzr_get_mass(self.h, _p(out))
        return out

    def pairs(self) -> np.ndarray:
        n = lib().kzr_num_pairs(self.h)
        out = np.zeros((n, 2), np.uint32)
        if n:
            lib().kzr_get_pairs(self.h, _p(out))
        return out

    def manifolds(self):
        n = lib().kzr_num_pairs(self.h)
        counts = np.zeros(n, np.uint32)
        pts = np.zeros((n, 8), S.CONTACT_DTYPE)
        keys = np.zeros(n, np.uint64)
        if n:
            lib().kzr_get_manifolds(self.h, _p(counts), _p(pts), _p(keys))
        return counts, pts, keys

    def constraints(self):
        n = lib().kzr_num_constraints(self.h)
        hdr = np.zeros((n, 4), np.uint32)
        pts = np.zeros((n, 8, POINT_FLOATS), np.float32)
        if n:
            lib().kzr_get_constraints(self.h, _p(hdr), _p(pts))
        return hdr, pts

    def active(self) -> np.ndarray:
        out = np.zeros(max(self.n, 1), np.uint32)
        k = lib().kzr_get_active(self.h, _p(out))
        return out[:k]

    def islands(self):
        n = lib().kzr_num_constraints(self.h)
        out = np.zeros(max(n, 1), np.uint32)
        k = lib().kzr_get_islands(self.h, _p(out))
        return k, out[:n]

    def state_hash(self) -> int:
        return int(lib().kzr_state_hash(self.h))


def hull_create(points: np.ndarray):
    """ConvexHullShape::Create -> (points (V,3), triangles (F,3)) or None."""
    pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    op = np.zeros((256, 3), np.float32)
    of = np.zeros((512, 3), np.uint32)
    nv, nf = np.zeros(1, np.uint32), np.zeros(1, np.uint32)
    rc = lib().kzr_hull_create(_p(pts), len(pts), _p(op), _p(nv), _p(of), _p(nf))
    if rc != 0:
        return None
    return op[: nv[0]].copy(), of[: nf[0]].copy()


def collide_batch(pb: S.PairBatch, max_separation=0.02, parallel=True):
    """NarrowPhase::Collide over a PairBatch -> (counts, points (n,8), ms)."""
    n = len(pb.shape_a)
    counts = np.zeros(n, np.uint32)
    pts = np.zeros((n, 8), S.CONTACT_DTYPE)
    shapes = np.ascontiguousarray(pb.shapes)
    hull = np.ascontiguousarray(pb.hull_points, dtype=np.float32)
    sa = np.ascontiguousarray(pb.shape_a, dtype=np.uint32)
    sb = np.ascontiguousarray(pb.shape_b, dtype=np.uint32)
    xa = np.ascontiguousarray(pb.xf_a, dtype=np.float32)
    xb = np.ascontiguousarray(pb.xf_b, dtype=np.float32)
    ms = lib().kzr_collide_batch(_p(shapes), len(shapes), _p(hull), _p(sa), _p(sb), _p(xa), _p(xb),
                                 n, float(max_separation), _p(counts), _p(pts), 1 if parallel else 0)
    if ms < 0:
        raise ValueError("bad shape in pair batch")
    return counts, pts, ms
