"""Python front end of the host world (include/kzw.h): the PhysicsWorld of
the reference with its contact path on the GPU.

    w = PhysicsWorld(scene)          # shapes + bodies from a scenes.Scene
    w.step_fixed(1/60)               # PhysicsWorld::StepFixed
    w.states(), w.stats()

``device=-1`` gives a host-only world (no GPU): hull building, mass data and
the broadphase work, stepping raises.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import device as D
from . import scenes as S

STATS_DTYPE = np.dtype([
    ("num_bodies", "<u4"), ("num_broadphase_pairs", "<u4"), ("num_contact_constraints", "<u4"),
    ("num_contact_points", "<u4"), ("broadphase_ms", "<f8"), ("contact_path_ms", "<f8"), ("step_ms", "<f8"),
    ("broadphase_move_ms", "<f8"), ("broadphase_query_ms", "<f8"), ("num_active_bodies", "<u4"),
    ("broadphase_reinserted", "<u4"),
], align=True)

_bound = False


def _lib():
    global _bound
    L = D.library()
    if not _bound:
        vp, u32, f32, i32 = C.c_void_p, C.c_uint32, C.c_float, C.c_int
        sig = {
            "kzw_create": ([vp, vp, i32], i32), "kzw_destroy": ([vp], None), "kzw_last_error": ([vp], C.c_char_p),
            "kzw_add_shapes": ([vp, vp, u32, vp, u32], i32), "kzw_add_bodies": ([vp, vp, u32], i32),
            "kzw_step": ([vp, f32], i32), "kzw_update_broadphase": ([vp], i32),
            "kzw_get_states": ([vp, vp, u32], i32), "kzw_set_states": ([vp, vp, u32], i32),
            "kzw_get_pairs": ([vp, vp, u32], i32), "kzw_get_stats": ([vp, vp], i32),
            "kzw_get_mass": ([vp, vp, u32], i32), "kzw_get_hull": ([vp, u32, vp, vp, vp, vp], i32),
            "kzw_device": ([vp], vp),
            "kzw_save": ([vp, vp, C.c_uint64, vp], i32), "kzw_load": ([vp, vp, C.c_uint64, vp, i32], i32),
            "kzw_set_device_broadphase": ([vp, i32], i32), "kzw_set_host_threads": ([vp, i32], i32),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _bound = True
    return L


KZW_EXPORTS = ["kzw_create", "kzw_destroy", "kzw_last_error", "kzw_add_shapes", "kzw_add_bodies", "kzw_step",
               "kzw_update_broadphase", "kzw_get_states", "kzw_set_states", "kzw_get_pairs", "kzw_get_stats",
               "kzw_get_mass", "kzw_get_hull", "kzw_device", "kzw_save", "kzw_load", "kzw_set_device_broadphase",
               "kzw_set_host_threads"]


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class PhysicsWorld:
    """A world built from a scenes.Scene; the contact path runs on the GPU."""

    def __init__(self, scene: S.Scene | None, device: int = 0):
        L = _lib()
        if scene is None:  # filled in by load()
            self.h = None
            self.device = device
            return
        self._world = np.array(scene.world, dtype=S.WORLD_DTYPE)
        h = C.c_void_p()
        rc = L.kzw_create(C.byref(h), _p(self._world), device)
        if rc != 0:
            raise D.KzgError(f"kzw_create failed ({rc})" + (": no usable GPU" if rc == -2 else ""))
        self.h = h
        self.device = device
        shapes = scene.shape_array()
        hull = np.ascontiguousarray(scene.hull_array(), np.float32)
        self._check(L.kzw_add_shapes(self.h, _p(shapes), len(shapes), _p(hull), len(hull)), "kzw_add_shapes")
        bodies = scene.body_array()
        self._check(L.kzw_add_bodies(self.h, _p(bodies), len(bodies)), "kzw_add_bodies")
        self.num_shapes = len(shapes)
        self.num_bodies = len(bodies)

    def save(self) -> bytes:
        """Snapshot (kzw_save): KZP1 world record + resumable extension."""
        size = C.c_uint64()
        self._check(_lib().kzw_save(self.h, None, 0, C.byref(size)), "kzw_save")
        buf = np.zeros(size.value, np.uint8)
        self._check(_lib().kzw_save(self.h, _p(buf), size.value, C.byref(size)), "kzw_save")
        return buf.tobytes()

    @classmethod
    def load(cls, data: bytes, device: int = 0, world: np.ndarray | None = None) -> "PhysicsWorld":
        """World from a snapshot or a plain KZP1 record (then `world`, a
        WORLD_DTYPE record, gives the settings)."""
        L = _lib()
        buf = np.frombuffer(data, np.uint8)
        ws = None if world is None else np.array(world, dtype=S.WORLD_DTYPE)
        h = C.c_void_p()
        rc = L.kzw_load(C.byref(h), _p(buf), len(buf), None if ws is None else _p(ws), device)
        if rc != 0:
            raise D.KzgError(f"kzw_load failed ({rc})")
        w = cls(None, device)
        w.h = h
        # body count from the record header (KZP1: magic, version, count)
        w.num_bodies = int(np.frombuffer(data[8:12], np.uint32)[0])
        w.num_shapes = None
        w._world = ws
        return w

    def close(self):
        if getattr(self, "h", None):
            _lib().kzw_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc < 0:
            msg = _lib().kzw_last_error(self.h)
            raise D.KzgError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
        return rc

    def step_fixed(self, dt=S.DT):
        self._check(_lib().kzw_step(self.h, float(dt)), "kzw_step")

    def set_host_threads(self, threads: int = 1):
        self._check(_lib().kzw_set_host_threads(self.h, int(threads)), "kzw_set_host_threads")

    def set_device_broadphase(self, on):
        """step_fixed's pair query on the device (default) or on host threads."""
        self._check(_lib().kzw_set_device_broadphase(self.h, int(on)), "kzw_set_device_broadphase")

    def update_broadphase(self):
        self._check(_lib().kzw_update_broadphase(self.h), "kzw_update_broadphase")

    def states(self) -> np.ndarray:
        out = np.zeros(self.num_bodies, S.STATE_DTYPE)
        self._check(_lib().kzw_get_states(self.h, _p(out), self.num_bodies), "kzw_get_states")
        return out

    def set_states(self, st):
        st = np.ascontiguousarray(st, dtype=S.STATE_DTYPE)
        self._check(_lib().kzw_set_states(self.h, _p(st), len(st)), "kzw_set_states")

    def pairs(self) -> np.ndarray:
        n = self._check(_lib().kzw_get_pairs(self.h, None, 0), "kzw_get_pairs")
        out = np.zeros((max(n, 1), 2), np.uint32)
        if n:
            _lib().kzw_get_pairs(self.h, _p(out), n)
        return out[:n]

    def stats(self) -> dict:
        s = np.zeros((), STATS_DTYPE)
        self._check(_lib().kzw_get_stats(self.h, _p(s)), "kzw_get_stats")
        return {k: s[k].item() for k in STATS_DTYPE.names}

    def mass(self) -> np.ndarray:
        out = np.zeros((self.num_bodies, 10), np.float32)
        self._check(_lib().kzw_get_mass(self.h, _p(out), self.num_bodies), "kzw_get_mass")
        return out

    def hull(self, shape: int):
        nv, nt = C.c_uint32(), C.c_uint32()
        self._check(_lib().kzw_get_hull(self.h, shape, None, None, C.byref(nv), C.byref(nt)), "kzw_get_hull")
        v = np.zeros((max(nv.value, 1), 3), np.float32)
        t = np.zeros((max(nt.value, 1), 3), np.uint32)
        _lib().kzw_get_hull(self.h, shape, _p(v), _p(t), None, None)
        return v[: nv.value], t[: nt.value]

    def hulls(self, shapes: np.ndarray) -> dict:
        return {i: self.hull(i) for i, s in enumerate(shapes) if s["type"] == S.SHAPE_HULL}

    def device_context_handle(self):
        return _lib().kzw_device(self.h)

    def device_context(self) -> D.DeviceContext:
        """The world's kzg context (borrowed: it lives as long as the world)."""
        h = self.device_context_handle()
        if not h:
            raise D.KzgError("the world has no device context")
        return D.DeviceContext.borrow(h, self.num_bodies)


def device_timings(world: PhysicsWorld) -> dict:
    """Per-phase device timings of the last step (HIP events, see kzg.h)."""
    h = world.device_context_handle()
    if not h:
        return {}
    t = np.zeros((), D.TIMINGS_DTYPE)
    rc = D.library().kzg_get_timings(h, _p(t))
    if rc != 0:
        raise D.KzgError(f"kzg_get_timings failed ({rc})")
    return {k: t[k].item() for k in D.TIMINGS_DTYPE.names}


def build_hulls(shapes: np.ndarray, hull_points: np.ndarray) -> dict:
    """ConvexHullShape::Create of every hull shape, on the host (no GPU):
    {shape index: (vertices (V, 3), triangles (T, 3))}."""
    sc = S.Scene("hulls", S.world_settings())
    sc.shapes = list(np.asarray(shapes, dtype=S.SHAPE_DTYPE))
    sc.hull_points = [tuple(p) for p in np.asarray(hull_points, np.float32).reshape(-1, 3)]
    w = PhysicsWorld(sc, device=-1)
    try:
        return w.hulls(sc.shape_array())
    finally:
        w.close()
