"""ctypes front end of the CPU oracle (oracle/_build/libkzoracle.so).

TEST INFRASTRUCTURE ONLY: the checker used by tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg.  The product never imports this module.

The oracle needs the per-shape and per-body data that the engine derives at
body creation: built convex hulls (vertices + triangles) and mass data
(inverse mass + inverse local inertia).  ``Derived`` carries them; tests fill
it from the reference build (oracle/refharness.py) or from the product, so
that each side's derived data is itself checked separately.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

from kizuriphysics3d_amd import scenes as S

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libkzoracle.so")
POINT_FLOATS = 44
HULL_REF_DTYPE = np.dtype([("vert_offset", "<u4"), ("vert_count", "<u4"),
                           ("tri_offset", "<u4"), ("tri_count", "<u4")])

_lib = None


def build() -> str:
    """Compile the oracle with its Makefile (g++)."""
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, u32, f32, i32 = C.c_void_p, C.c_uint32, C.c_float, C.c_int
        L.kzo_world_create.argtypes = [i32, vp, vp, u32, vp, vp, vp, vp, u32, vp]
        L.kzo_world_create.restype = vp
        L.kzo_world_destroy.argtypes = [vp]
        L.kzo_set_warm_start.argtypes = [vp, i32]
        L.kzo_step.argtypes = [vp, f32, vp, u32, i32]
        L.kzo_num_constraints.argtypes = [vp]; L.kzo_num_constraints.restype = u32
        L.kzo_get_states.argtypes = [vp, vp]
        L.kzo_get_states_f64.argtypes = [vp, vp]
        L.kzo_set_states.argtypes = [vp, vp]
        L.kzo_get_manifolds.argtypes = [vp, vp, vp, vp]
        L.kzo_get_constraints.argtypes = [vp, vp, vp]
        L.kzo_get_residuals.argtypes = [vp, vp, u32]; L.kzo_get_residuals.restype = u32
        L.kzo_collide_batch.argtypes = [i32, vp, u32, vp, vp, vp, vp, vp, vp, vp, u32, f32, vp, vp, i32]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


@dataclass
class Derived:
    """Built hulls (packed) and body mass data for one shape/body table."""

    hull_refs: np.ndarray    # HULL_REF_DTYPE per shape
    hull_verts: np.ndarray   # (V, 3) float32
    hull_tris: np.ndarray    # (T, 3) uint32
    mass: np.ndarray         # (num_bodies, 10) float32

    @staticmethod
    def pack(shapes: np.ndarray, hulls: dict, mass: np.ndarray | None) -> "Derived":
        refs = np.zeros(len(shapes), HULL_REF_DTYPE)
        verts, tris = [np.zeros((0, 3), np.float32)], [np.zeros((0, 3), np.uint32)]
        nv = nt = 0
        for i, (v, t) in sorted(hulls.items()):
            refs[i] = (nv, len(v), nt, len(t))
            verts.append(np.asarray(v, np.float32))
            tris.append(np.asarray(t, np.uint32))
            nv += len(v)
            nt += len(t)
        hv = np.ascontiguousarray(np.concatenate(verts) if nv else np.zeros((1, 3), np.float32), np.float32)
        ht = np.ascontiguousarray(np.concatenate(tris) if nt else np.zeros((1, 3), np.uint32), np.uint32)
        m = np.zeros((0, 10), np.float32) if mass is None else np.ascontiguousarray(mass, np.float32)
        return Derived(refs, hv, ht, m)


class OracleWorld:
    """The oracle's world: per-step path over externally supplied pairs."""

    def __init__(self, scene: S.Scene, derived: Derived, precision: int = 0):
        L = lib()
        self._world = np.array(scene.world, dtype=S.WORLD_DTYPE)
        self._shapes = scene.shape_array()
        self._bodies = scene.body_array()
        self._d = derived
        self.n = len(self._bodies)
        assert derived.mass.shape == (self.n, 10)
        self.h = L.kzo_world_create(precision, _p(self._world), _p(self._shapes), len(self._shapes),
                                    _p(derived.hull_refs), _p(derived.hull_verts), _p(derived.hull_tris),
                                    _p(self._bodies), self.n, _p(derived.mass))

    def close(self):
        if self.h:
            lib().kzo_world_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_warm_start(self, mode: int):
        lib().kzo_set_warm_start(self.h, mode)

    def step(self, pairs: np.ndarray, dt=S.DT, stop_after: int = 5):
        pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
        lib().kzo_step(self.h, float(dt), _p(pairs), len(pairs), stop_after)

    def states(self) -> np.ndarray:
        out = np.zeros(self.n, S.STATE_DTYPE)
        lib().kzo_get_states(self.h, _p(out))
        return out

    def states_f64(self) -> np.ndarray:
        out = np.zeros((self.n, 13), np.float64)
        lib().kzo_get_states_f64(self.h, _p(out))
        return out

    def set_states(self, st: np.ndarray):
        st = np.ascontiguousarray(st, dtype=S.STATE_DTYPE)
        lib().kzo_set_states(self.h, _p(st))

    def manifolds(self, num_pairs: int):
        counts = np.zeros(num_pairs, np.uint32)
        pts = np.zeros((num_pairs, 8), S.CONTACT_DTYPE)
        keys = np.zeros(num_pairs, np.uint64)
        if num_pairs:
            lib().kzo_get_manifolds(self.h, _p(counts), _p(pts), _p(keys))
        return counts, pts, keys

    def constraints(self):
        n = lib().kzo_num_constraints(self.h)
        hdr = np.zeros((n, 4), np.uint32)
        pts = np.zeros((n, 8, POINT_FLOATS), np.float32)
        if n:
            lib().kzo_get_constraints(self.h, _p(hdr), _p(pts))
        return hdr, pts

    def residuals(self) -> np.ndarray:
        out = np.zeros(64, np.float64)
        k = lib().kzo_get_residuals(self.h, _p(out), 64)
        return out[: min(k, 64)].copy()


def collide_batch(pb: S.PairBatch, derived: Derived, max_separation=0.02, precision=0, threads=1):
    """Oracle narrowphase over a PairBatch -> (counts, points (n, 8))."""
    n = len(pb.shape_a)
    counts = np.zeros(n, np.uint32)
    pts = np.zeros((n, 8), S.CONTACT_DTYPE)
    shapes = np.ascontiguousarray(pb.shapes)
    sa = np.ascontiguousarray(pb.shape_a, np.uint32)
    sb = np.ascontiguousarray(pb.shape_b, np.uint32)
    xa = np.ascontiguousarray(pb.xf_a, np.float32)
    xb = np.ascontiguousarray(pb.xf_b, np.float32)
    lib().kzo_collide_batch(precision, _p(shapes), len(shapes), _p(derived.hull_refs), _p(derived.hull_verts),
                            _p(derived.hull_tris), _p(sa), _p(sb), _p(xa), _p(xb), n, float(max_separation),
                            _p(counts), _p(pts), threads)
    return counts, pts


def _qrot(q: np.ndarray, v: np.ndarray) -> np.ndarray:
    """Rotate rows of v by quaternions q (xyzw), float64: v + 2w(u x v) + 2u x (u x v)."""
    u, w = q[:, :3], q[:, 3:4]
    t = 2.0 * np.cross(u, v)
    return v + w * t + np.cross(u, t)


def natural_residual(hdr: np.ndarray, pts: np.ndarray, states: np.ndarray, velocity_bias: bool = False) -> dict:
    """Natural-map residual of a solved iterate, in float64 (the device's
    kzg_get_natural_residual restated).  hdr / pts: a constraint dump
    (body A, body B, points, flags; 44 floats per point); states: the bodies
    after the solve (velocities) with the rotations the solve used.  Per
    non-sensor point, rA = qA localAnchorA, vRel = (vA + wA x rA) - (vB + wB x rB):
      normal   min(lambda_n, m_n (vRel . n - bias))
      friction lambda_t - clamp(lambda_t - m_t (vRel . t), +-mu lambda_n) for t1, t2
      cone     max(0, |lambda_t| - mu lambda_n) for t1, t2
    bias = restitution bias (+ position bias with velocity bias).  Returns
    the RMS values (N s) of normal + friction, normal, friction and cone."""
    k = hdr[:, 2].astype(np.int64)
    live = (np.arange(8)[None, :] < k[:, None]) & ((hdr[:, 3] & 1) == 0)[:, None]
    ci, pi = np.nonzero(live)
    P = pts[ci, pi].astype(np.float64)
    a, b = hdr[ci, 0], hdr[ci, 1]
    q = lambda i: np.stack([states["rotation"][i][:, j] for j in range(4)], 1).astype(np.float64)
    ra = _qrot(q(a), P[:, 12:15])
    rb = _qrot(q(b), P[:, 15:18])
    va, wa = states["linear_velocity"][a].astype(np.float64), states["angular_velocity"][a].astype(np.float64)
    vb, wb = states["linear_velocity"][b].astype(np.float64), states["angular_velocity"][b].astype(np.float64)
    vr = (va + np.cross(wa, ra)) - (vb + np.cross(wb, rb))
    n, t1, t2 = P[:, 18:21], P[:, 21:24], P[:, 24:27]
    mn, mt = P[:, 27], (P[:, 28], P[:, 29])
    ln, lt = P[:, 30], (P[:, 31], P[:, 32])
    bias = P[:, 34] + (P[:, 35] if velocity_bias else 0.0)
    mu = P[:, 37]
    lim = mu * ln
    rn = np.minimum(ln, mn * ((vr * n).sum(1) - bias))
    rt = np.zeros_like(rn)
    rc = np.zeros_like(rn)
    for tk, mk, lk in ((t1, mt[0], lt[0]), (t2, mt[1], lt[1])):
        step = lk - mk * (vr * tk).sum(1)
        rt += (lk - np.clip(step, -lim, lim)) ** 2
        rc += np.maximum(0.0, np.abs(lk) - lim) ** 2
    cnt = max(len(rn), 1)
    return {"total": float(np.sqrt(((rn ** 2).sum() + rt.sum()) / cnt)), "normal": float(np.sqrt((rn ** 2).sum() / cnt)),
            "friction": float(np.sqrt(rt.sum() / cnt)), "cone": float(np.sqrt(rc.sum() / cnt))}
