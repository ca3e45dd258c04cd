"""ctypes binding of the kzg_* C-ABI (include/kzg.h) in libkizuri_amd.so.

There is no CPU fallback: loading fails loudly when the library is missing and
``DeviceContext`` raises when no GPU is usable.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import scenes as S

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkizuri_amd.so")
POINT_FLOATS = 44

KZG_BODY_DTYPE = np.dtype([
    ("packed_id", "<u4"), ("shape", "<u4"), ("motion_type", "<u4"), ("is_sensor", "<u4"),
    ("position", "<f4", (3,)), ("rotation", "<f4", (4,)), ("linear_velocity", "<f4", (3,)),
    ("angular_velocity", "<f4", (3,)), ("inverse_mass", "<f4"), ("inverse_inertia_local", "<f4", (9,)),
    ("friction", "<f4"), ("restitution", "<f4"), ("linear_damping", "<f4"), ("angular_damping", "<f4"),
    ("gravity_factor", "<f4"), ("max_linear_velocity", "<f4"), ("max_angular_velocity", "<f4"),
    ("allow_sleeping", "<u4"),
])
HULL_RANGE_DTYPE = np.dtype([("vert_offset", "<u4"), ("vert_count", "<u4"), ("tri_offset", "<u4"), ("tri_count", "<u4")])
TIMINGS_DTYPE = np.dtype([
    ("narrowphase_ms", "<f4"), ("build_ms", "<f4"), ("integrate_velocities_ms", "<f4"), ("solve_ms", "<f4"),
    ("integrate_positions_ms", "<f4"), ("solve_kernel_ms", "<f4"), ("velocity_levels", "<u4"),
    ("bias_levels", "<u4"), ("groups", "<u4"), ("islands", "<u4"), ("build_kernel_ms", "<f4"),
    ("cooperative_blocks", "<u4"), ("cooperative_ws_points", "<u4"), ("cooperative_vel_points", "<u4"),
    ("cooperative_kernel_ms", "<f4"), ("schedule_ms", "<f4"), ("build_device_ms", "<f4"),
    ("schedule_ws_ms", "<f4"), ("schedule_wait_ms", "<f4"), ("early_warm_start_ms", "<f4"),
    ("schedule_velocity_after_ms", "<f4"), ("overlap_ms", "<f4"),
])
assert KZG_BODY_DTYPE.itemsize == 140

EXPORTS = [
    "kzg_create", "kzg_destroy", "kzg_last_error", "kzg_upload_shapes", "kzg_upload_bodies", "kzg_upload_states",
    "kzg_narrowphase", "kzg_broadphase_pairs", "kzg_broadphase_pairs_update", "kzg_broadphase_set_proxies",
    "kzg_broadphase_move", "kzg_broadphase_download", "kzg_narrowphase_resident", "kzg_download_pairs", "kzg_build_constraints", "kzg_integrate_velocities", "kzg_solve", "kzg_integrate_positions",
    "kzg_update_sleeping", "kzg_get_active_bodies", "kzg_set_active_bodies", "kzg_set_sleep_thresholds",
    "kzg_step", "kzg_query_sizes", "kzg_download_states", "kzg_download_manifolds", "kzg_download_constraints",
    "kzg_get_residuals", "kzg_get_natural_residual", "kzg_get_timings", "kzg_collide_batch", "kzg_collide_batch_device", "kzg_schedule", "kzg_set_warm_start",
    "kzg_set_iterations", "kzg_set_cooperative", "kzg_set_sweep_units", "kzg_set_early_warm_start", "kzg_set_narrowphase_waves", "kzg_set_island_lds_limit", "kzg_set_cooperative_placement", "kzg_set_spin_limit", "kzg_set_device_bucketing", "kzg_cache_info", "kzg_download_cache", "kzg_upload_cache",
    "kzg_mlp_init", "kzg_mlp_set_params", "kzg_mlp_get_params", "kzg_mlp_set_normalization",
    "kzg_mlp_get_normalization", "kzg_mlp_collect", "kzg_mlp_num_samples", "kzg_mlp_get_samples",
    "kzg_mlp_fit_normalization", "kzg_mlp_train", "kzg_mlp_set_target_clip", "kzg_set_learned_features", "kzg_mlp_forward", "kzg_mlp_gradients",
]

# Warm-start modes (kzg_warm_start in include/kzg.h).
WARM_CACHE, WARM_COLD, WARM_LEARNED, WARM_LEARNED_MISS, WARM_NO_STORE, WARM_NO_COLLECT = 0, 1, 2, 3, 0x100, 0x200
COOP_ONE_XCD, COOP_SPREAD = 0, 1
MLP_PARAMS, MLP_FEATURES, MLP_FEATURE_ROW, MLP_SAMPLE_ROW = 5443, 16, 16, 20
FEATURES_LOAD = 1

_lib = None


class KzgError(RuntimeError):
    pass


def library():
    """Load libkizuri_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is missing: run __graft_entry__.build() or "
                                    "python -m kizuriphysics3d_amd.build")
        L = C.CDLL(LIB_PATH)
        vp, u32, f32, i32 = C.c_void_p, C.c_uint32, C.c_float, C.c_int
        sig = {
            "kzg_create": ([vp, vp, i32], i32), "kzg_destroy": ([vp], None),
            "kzg_last_error": ([vp], C.c_char_p),
            "kzg_upload_shapes": ([vp, vp, u32, vp, vp, u32, vp, u32], i32),
            "kzg_upload_bodies": ([vp, vp, u32], i32), "kzg_upload_states": ([vp, vp, u32], i32),
            "kzg_narrowphase": ([vp, vp, u32], i32), "kzg_build_constraints": ([vp, i32], i32),
            "kzg_broadphase_pairs": ([vp, vp, u32, u32, vp, vp, u32, vp], i32),
            "kzg_narrowphase_resident": ([vp], i32), "kzg_download_pairs": ([vp, vp, u32], i32),
            "kzg_broadphase_pairs_update": ([vp, vp, vp, u32, u32, u32, vp], i32),
            "kzg_broadphase_set_proxies": ([vp, vp, vp, u32], i32),
            "kzg_broadphase_move": ([vp, vp, u32, vp], i32), "kzg_broadphase_download": ([vp, vp, u32], i32),
            "kzg_integrate_velocities": ([vp, f32], i32), "kzg_solve": ([vp, f32], i32),
            "kzg_integrate_positions": ([vp, f32], i32), "kzg_step": ([vp, vp, u32, f32], i32),
            "kzg_update_sleeping": ([vp, f32], i32), "kzg_get_active_bodies": ([vp, vp, u32, vp], i32),
            "kzg_set_active_bodies": ([vp, vp, u32], i32), "kzg_set_sleep_thresholds": ([vp, f32, f32, f32], i32),
            "kzg_query_sizes": ([vp, vp, vp, vp], i32), "kzg_download_states": ([vp, vp, u32], i32),
            "kzg_download_manifolds": ([vp, vp, vp], i32), "kzg_download_constraints": ([vp, vp, vp], i32),
            "kzg_get_residuals": ([vp, vp, u32], i32), "kzg_get_timings": ([vp, vp], i32),
            "kzg_get_natural_residual": ([vp, vp], i32),
            "kzg_collide_batch": ([vp, vp, vp, vp, vp, u32, f32, vp, vp, vp], i32),
            "kzg_collide_batch_device": ([vp, vp, vp, vp, vp, u32, f32, vp, vp], i32),
            "kzg_schedule": ([u32, vp, u32, vp, u32, u32, vp, vp, vp, vp, vp, vp], i32),
            "kzg_set_warm_start": ([vp, i32], i32), "kzg_set_iterations": ([vp, u32, u32], i32),
            "kzg_set_cooperative": ([vp, u32], i32),
            "kzg_set_sweep_units": ([vp, u32, u32], i32),
            "kzg_set_early_warm_start": ([vp, u32], i32),
            "kzg_set_narrowphase_waves": ([vp, u32], i32),
            "kzg_set_island_lds_limit": ([vp, u32], i32),
            "kzg_set_cooperative_placement": ([vp, u32], i32),
            "kzg_set_spin_limit": ([vp, u32], i32), "kzg_set_device_bucketing": ([vp, u32], i32),
            "kzg_cache_info": ([vp, vp, vp], i32), "kzg_download_cache": ([vp, vp, C.c_uint64], i32),
            "kzg_upload_cache": ([vp, vp, C.c_uint64, u32, u32], i32),
            "kzg_mlp_init": ([vp, C.c_uint64], i32), "kzg_mlp_set_params": ([vp, vp, u32], i32),
            "kzg_mlp_get_params": ([vp, vp, u32], i32), "kzg_mlp_set_normalization": ([vp, vp, vp], i32),
            "kzg_mlp_get_normalization": ([vp, vp, vp], i32), "kzg_mlp_collect": ([vp, u32], i32),
            "kzg_mlp_num_samples": ([vp, vp, vp], i32), "kzg_mlp_get_samples": ([vp, u32, u32, vp], i32),
            "kzg_mlp_fit_normalization": ([vp], i32), "kzg_mlp_set_target_clip": ([vp, f32], i32), "kzg_set_learned_features": ([vp, u32], i32), "kzg_mlp_train": ([vp, u32, u32, f32, u32, vp], i32),
            "kzg_mlp_forward": ([vp, vp, u32, vp], i32), "kzg_mlp_gradients": ([vp, vp, u32, vp, vp], i32),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def pack_hulls(shapes: np.ndarray, hulls: dict):
    """Built hulls {shape index: (verts (V,3), tris (T,3))} -> ranges + pools."""
    ranges = np.zeros(len(shapes), HULL_RANGE_DTYPE)
    vs, ts = [np.zeros((0, 3), np.float32)], [np.zeros((0, 3), np.uint32)]
    nv = nt = 0
    for i, (v, t) in sorted(hulls.items()):
        ranges[i] = (nv, len(v), nt, len(t))
        vs.append(np.asarray(v, np.float32))
        ts.append(np.asarray(t, np.uint32))
        nv += len(v)
        nt += len(t)
    verts = np.ascontiguousarray(np.concatenate(vs), np.float32) if nv else np.zeros((1, 3), np.float32)
    tris = np.ascontiguousarray(np.concatenate(ts), np.uint32) if nt else np.zeros((1, 3), np.uint32)
    return ranges, verts, tris, nv, nt


def device_bodies(bodies: np.ndarray, mass: np.ndarray) -> np.ndarray:
    """kz_body_desc table + mass data (inverse mass, inverse local inertia) -> kzg_body table."""
    out = np.zeros(len(bodies), KZG_BODY_DTYPE)
    out["packed_id"] = np.arange(len(bodies), dtype=np.uint32)
    for f in ("shape", "motion_type", "is_sensor", "position", "rotation", "linear_velocity", "angular_velocity",
              "friction", "restitution", "linear_damping", "angular_damping", "gravity_factor",
              "max_linear_velocity", "max_angular_velocity", "allow_sleeping"):
        out[f] = bodies[f]
    out["inverse_mass"] = mass[:, 0]
    out["inverse_inertia_local"] = mass[:, 1:10]
    return out


def schedule(dynamic: np.ndarray, info: np.ndarray, use_simd: bool = True, max_groups: int = 1024,
             units: bool = False, with_units: bool = False, defer: bool = False):
    """Host-only solve schedule (kzg_schedule): per-point velocity level,
    warm-start level, group, and (islands, groups, vel depth, ws depth, vel
    units, ws units).  units=True joins points into multi-point units as the
    device solve does; with_units=True also returns the per-point velocity and
    warm-start unit numbers.  defer=True builds it the way an early warm
    start does (the velocity order finished on the helper thread)."""
    dyn = np.ascontiguousarray(dynamic, np.uint8)
    info = np.ascontiguousarray(info, np.uint32).reshape(-1, 4)
    npts = int((info[:, 3] & 0xFF).sum()) if len(info) else 0
    vl, wl, gr, vu, wu = (np.zeros(max(npts, 1), np.uint32) for _ in range(5))
    st = np.zeros(6, np.uint32)
    flags = (1 if use_simd else 0) | (2 if units else 0) | (4 if defer else 0)
    rc = library().kzg_schedule(len(dyn), _p(dyn), len(info), _p(info), flags, max_groups, _p(vl), _p(wl), _p(gr),
                                _p(vu), _p(wu), _p(st))
    if rc != 0:
        raise KzgError(f"kzg_schedule failed ({rc})")
    if with_units:
        return vl[:npts], wl[:npts], gr[:npts], st, vu[:npts], wu[:npts]
    return vl[:npts], wl[:npts], gr[:npts], st


class DeviceContext:
    """A kzg_ctx: the device contact path for one world."""

    def __init__(self, world: np.ndarray, device: int = 0):
        L = library()
        self._world = np.array(world, dtype=S.WORLD_DTYPE)
        h = C.c_void_p()
        rc = L.kzg_create(C.byref(h), _p(self._world), device)
        if rc != 0:
            raise KzgError(f"kzg_create failed ({rc}): no usable GPU")
        self.h = h
        self.num_bodies = 0
        self.num_pairs = 0

    @classmethod
    def borrow(cls, handle, num_bodies: int) -> "DeviceContext":
        """Wrap a context owned elsewhere (a PhysicsWorld's); never destroyed here."""
        self = cls.__new__(cls)
        self.h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle
        self._borrowed = True
        self.num_bodies = num_bodies
        self.num_pairs = 0
        return self

    def close(self):
        if getattr(self, "h", None) and not getattr(self, "_borrowed", False):
            library().kzg_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str):
        if rc < 0:
            msg = library().kzg_last_error(self.h)
            raise KzgError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
        return rc

    def upload_shapes(self, shapes: np.ndarray, hulls: dict):
        shapes = np.ascontiguousarray(shapes)
        ranges, verts, tris, nv, nt = pack_hulls(shapes, hulls)
        self._check(library().kzg_upload_shapes(self.h, _p(shapes), len(shapes), _p(ranges), _p(verts), nv,
                                                _p(tris), nt), "kzg_upload_shapes")

    def upload_bodies(self, bodies: np.ndarray):
        bodies = np.ascontiguousarray(bodies, dtype=KZG_BODY_DTYPE)
        self._check(library().kzg_upload_bodies(self.h, _p(bodies), len(bodies)), "kzg_upload_bodies")
        self.num_bodies = len(bodies)

    def upload_states(self, st: np.ndarray):
        st = np.ascontiguousarray(st, dtype=S.STATE_DTYPE)
        self._check(library().kzg_upload_states(self.h, _p(st), len(st)), "kzg_upload_states")

    def narrowphase(self, pairs: np.ndarray):
        pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
        self.num_pairs = len(pairs)
        self._check(library().kzg_narrowphase(self.h, _p(pairs), len(pairs)), "kzg_narrowphase")

    def build_constraints(self, warm_start_mode: int = 0):
        self._check(library().kzg_build_constraints(self.h, warm_start_mode), "kzg_build_constraints")

    def integrate_velocities(self, dt=S.DT):
        self._check(library().kzg_integrate_velocities(self.h, float(dt)), "kzg_integrate_velocities")

    def solve(self, dt=S.DT):
        self._check(library().kzg_solve(self.h, float(dt)), "kzg_solve")

    def integrate_positions(self, dt=S.DT):
        self._check(library().kzg_integrate_positions(self.h, float(dt)), "kzg_integrate_positions")

    def update_sleeping(self, dt=S.DT):
        """UpdateSleeping (a no-op unless the world allows sleeping)."""
        self._check(library().kzg_update_sleeping(self.h, float(dt)), "kzg_update_sleeping")

    def active_bodies(self) -> np.ndarray:
        """The active body list in order (body slots)."""
        n = C.c_uint32(0)
        self._check(library().kzg_get_active_bodies(self.h, None, 0, C.byref(n)), "kzg_get_active_bodies")
        out = np.zeros(max(n.value, 1), np.uint32)
        self._check(library().kzg_get_active_bodies(self.h, _p(out), n.value, C.byref(n)), "kzg_get_active_bodies")
        return out[:n.value]

    def set_active_bodies(self, bodies):
        lst = np.ascontiguousarray(bodies, np.uint32)
        self._check(library().kzg_set_active_bodies(self.h, _p(lst) if len(lst) else None, len(lst)),
                    "kzg_set_active_bodies")

    def set_sleep_thresholds(self, time=0.5, linear=0.05, angular=0.1):
        self._check(library().kzg_set_sleep_thresholds(self.h, float(time), float(linear), float(angular)),
                    "kzg_set_sleep_thresholds")

    def step(self, pairs: np.ndarray, dt=S.DT, stop_after: int = 5, warm_start_mode: int = 0):
        """StepFixed's contact path over `pairs`; stop_after as in the oracle."""
        self.narrowphase(pairs)
        if stop_after >= 2:
            self.build_constraints(warm_start_mode)
        if stop_after >= 3:
            self.integrate_velocities(dt)
        if stop_after >= 4:
            self.solve(dt)
        if stop_after >= 5:
            self.integrate_positions(dt)
            self.update_sleeping(dt)

    def sizes(self):
        a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
        self._check(library().kzg_query_sizes(self.h, C.byref(a), C.byref(b), C.byref(c)), "kzg_query_sizes")
        return a.value, b.value, c.value

    def states(self) -> np.ndarray:
        out = np.zeros(self.num_bodies, S.STATE_DTYPE)
        self._check(library().kzg_download_states(self.h, _p(out), self.num_bodies), "kzg_download_states")
        return out

    def manifolds(self):
        n = self.num_pairs
        counts = np.zeros(max(n, 1), np.uint32)
        pts = np.zeros((max(n, 1), 4), S.CONTACT_DTYPE)
        if n:
            self._check(library().kzg_download_manifolds(self.h, _p(counts), _p(pts)), "kzg_download_manifolds")
        return counts[:n], pts[:n]

    def constraints(self):
        _, nc, _ = self.sizes()
        hdr = np.zeros((max(nc, 1), 4), np.uint32)
        pts = np.zeros((max(nc, 1), 8, POINT_FLOATS), np.float32)
        if nc:
            self._check(library().kzg_download_constraints(self.h, _p(hdr), _p(pts)), "kzg_download_constraints")
        return hdr[:nc], pts[:nc]

    def residuals(self) -> np.ndarray:
        out = np.zeros(64, np.float64)
        k = self._check(library().kzg_get_residuals(self.h, _p(out), 64), "kzg_get_residuals")
        return out[:k].copy()

    def natural_residual(self) -> dict:
        """Natural-map residual of the iterate the last solve left (before
        integrate_positions): RMS of normal + friction, normal, friction and
        cone terms (N s)."""
        out = np.zeros(4, np.float64)
        self._check(library().kzg_get_natural_residual(self.h, _p(out)), "kzg_get_natural_residual")
        return {"total": float(out[0]), "normal": float(out[1]), "friction": float(out[2]), "cone": float(out[3])}

    def timings(self) -> dict:
        t = np.zeros((), TIMINGS_DTYPE)
        self._check(library().kzg_get_timings(self.h, _p(t)), "kzg_get_timings")
        return {k: t[k].item() for k in TIMINGS_DTYPE.names}

    # ---- warm-start mode, iterations, learned warm start --------------------
    def set_warm_start(self, mode: int):
        self._check(library().kzg_set_warm_start(self.h, int(mode)), "kzg_set_warm_start")

    def set_iterations(self, velocity: int, position: int):
        self._check(library().kzg_set_iterations(self.h, int(velocity), int(position)), "kzg_set_iterations")

    def set_cooperative(self, blocks: int):
        """Workgroups sharing the largest group's levels: 0 auto, 1 off."""
        self._check(library().kzg_set_cooperative(self.h, int(blocks)), "kzg_set_cooperative")

    def set_sweep_units(self, velocity: int, warm_start: int | None = None):
        """Most points of one body pair a sweep thread applies in a row, in the
        velocity and the warm-start / bias sweeps: 0 default, 1..4."""
        ws = velocity if warm_start is None else warm_start
        self._check(library().kzg_set_sweep_units(self.h, int(velocity), int(ws)), "kzg_set_sweep_units")

    def set_narrowphase_waves(self, waves: int):
        """Waves of the EPA / clipping kernel: 0 auto (16 per CU)."""
        self._check(library().kzg_set_narrowphase_waves(self.h, int(waves)), "kzg_set_narrowphase_waves")

    def set_cooperative_placement(self, placement: int):
        """COOP_ONE_XCD (0, default) or COOP_SPREAD (1)."""
        self._check(library().kzg_set_cooperative_placement(self.h, int(placement)), "kzg_set_cooperative_placement")

    def set_device_bucketing(self, mode: int):
        """0 (default): an early warm start's velocity schedule is bucketed
        by level on the device; 1: on the host."""
        self._check(library().kzg_set_device_bucketing(self.h, int(mode)), "kzg_set_device_bucketing")

    def set_spin_limit(self, polls: int):
        """Most polls of one wait of a cooperative sweep (0 default)."""
        self._check(library().kzg_set_spin_limit(self.h, int(polls)), "kzg_set_spin_limit")

    def set_island_lds_limit(self, max_bodies: int):
        """Most bodies for the one-workgroup LDS union-find (0 default, 1:
        always the global-memory union-find)."""
        self._check(library().kzg_set_island_lds_limit(self.h, int(max_bodies)), "kzg_set_island_lds_limit")

    def set_early_warm_start(self, mode: int):
        """Sweep the warm start while the host finishes the velocity schedule:
        0 auto (16,384 or more points), 1 off, 2 always."""
        self._check(library().kzg_set_early_warm_start(self.h, int(mode)), "kzg_set_early_warm_start")

    def mlp_init(self, seed: int = 0x4B5A0005):
        self._check(library().kzg_mlp_init(self.h, C.c_uint64(seed)), "kzg_mlp_init")

    def mlp_params(self) -> np.ndarray:
        out = np.zeros(MLP_PARAMS, np.float32)
        self._check(library().kzg_mlp_get_params(self.h, _p(out), MLP_PARAMS), "kzg_mlp_get_params")
        return out

    def mlp_set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        self._check(library().kzg_mlp_set_params(self.h, _p(p), len(p)), "kzg_mlp_set_params")

    def mlp_normalization(self):
        mean, istd = np.zeros(MLP_FEATURES, np.float32), np.zeros(MLP_FEATURES, np.float32)
        self._check(library().kzg_mlp_get_normalization(self.h, _p(mean), _p(istd)), "kzg_mlp_get_normalization")
        return mean, istd

    def mlp_set_normalization(self, mean, inv_std):
        m = np.ascontiguousarray(mean, np.float32)
        s = np.ascontiguousarray(inv_std, np.float32)
        assert len(m) == len(s) == MLP_FEATURES
        self._check(library().kzg_mlp_set_normalization(self.h, _p(m), _p(s)), "kzg_mlp_set_normalization")

    def mlp_collect(self, capacity: int):
        self._check(library().kzg_mlp_collect(self.h, int(capacity)), "kzg_mlp_collect")

    def mlp_num_samples(self):
        w, h = C.c_uint64(), C.c_uint32()
        self._check(library().kzg_mlp_num_samples(self.h, C.byref(w), C.byref(h)), "kzg_mlp_num_samples")
        return w.value, h.value

    def mlp_samples(self, first: int = 0, count: int | None = None) -> np.ndarray:
        if count is None:
            count = self.mlp_num_samples()[1] - first
        out = np.zeros((max(count, 1), MLP_SAMPLE_ROW), np.float32)
        self._check(library().kzg_mlp_get_samples(self.h, int(first), int(count), _p(out)), "kzg_mlp_get_samples")
        return out[:count]

    def mlp_fit_normalization(self):
        self._check(library().kzg_mlp_fit_normalization(self.h), "kzg_mlp_fit_normalization")

    def mlp_train(self, steps: int, batch: int = 4096, lr: float = 1e-3, seed: int = 1) -> np.ndarray:
        losses = np.zeros(max(steps, 1), np.float32)
        self._check(library().kzg_mlp_train(self.h, int(steps), int(batch), float(lr), int(seed) & 0xFFFFFFFF,
                                            _p(losses)), "kzg_mlp_train")
        return losses[:steps]

    def cache_image(self):
        """(image bytes, capacity, used) of the warm-start cache."""
        cap, used = C.c_uint32(), C.c_uint32()
        self._check(library().kzg_cache_info(self.h, C.byref(cap), C.byref(used)), "kzg_cache_info")
        img = np.zeros(max(cap.value, 1) * 108, np.uint8)
        self._check(library().kzg_download_cache(self.h, _p(img), cap.value * 108), "kzg_download_cache")
        return img[:cap.value * 108], cap.value, used.value

    def set_cache_image(self, image):
        img, cap, used = image
        self._check(library().kzg_upload_cache(self.h, _p(img), len(img), cap, used), "kzg_upload_cache")

    def mlp_set_target_clip(self, clip: float = 0.0):
        """Clamp training targets to [-clip, clip] (0: off) in mlp_train and mlp_gradients."""
        self._check(library().kzg_mlp_set_target_clip(self.h, float(clip)), "kzg_mlp_set_target_clip")

    def set_learned_features(self, flags: int = 0):
        """FEATURES_LOAD (1): features 14-15 carry the loads of the point's
        bodies (summed cached normal impulses in body weights); 0: zero."""
        self._check(library().kzg_set_learned_features(self.h, int(flags)), "kzg_set_learned_features")

    def mlp_forward(self, features) -> np.ndarray:
        f = np.ascontiguousarray(features, np.float32).reshape(-1, MLP_FEATURE_ROW)
        out = np.zeros((max(len(f), 1), 4), np.float32)
        self._check(library().kzg_mlp_forward(self.h, _p(f), len(f), _p(out)), "kzg_mlp_forward")
        return out[:len(f)]

    def mlp_gradients(self, samples):
        s = np.ascontiguousarray(samples, np.float32).reshape(-1, MLP_SAMPLE_ROW)
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = C.c_float()
        self._check(library().kzg_mlp_gradients(self.h, _p(s), len(s), _p(g), C.byref(loss)), "kzg_mlp_gradients")
        return g, loss.value

    def collide_batch(self, shape_a, shape_b, xf_a, xf_b, max_separation=0.02):
        sa = np.ascontiguousarray(shape_a, np.uint32)
        sb = np.ascontiguousarray(shape_b, np.uint32)
        xa = np.ascontiguousarray(xf_a, np.float32)
        xb = np.ascontiguousarray(xf_b, np.float32)
        n = len(sa)
        counts = np.zeros(max(n, 1), np.uint32)
        pts = np.zeros((max(n, 1), 4), S.CONTACT_DTYPE)
        ms = C.c_float()
        self.num_pairs = n
        self._check(library().kzg_collide_batch(self.h, _p(sa), _p(sb), _p(xa), _p(xb), n, float(max_separation),
                                                _p(counts), _p(pts), C.byref(ms)), "kzg_collide_batch")
        return counts[:n], pts[:n], ms.value
