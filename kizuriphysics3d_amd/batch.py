"""Batched worlds (include/kzb.h): thousands of independent worlds stepped
together on one GPU, one captured HIP graph per step."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import device as D
from . import scenes as S

CONFIG_DTYPE = np.dtype([("narrowphase_waves", "<u4"), ("solve_threads", "<u4"), ("broadphase_threads", "<u4"),
                         ("flags", "<u4")])
STATS_DTYPE = np.dtype([
    ("num_worlds", "<u4"), ("num_bodies", "<u4"), ("num_pairs", "<u4"), ("num_constraints", "<u4"),
    ("num_points", "<u4"), ("max_velocity_levels", "<u4"), ("max_bias_levels", "<u4"), ("reserved", "<u4"),
    ("broadphase_ms", "<f8"), ("device_ms", "<f8"), ("step_ms", "<f8"), ("device_with_copies_ms", "<f8"),
    ("launch_ms", "<f8"), ("phase_ms", "<f4", (6,)),
])
DIRECT_LAUNCH = 1  # kzb_config.flags: kernels one by one with phase events, no graph
KZB_EXPORTS = ["kzb_create", "kzb_destroy", "kzb_last_error", "kzb_add_world", "kzb_step", "kzb_num_worlds",
               "kzb_get_states", "kzb_get_stats", "kzb_set_config", "kzb_get_config", "kzb_autotune",
               "kzb_autotune_candidates", "kzb_world_schedule"]

_bound = False


def _lib():
    global _bound
    L = D.library()
    if not _bound:
        vp, u32, f32, i32 = C.c_void_p, C.c_uint32, C.c_float, C.c_int
        sig = {
            "kzb_create": ([vp, vp, i32], i32), "kzb_destroy": ([vp], None), "kzb_last_error": ([vp], C.c_char_p),
            "kzb_add_world": ([vp, vp, u32, vp, u32, vp, u32], i32), "kzb_step": ([vp, f32], i32),
            "kzb_num_worlds": ([vp], i32), "kzb_get_states": ([vp, u32, vp, u32], i32),
            "kzb_get_stats": ([vp, vp], i32), "kzb_set_config": ([vp, vp], i32), "kzb_get_config": ([vp, vp], i32),
            "kzb_autotune": ([vp, f32, u32, vp, vp, u32], i32), "kzb_autotune_candidates": ([vp, vp, u32], i32),
            "kzb_world_schedule": ([vp, u32, vp, vp, u32], i32),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _bound = True
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class BatchedWorlds:
    """Independent worlds (scenes.Scene each, one shared world setting) on one GPU."""

    def __init__(self, scenes: list, device: int = 0):
        L = _lib()
        self._settings = np.array(scenes[0].world, dtype=S.WORLD_DTYPE)
        h = C.c_void_p()
        rc = L.kzb_create(C.byref(h), _p(self._settings), device)
        if rc != 0:
            raise D.KzgError(f"kzb_create failed ({rc})" + (": no usable GPU" if rc == -2 else ""))
        self.h = h
        self.sizes = []
        for sc in scenes:
            shapes = sc.shape_array()
            hull = np.ascontiguousarray(sc.hull_array(), np.float32)
            bodies = sc.body_array()
            self._check(L.kzb_add_world(self.h, _p(shapes), len(shapes), _p(hull), len(hull), _p(bodies),
                                        len(bodies)), "kzb_add_world")
            self.sizes.append(len(bodies))

    def close(self):
        if getattr(self, "h", None):
            _lib().kzb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc < 0:
            msg = _lib().kzb_last_error(self.h)
            raise D.KzgError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
        return rc

    def step(self, dt=S.DT):
        self._check(_lib().kzb_step(self.h, float(dt)), "kzb_step")

    def states(self, world: int) -> np.ndarray:
        out = np.zeros(self.sizes[world], S.STATE_DTYPE)
        self._check(_lib().kzb_get_states(self.h, world, _p(out), len(out)), "kzb_get_states")
        return out

    def stats(self) -> dict:
        s = np.zeros((), STATS_DTYPE)
        self._check(_lib().kzb_get_stats(self.h, _p(s)), "kzb_get_stats")
        return {k: s[k].tolist() for k in STATS_DTYPE.names}

    def config(self) -> dict:
        c = np.zeros((), CONFIG_DTYPE)
        self._check(_lib().kzb_get_config(self.h, _p(c)), "kzb_get_config")
        return {k: c[k].item() for k in CONFIG_DTYPE.names}

    def set_config(self, narrowphase_waves=0, solve_threads=0, broadphase_threads=0, flags=0):
        c = np.zeros((), CONFIG_DTYPE)
        c["narrowphase_waves"], c["solve_threads"], c["broadphase_threads"], c["flags"] = narrowphase_waves, \
            solve_threads, broadphase_threads, flags
        self._check(_lib().kzb_set_config(self.h, _p(c)), "kzb_set_config")

    def candidates(self) -> list:
        n = self._check(_lib().kzb_autotune_candidates(self.h, None, 0), "kzb_autotune_candidates")
        c = np.zeros(n, CONFIG_DTYPE)
        _lib().kzb_autotune_candidates(self.h, _p(c), n)
        return [{k: r[k].item() for k in CONFIG_DTYPE.names} for r in c]

    def autotune(self, repeats: int = 3, dt=S.DT):
        """Steps len(candidates) * repeats times; returns (best index, medians in ms)."""
        n = len(self.candidates())
        times = np.zeros(n, np.float32)
        best = np.zeros((), CONFIG_DTYPE)
        bi = self._check(_lib().kzb_autotune(self.h, float(dt), repeats, _p(best), _p(times), n), "kzb_autotune")
        return bi, times

    def world_schedule(self, world: int):
        n = self._check(_lib().kzb_world_schedule(self.h, world, None, None, 0), "kzb_world_schedule")
        vl = np.zeros(max(n, 1), np.uint32)
        wl = np.zeros(max(n, 1), np.uint32)
        _lib().kzb_world_schedule(self.h, world, _p(vl), _p(wl), n)
        return vl[:n], wl[:n]
