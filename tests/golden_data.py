"""Loaders for the golden fixtures in tests/golden/ (data written from the
reference engine by tests/golden/make_fixtures.py)."""
from __future__ import annotations

import hashlib
import os

import numpy as np

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# Scene parameters the fixtures were generated with (make_fixtures.py).
PILE_SEED, PILE_N = 0x4B5A0202, 256
BATCH_SEED, BATCH_N, BATCH_SHAPES = 0x4B5A0303, 4096, 512
PLANE_SEED, PLANE_N, PLANE_SHAPES = 0x4B5A0107, 2048, 256

_cache = {}


def load_fixture(name: str):
    if name not in _cache:
        _cache[name] = dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))
    return _cache[name]


def state_digest(st: np.ndarray) -> np.uint64:
    return np.frombuffer(hashlib.sha256(st.tobytes()).digest()[:8], np.uint64)[0]


def fixture_pairs(z, step: int) -> np.ndarray:
    o = z["pair_offsets"]
    return z["pairs"][o[step]: o[step + 1]].astype(np.uint32)


def fixture_hulls(z) -> dict:
    out = {}
    for i, (vo, vc, to, tc) in enumerate(z["hull_refs"]):
        if vc:
            out[i] = (z["hull_verts"][vo: vo + vc], z["hull_tris"][to: to + tc])
    return out


def fixture_derived(z, with_mass: bool = True) -> O.Derived:
    refs = z["hull_refs"]
    shapes_stub = np.zeros(len(refs))
    mass = z["mass"][:, :10] if with_mass else None
    return O.Derived.pack(shapes_stub, fixture_hulls(z), mass)
