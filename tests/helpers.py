"""Shared test helpers: derived scene data from the reference build and the
product, and comparison utilities."""
from __future__ import annotations

import numpy as np

from kizuriphysics3d_amd import scenes as S


def reference_hulls(shapes: np.ndarray, hull_points: np.ndarray) -> dict:
    """Built hulls per hull shape, from the reference's ConvexHullShape::Create."""
    from oracle import refharness as R
    out = {}
    for i, s in enumerate(shapes):
        if s["type"] == S.SHAPE_HULL:
            h = R.hull_create(hull_points[s["hull_offset"]: s["hull_offset"] + s["hull_count"]])
            assert h is not None, f"reference rejected hull {i}"
            out[i] = h
    return out


def first_mismatch(a: np.ndarray, b: np.ndarray):
    """Index of the first record whose bytes differ, or None."""
    av = a.view(np.uint8).reshape(len(a), -1)
    bv = b.view(np.uint8).reshape(len(b), -1)
    bad = np.nonzero((av != bv).any(1))[0]
    return None if len(bad) == 0 else int(bad[0])


def bits(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a).view(np.uint8)
