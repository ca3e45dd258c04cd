"""The C-ABI libraries load and export every function their headers declare
(no compute calls: these run without a GPU)."""
import os
import re

import pytest

from kizuriphysics3d_amd import device as D

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kz[a-z]_[a-z0-9_]+)\s*\(", text)))


def test_kzg_exports_every_declared_function():
    lib = D.library()
    names = _declared("kzg.h")
    assert len(names) >= 19
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(D.EXPORTS)


def test_no_gpu_means_a_loud_error():
    try:
        ctx = D.DeviceContext(__import__("kizuriphysics3d_amd.scenes", fromlist=["x"]).world_settings())
    except D.KzgError as e:
        assert "no usable GPU" in str(e)
        return
    ctx.close()
    pytest.skip("a GPU is present: the no-GPU error path cannot be exercised here")
