import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")
    config.addinivalue_line("markers", "reference: needs the reference build oracle/_ref/libkzref.so")


def pytest_collection_modifyitems(config, items):
    from oracle import refharness
    if refharness.available():
        return
    skip = pytest.mark.skip(reason="oracle/_ref/libkzref.so is not built (make -C oracle ref)")
    for item in items:
        if "reference" in item.keywords:
            item.add_marker(skip)
