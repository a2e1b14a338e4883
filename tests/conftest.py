"""Shared test setup: the `gpu` marker, and the package (whose directory name,
rivet-vp9_amd, is not an identifier) importable as `rivet_vp9_amd`."""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def _load_package():
    if "rivet_vp9_amd" in sys.modules:
        return sys.modules["rivet_vp9_amd"]
    path = os.path.join(ROOT, "rivet-vp9_amd", "__init__.py")
    spec = importlib.util.spec_from_file_location("rivet_vp9_amd", path,
                                                  submodule_search_locations=[os.path.dirname(path)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["rivet_vp9_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


_load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")
