"""Shared fixtures: the `gpu` marker, the package loader and the oracle."""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "arap-deformation-of-gaussian-radiance-fields_amd")
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def load_package():
    """Import the package directory under the module name ``arapgs_amd``."""
    if "arapgs_amd" in sys.modules:
        return sys.modules["arapgs_amd"]
    spec = importlib.util.spec_from_file_location("arapgs_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["arapgs_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def ctx():
    import torch
    from arapgs_amd import arapgs
    assert torch.cuda.is_available(), "GPU test without a GPU"
    c = arapgs.Context(0)
    yield c
    c.close()
