import importlib.util
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "tests"))
sys.path.insert(0, REPO)


def load_pkg():
    """Import master-of-puppets_amd/ as the package `mop_amd`."""
    if "mop_amd" in sys.modules:
        return sys.modules["mop_amd"]
    pkg_dir = os.path.join(REPO, "master-of-puppets_amd")
    spec = importlib.util.spec_from_file_location(
        "mop_amd", os.path.join(pkg_dir, "__init__.py"), submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mop_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


load_pkg()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")
