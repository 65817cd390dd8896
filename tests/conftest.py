"""Shared test helpers: the `gpu` marker, the product package and the oracle.

The oracle (oracle/liboracle.so) is the checker only; the product package is
loaded from fherma-polymul-baseline_amd/ and calls the HIP library.
"""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "fherma-polymul-baseline_amd")
sys.path.insert(0, os.path.join(ROOT, "tests"))


def load_package():
    name = "polymul_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def pm():
    return load_package()


@pytest.fixture(scope="session")
def workload(pm):
    return pm.workload


@pytest.fixture(scope="session")
def orc():
    import oracle_lib
    return oracle_lib.Oracle()
