"""Shared test set-up: the `gpu` marker, package and oracle loading."""
import importlib.util
import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG_DIR = ROOT / "peregrine-desktop_amd"


def load_package():
    name = "peregrine_desktop_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(
        name, PKG_DIR / "__init__.py", submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def pg():
    return load_package()


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.lib()
    return oracle


def oracle_problem(orc_mod, prob):
    """oracle.Problem from a configs.Problem."""
    from peregrine_desktop_amd.model import ctl_soa
    return orc_mod.Problem(prob.grid.c(), prob.model.c_struct(), prob.params.c(), ctl_soa(prob.ctl),
                           prob.occ)


def nthreads():
    return min(16, os.cpu_count() or 1)
