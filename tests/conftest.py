"""Shared test set-up: the package loader, the oracle and the ``gpu`` marker."""
import importlib.util
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def load_package():
    """Import ``trixiparticles.jl_amd/`` (a directory name with a dot) as
    ``trixiparticles_jl_amd``."""
    name = "trixiparticles_jl_amd"
    if name in sys.modules:
        return sys.modules[name]
    pkg_dir = os.path.join(REPO, "trixiparticles.jl_amd")
    spec = importlib.util.spec_from_file_location(
        name, os.path.join(pkg_dir, "__init__.py"), submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def tp():
    return load_package()


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    return oracle


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
