"""Shared test set-up: the `gpu` marker, the package loader and the oracle."""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libdiffmom.so on cuda:0)")


def load_package():
    """Import diffmom.jl_amd (a directory name that is not an identifier) as diffmom_jl_amd."""
    name = "diffmom_jl_amd"
    if name in sys.modules:
        return sys.modules[name]
    pkg_dir = os.path.join(ROOT, "diffmom.jl_amd")
    spec = importlib.util.spec_from_file_location(name, os.path.join(pkg_dir, "__init__.py"),
                                                  submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def dm():
    return load_package()


@pytest.fixture(scope="session")
def O():
    from oracle import oracle
    return oracle


def oracle_rwg(O, mesh, rwg):
    """Oracle RWG over the same mesh arrays and basis order as a package RWGData."""
    om = O.Mesh(mesh.xyz, mesh.tri)
    return O.RWG(om, rwg.tplus, rwg.tminus, rwg.evert, rwg.vplus_opp, rwg.vminus_opp, rwg.len,
                 rwg.area_plus, rwg.area_minus, rwg.coeff_plus, rwg.coeff_minus)


def rel_fro(a, b):
    import numpy as np
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
