"""Shared test setup: registers the package (directory ``polfed.jl_amd``) under
the import name ``polfed_amd`` and the ``gpu`` marker."""
import importlib.util
import pathlib
import sys

import pytest

REPO = pathlib.Path(__file__).resolve().parent.parent
if str(REPO) not in sys.path:
    sys.path.insert(0, str(REPO))


def _load_package():
    if "polfed_amd" in sys.modules:
        return sys.modules["polfed_amd"]
    pkg_dir = REPO / "polfed.jl_amd"
    spec = importlib.util.spec_from_file_location(
        "polfed_amd", pkg_dir / "__init__.py", submodule_search_locations=[str(pkg_dir)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["polfed_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


_load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X)")


@pytest.fixture(scope="session")
def dev():
    from polfed_amd.device import Device
    return Device(0)


def pytest_collection_modifyitems(config, items):
    """Run CPU tests first; GPU tests keep their file order."""
    items.sort(key=lambda it: 1 if it.get_closest_marker("gpu") else 0)
