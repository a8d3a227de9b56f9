import importlib.util
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)
ORACLE_DIR = os.path.join(REPO, "oracle")
if ORACLE_DIR not in sys.path:
    sys.path.insert(0, ORACLE_DIR)
REFERENCE = "/root/reference"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def load_package():
    """Import the rustmpeg-next_amd package (its directory name has a hyphen)."""
    name = "rustmpeg_next_amd"
    if name in sys.modules:
        return sys.modules[name]
    path = os.path.join(REPO, "rustmpeg-next_amd", "__init__.py")
    spec = importlib.util.spec_from_file_location(name, path,
                                                  submodule_search_locations=[os.path.dirname(path)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def po():
    import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def rmh():
    return load_package()


def reference_text(rel):
    path = os.path.join(REFERENCE, rel)
    if not os.path.exists(path):
        pytest.skip("reference source not present")
    with open(path) as f:
        return f.read()


@pytest.fixture(scope="session")
def acc(rmh):
    """A device context; GPU tests fail (not skip) when it cannot be created."""
    a = rmh.Accelerator(0)
    yield a
    a.close()
