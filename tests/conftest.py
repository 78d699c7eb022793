import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
if HERE not in sys.path:
    sys.path.insert(0, HERE)
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def orc():
    import oracle_lib
    return oracle_lib.load()


def load_package():
    """Import micsa-ea-mc_amd/ (hyphenated directory) as module micsa_ea_mc_amd."""
    import importlib.util
    if "micsa_ea_mc_amd" in sys.modules:
        return sys.modules["micsa_ea_mc_amd"]
    path = os.path.join(REPO, "micsa-ea-mc_amd", "__init__.py")
    spec = importlib.util.spec_from_file_location(
        "micsa_ea_mc_amd", path, submodule_search_locations=[os.path.dirname(path)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["micsa_ea_mc_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def pkg():
    return load_package()
