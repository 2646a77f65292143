import importlib.util
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
PKG_DIR = ROOT / "bethe-salpeter-equation_amd"


def _load_pkg():
    if "bse_amd" in sys.modules:
        return sys.modules["bse_amd"]
    spec = importlib.util.spec_from_file_location(
        "bse_amd", PKG_DIR / "__init__.py", submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["bse_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


_load_pkg()
# The oracle is the checker: tests import it by path.
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def bse_amd():
    return _load_pkg()
