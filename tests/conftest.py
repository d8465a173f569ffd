import importlib
import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
if str(ROOT / "tests") not in sys.path:
    sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def wf():
    return importlib.import_module("wavefront-renderer_amd")


@pytest.fixture(scope="session")
def scenes():
    return importlib.import_module("wavefront-renderer_amd.scenes")


@pytest.fixture(scope="session")
def orc():
    import oracle_lib
    return oracle_lib.Oracle()


@pytest.fixture(scope="session")
def ref():
    import oracle_lib
    r = oracle_lib.reference()
    if r is None:
        pytest.skip("oracle/_ref/libref.so not built (reference tree absent)")
    return r


@pytest.fixture(scope="session")
def ctx(wf):
    c = wf.Context(0)
    yield c
    c.close()


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
