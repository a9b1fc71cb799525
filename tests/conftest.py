import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import repeat_oracle as O
    if not os.path.exists(O._LIB_PATH):
        O.build()
    return O


@pytest.fixture(scope="session")
def gpu_ctx():
    from pyrepeat_amd import _abi
    ctx = _abi.Context(0, 0)
    yield ctx
    ctx.close()


def golden(name):
    import json
    with open(os.path.join(ROOT, "tests", "golden", f"{name}.json")) as f:
        return json.load(f)
