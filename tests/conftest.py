import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

GOLDEN = os.path.join(ROOT, "tests", "golden")
INSTANCES = os.path.join(GOLDEN, "instances")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def problem_type_of(name: str) -> int:
    if name.startswith("KP_") or name.startswith("knapsack"):
        return 1
    if name.startswith("bp-"):
        return 2
    if name.startswith("tsp-"):
        return 3
    raise ValueError(name)


def instance_path(name: str) -> str:
    return os.path.join(INSTANCES, name)


@pytest.fixture(scope="session")
def engine():
    from pmdd_amd.engine import Engine
    if Engine.device_count() < 1:
        pytest.fail("no HIP device visible to a gpu-marked test")
    eng = Engine(0)
    yield eng
    eng.close()
