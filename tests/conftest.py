import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "diffpumas.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def physics():
    import diffpumas_hip as dp
    return dp.create_physics_csda(("StandardRock", "Air"))


@pytest.fixture(scope="session")
def ctx():
    import diffpumas_hip as dp
    return dp.default_context()
