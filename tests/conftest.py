import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
EXAMPLE = os.path.join(GOLDEN, "example")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def example():
    """The reference's minimal reproducible example (user_parameters:4-145):
    J=2, r=4, M=5, user knots + parent file, chordal Wendland fields."""
    from tests.cases import example_case
    return example_case()
