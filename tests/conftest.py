import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def pkg():
    import helpers
    return helpers.load_pkg()


@pytest.fixture(scope="session")
def solver(pkg):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    cfg = pkg.default_config(max_knots=512, max_instances=64, max_total_knots=8192)
    s = pkg.Solver(cfg)
    yield s
    s.close()
