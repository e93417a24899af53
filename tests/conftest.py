import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("oracle", "cosmic-fields-lite_amd", ""):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: CPU test that takes tens of seconds")


@pytest.fixture(scope="session")
def O():
    import cf_oracle
    return cf_oracle


@pytest.fixture(scope="session")
def cfl():
    import torch
    if not torch.cuda.is_available():
        # GPU tests are selected with -m gpu on a GPU machine; not seeing the
        # device there is a failure, not a reason to skip.
        pytest.fail("gpu test selected but no GPU is visible to torch")
    import cfl_amd
    return cfl_amd
