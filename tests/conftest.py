import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

# the autotuner's plan file: the test session's own, not the user's cache
if "FFCZ_PLAN_CACHE" not in os.environ:
    import tempfile

    os.environ["FFCZ_PLAN_CACHE"] = str(Path(tempfile.mkdtemp(prefix="ffcz-plans-")) / "plans.txt")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libffcz.so")


@pytest.fixture(scope="session")
def gpu():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test selected but no HIP device is visible")
    import ffcz_amd

    ffcz_amd.load()  # raises if the HIP library is missing: no fallback
    return torch.device("cuda:0")
