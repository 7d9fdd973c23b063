import json
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device; runs through the HIP C ABI")
    # stores in tests start from the built-in defaults, so each test runs the
    # executor path it names; tests of tuning files pass them explicitly
    os.environ.setdefault("SPARKALOG_HIP_TUNING", "")


@pytest.fixture(scope="session")
def golden():
    return json.loads((GOLDEN / "reference_vectors.json").read_text())


@pytest.fixture(scope="session")
def expected():
    return json.loads((GOLDEN / "workload_expected.json").read_text())


@pytest.fixture(scope="session")
def stream():
    import torch

    from sparkalog_amd.execution import HipStream

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.init()
    return HipStream()
