import json
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "cuda-cube-solver_amd"))
sys.path.insert(0, str(ROOT / "oracle"))
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def index_vectors():
    return json.loads((GOLDEN / "index_vectors.json").read_text())


@pytest.fixture(scope="session")
def reference_results():
    return json.loads((GOLDEN / "reference_results.json").read_text())


@pytest.fixture(scope="session")
def table_hashes():
    return json.loads((GOLDEN / "table_sha256.json").read_text())
