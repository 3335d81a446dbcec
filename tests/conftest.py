import json
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def fixtures():
    return json.loads((GOLDEN / "reference_fixtures.json").read_text())


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    pyoracle.lib()
    return pyoracle


def ref_index_files():
    return sorted(GOLDEN.glob("ref_index_*.npz"))


@pytest.fixture(scope="session")
def ctx():
    import torch
    import simit_amd
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = simit_amd.Context(0)
    yield c
    c.close()


def to_dev(a, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()
