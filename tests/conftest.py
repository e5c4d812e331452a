import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "hybrid-aphi-solver_amd"))
sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests proper")


def golden(name):
    """(meta, arrays) of a fixture written by tests/golden/make_golden.py."""
    meta = json.load(open(os.path.join(GOLDEN, name + ".json")))
    path = os.path.join(GOLDEN, name + ".npz")
    arrays = dict(np.load(path)) if os.path.exists(path) else {}
    return meta, arrays


def conditioning(name):
    """Condition numbers of a golden numeric case (tests/golden/make_conditioning.py)."""
    return json.load(open(os.path.join(GOLDEN, "conditioning.json")))[name]


def cube_config(meta):
    import aphi_amd as A
    # mode 3 fixtures drive the z = side port with a 1 A boundary current
    return A.CubeConfig(n=meta["n"], side=meta["side"], bar=tuple(meta["bar"]), sigma=meta["sigma"],
                        i_top=1.0 if meta.get("mode") == 3 else None)


@pytest.fixture(scope="session")
def gpu_context():
    import torch
    import aphi_amd as A
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    ctx = A.Context(0)
    yield ctx
    ctx.close()
