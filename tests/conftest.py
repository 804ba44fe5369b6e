import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)
GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and libpmw.so")


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def rel_l2(a, b):
    """||a - b|| / ||b|| over the arrays given (inf if b is zero and a is not)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    nd = np.linalg.norm(a - b)
    if nb == 0.0:
        return 0.0 if nd == 0.0 else np.inf
    return nd / nb


def interior(s, hs=2):
    return s[..., hs:-hs, hs:-hs]


@pytest.fixture
def golden():
    return load_golden
