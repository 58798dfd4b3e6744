import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")
PTNI_CE = os.path.join(GOLDEN, "icet_c8_6_5.ce")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (ROCm GPU) and the built libneuralce_hip.so")


@pytest.fixture(scope="session")
def ptni_ce():
    from neuralce2_amd.ce_io import ClusterExpansion
    return ClusterExpansion.read(PTNI_CE)


@pytest.fixture(scope="session")
def ptni_template(ptni_ce):
    from neuralce2_amd.ce.ce_gpu_implicit import CETemplate
    return CETemplate(ptni_ce.cluster_space, ptni_ce.parameters)


def random_fcc(n, x=0.5, seed=1, a=3.924):
    """random Ni(1-x)Pt(x) on an n x n x n conventional fcc cell (the reference's test fixture for n = 4,
    tests/conftest.py:23-33, uses RandomState(1) and x = 0.5)."""
    from neuralce2_amd.lattice import bulk_fcc
    at = bulk_fcc("Ni", a, cubic=True).repeat((n, n, n))
    r = np.random.RandomState(seed)
    Z = np.full(len(at), 28, np.int64)
    Z[r.choice(len(at), int(round(len(at) * x)), replace=False)] = 78
    at.numbers = Z
    return at


@pytest.fixture(scope="session")
def ptni_atoms_256():
    """random Ni0.5Pt0.5 on a 4x4x4 conventional fcc cell (256 atoms), as the reference's fixture."""
    from neuralce2_amd.lattice import bulk_fcc
    a = bulk_fcc("Ni", 3.924, cubic=True).repeat((4, 4, 4))
    r = np.random.RandomState(1)
    Z = np.full(len(a), 78, np.int64)
    Z[r.choice(len(a), len(a) // 2, replace=False)] = 28
    a.numbers = Z
    return a


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
