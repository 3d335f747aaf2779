import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library through the C-ABI)")


def load_fixture_rank1():
    """Inputs of rank 1 of the reference's C test data (mappraiser/test/data_bin),
    rebuilt into MLmap's layout (pix = 3p+k, weights = (1, q, u))."""
    z = np.load(os.path.join(GOLDEN, "fixture_rank1.npz"))
    p = z["p"].astype(np.int64)
    m = p.size
    pix = np.empty(3 * m, np.int32)
    w = np.empty(3 * m, np.float64)
    for k in range(3):
        pix[k::3] = 3 * p + k
    w[0::3] = 1.0
    w[1::3] = z["wq"]
    w[2::3] = z["wu"]
    return dict(pix=pix, weights=w, signal=z["signal"], noise=z["noise"], invtt=z["invtt"],
                block_sizes=z["block_sizes"], lam=int(z["lam"]))


@pytest.fixture(scope="session")
def fixture_rank1():
    return load_fixture_rank1()


@pytest.fixture(scope="session")
def ref_rank1():
    return dict(np.load(os.path.join(GOLDEN, "ref_mapmat_rank1.npz")))


@pytest.fixture(scope="session")
def ref_flags():
    return dict(np.load(os.path.join(GOLDEN, "ref_mapmat_flags.npz")))


@pytest.fixture(scope="session")
def lapack_golden():
    return dict(np.load(os.path.join(GOLDEN, "lapack_3x3.npz")))


@pytest.fixture(scope="session")
def gpu():
    from midapack_amd import lib as G

    G.init(0)
    return G
