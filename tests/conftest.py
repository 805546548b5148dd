import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (parity tests through the C ABI)")
    config.addinivalue_line("markers", "slow: longer CPU/GPU runs")


@pytest.fixture(scope="session")
def hg():
    import hacc_gpm_amd
    hacc_gpm_amd.lib()
    return hacc_gpm_amd


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.L()
    return oracle


def split_of(x, ng):
    """Grid-unit positions (FP64) in the split layout as the engine holds them: the
    fixed-point coordinate rint(x 2^22) mod ng 2^22 as (cell int32, offset float32 on the
    2^-22 grid)."""
    X = np.mod(np.rint(np.asarray(x, np.float64) * 2.0 ** 22).astype(np.int64), np.int64(ng) << 22)
    return (X >> 22).astype(np.int32), ((X & ((1 << 22) - 1)).astype(np.float64) * 2.0 ** -22).astype(np.float32)


def rand_split(rng, n, ng, clustered=False):
    """Random particles in the split layout (cell int32, offset float32 in [0,1) on the
    2^-22 grid)."""
    if clustered:
        centre = rng.integers(0, ng, 3)
        x = (centre + rng.normal(0, 1.5, (n, 3))) % ng
    else:
        x = rng.random((n, 3)) * ng
    return split_of(x, ng)


def lattice_split(rng, np_, ng, sigma, outliers=0):
    """Lattice particles (reference order, spacing ng/np_) displaced by Gaussian offsets of
    sigma cells, as the Zel'dovich initial conditions and early steps place them; the
    first `outliers` particles are thrown anywhere in the box."""
    q = np.indices((np_,) * 3).reshape(3, -1).T.astype(np.float64) * (ng / np_)
    x = (q + rng.normal(0, sigma, q.shape)) % ng
    if outliers:
        x[:outliers] = rng.random((outliers, 3)) * ng
    return split_of(x, ng)
