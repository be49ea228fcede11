import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "p9-distributed-reconstruction_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")


@pytest.fixture(scope="session")
def O():
    import oracle

    return oracle


def oracle_constants(c):
    import oracle

    return oracle.make_constants(c.VisibilitiesCount, c.GridSize, c.SubgridSize, c.KernelSize,
                                 c.MaxTimestepsPerSubgrid, float(c.ScaleArcSec), c.WLayerCount, c.WStep)


def rel_err(a, b):
    import numpy as np

    return float(np.abs(a - b).max() / np.abs(b).max())
