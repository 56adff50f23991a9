import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


def rel_up_to_normaliser(Ma, Mb):
    """(relative L2 difference after matching the global scale, relative scale difference).
    The model is divided by its mean over positive voxels; after the blur, voxels far from any
    data hold values at the rounding level whose sign can follow the (unordered) float atomics of
    the insertion, so the positive-voxel count, and with it the scale, can move by 1/count."""
    top = max(float(Ma.max()), float(Mb.max()))
    sel = (Ma > 1e-3 * top) & (Mb > 1e-3 * top)
    c = float(Mb[sel].astype(np.float64).sum() / Ma[sel].astype(np.float64).sum())
    rel = float(np.linalg.norm(c * Ma.astype(np.float64) - Mb) / np.linalg.norm(Mb))
    return rel, abs(c - 1.0)
