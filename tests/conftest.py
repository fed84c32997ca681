import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "bem-mececcs_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


@pytest.fixture(scope="session")
def orc():
    import oracle
    if not os.path.exists(oracle._LIBS["c"]):
        oracle.build()
    return oracle.Oracle("c")


@pytest.fixture(scope="session")
def ref():
    import oracle
    if not oracle.ref_available():
        pytest.skip("reference build oracle/_ref not present")
    return oracle.Oracle("ref")


def block_scale(b):
    """max |entry| of the 3x3 block each entry belongs to.

    An (n, n) matrix with n = 3 NBE is cut into its (collocation, element)
    3x3 blocks; a (k, 9) array is k blocks of 9; a (3, 3) array is one."""
    ab = np.abs(np.asarray(b))
    if ab.ndim == 2 and ab.shape[0] == ab.shape[1] and ab.shape[0] % 3 == 0 and ab.shape[0] > 3:
        m = ab.shape[0] // 3
        bm = ab.reshape(m, 3, m, 3).max(axis=(1, 3))
        return np.repeat(np.repeat(bm, 3, axis=0), 3, axis=1)
    if ab.ndim == 2 and ab.shape[1] == 9:
        return np.repeat(ab.max(axis=1, keepdims=True), 9, axis=1)
    return np.full(ab.shape, ab.max())


def close_blocks(a, b, rtol):
    """Entrywise |a-b| <= rtol * max(|b|, max|b| over its 3x3 block).

    This is the per-entry relative check with a norm floor: exactly-zero and
    cancellation-born entries (flat-mesh zeros, the off-diagonal terms of the
    singular diagonal blocks) are judged against their block's scale, since
    their own magnitude carries no significant digits.  Returns (ok, worst
    relative deviation)."""
    a = np.asarray(a)
    b = np.asarray(b)
    scale = np.maximum(np.abs(b), block_scale(b))
    rel = np.abs(a - b) / np.maximum(scale, 1e-300)
    return bool(np.all(rel <= rtol)), float(rel.max())
