import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def engine():
    import afmm2d_amd
    afmm2d_amd.lib()
    return afmm2d_amd


def uniform_points(n, seed):
    import numpy as np
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.0, 1.0, size=(n, 2)), rng.uniform(-1.0, 1.0, size=n)


def mixture_points(n, seed, clusters=16, sigma=0.05):
    """Points from a mixture of isotropic Gaussians: centres uniform in
    [-1,1]^2, standard deviation sigma, clipped to the box; and U(-1,1)
    charges."""
    import numpy as np
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-1.0, 1.0, size=(clusters, 2))
    lab = rng.integers(0, clusters, size=n)
    xy = np.clip(centres[lab] + sigma * rng.standard_normal((n, 2)), -1.0, 1.0)
    return xy, rng.uniform(-1.0, 1.0, size=n)


def oracle_kernel(name):
    """The oracle's kernel id for an engine kernel name (a user kernel
    declared symmetric is the user kernel itself)."""
    import afmm2d_amd
    kid = afmm2d_amd.KERNELS[name]
    return 4 if kid == 7 else kid


def bfp48_round(block):
    """A block as the 48-bit block floating-point M2L storage keeps it: one
    exponent s for the block, the smallest with every |entry| / 2^s rounded
    to an integer of magnitude <= 2^46 (s = e - 46 for max |entry| = f 2^e,
    1/2 <= f < 1), each entry rounded to nearest (ties to even) on that
    grid."""
    import numpy as np
    a = np.asarray(block, dtype=np.float64)
    mx = float(np.max(np.abs(a))) if a.size else 0.0
    if mx == 0.0:
        return a.copy(), 0
    s = int(np.frexp(mx)[1]) - 46
    return np.ldexp(np.rint(np.ldexp(a, -s)), s), s
