import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "dennou-ogcm_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def rel_l2(x, ref):
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    d = np.linalg.norm((x - ref).ravel())
    n = np.linalg.norm(ref.ravel())
    return d / n if n > 0 else d


def red_spectrum(nmax, ncols, seed, amp=1.0):
    """Complex Gaussian coefficients with variance ~ (n+1)^-3 (Im = 0 at m = 0), SPMODEL order."""
    rng = np.random.default_rng(seed)
    L = (nmax + 1) ** 2
    n = np.floor(np.sqrt(np.arange(L))).astype(int)
    sd = amp * np.sqrt(0.5 * (n + 1.0) ** -3)
    return rng.standard_normal((ncols, L)) * sd


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import dsht
    dsht.load()
    return torch.device("cuda:0")
