import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "seismic.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libseiswem.so")


def rel_l2(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    a = a.astype(np.complex128) if np.iscomplexobj(a) or np.iscomplexobj(b) else a.astype(np.float64)
    b = b.astype(a.dtype)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="session")
def gpu():
    """The loaded library with at least one gfx950 device; errors (does not skip) otherwise."""
    import seiswem

    lib = seiswem.load()
    n = seiswem.device_count()
    if n < 1:
        raise RuntimeError("GPU test requested but no HIP device is visible")
    return lib
