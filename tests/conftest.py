import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "ct-research_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libctr.so")


def gpu_available():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def ctx():
    # torch carries its own HIP runtime: initialise it on the device before the
    # library's runtime creates its streams, so that the tests which hand torch
    # tensors to the library find torch's side ready whatever ran before them
    import torch
    torch.cuda.init()
    import ctr
    c = ctr.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        O.build()
    return O


def report(name, **vals):
    """Print a parity line (visible with -s / -v on failures) and flush it."""
    msg = " ".join(f"{k}={v:.3e}" if isinstance(v, float) else f"{k}={v}" for k, v in vals.items())
    print(f"[parity] {name}: {msg}", flush=True)


def errs(x, ref):
    x = np.asarray(x, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    scale = max(np.abs(ref).max(), 1e-30)
    rel_l2 = np.linalg.norm(x - ref) / max(np.linalg.norm(ref), 1e-30)
    max_rel = np.abs(x - ref).max() / scale
    return float(rel_l2), float(max_rel)
