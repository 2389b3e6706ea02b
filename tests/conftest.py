import os
import sys

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def golden():
    import numpy as np
    path = os.path.join(ROOT, "tests", "golden", "gsks_golden.npz")
    return dict(np.load(path, allow_pickle=False))


def golden_cases(g):
    """Yield (key, dict) for every single-call case of the golden fixture."""
    keys = sorted({k.split("_")[0] for k in g if k[0] in "cd" and k.endswith("_meta")})
    for key in keys:
        d = {name[len(key) + 1:]: g[name] for name in g if name.startswith(key + "_")}
        yield key, d
