import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")


def gpu_ready() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def dev():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from phrike_amd import _lib

    _lib.load()  # a GPU run without the library must fail, not skip
    return torch.device("cuda:0")


def golden(name):
    import numpy as np

    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        pytest.skip(f"golden file {name} not generated")
    return np.load(path, allow_pickle=False)


def rel_l2(a, b):
    import numpy as np

    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def field_errors(a, b, floor=None):
    """per-field relative L2 (absolute, scaled by the state norm, when a
    field's reference norm is ~0, e.g. rho*w of the TGV)"""
    import numpy as np

    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.linalg.norm(b) / np.sqrt(b.shape[0])
    errs = []
    for f in range(b.shape[0]):
        nb = np.linalg.norm(b[f])
        den = nb if nb > (floor if floor is not None else 1e-8 * scale) else scale
        errs.append(np.linalg.norm(a[f] - b[f]) / den)
    return np.array(errs)


def scaled_err(got, ref, scale):
    """||got - ref|| / max(||ref||, scale): for outputs that can be ~0 by
    construction (e.g. a derivative of a field whose modes are all masked)."""
    import numpy as np

    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    return np.linalg.norm(got - ref) / max(np.linalg.norm(ref), float(scale))


def group_errors(a, b, groups):
    """per-field L2 error over the largest reference norm in the field's group
    (density / momenta / energy): FP32 error on a momentum component that is
    ~0 (rho*w of the TGV) is measured against the momentum scale."""
    import numpy as np

    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    norms = np.array([np.linalg.norm(b[f]) for f in range(b.shape[0])])
    out = []
    for f in range(b.shape[0]):
        g = [grp for grp in groups if f in grp][0]
        out.append(np.linalg.norm(a[f] - b[f]) / max(norms[list(g)].max(), 1e-300))
    return np.array(out)
