import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "dcu-winograd-conv2d_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on cuda:0)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    return O


@pytest.fixture(scope="session")
def wc():
    # PyTorch-ROCm ships its own HIP runtime under the same soname: importing
    # torch first makes libwinoconv bind to that one runtime instead of
    # loading a second runtime that would compete for the device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    import winoconv
    winoconv.lib()
    return winoconv


@pytest.fixture(scope="session")
def gpu():
    """Device helpers for GPU tests (torch is plumbing for device memory)."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gpu_helpers
    return gpu_helpers


def pytest_collection_modifyitems(config, items):
    # keep CPU-only sessions fast: nothing to adjust, markers drive selection
    return


def rng_inputs(seed, n, c, h, w, k, dist="u11"):
    """Seeded synthetic inputs of the configured distributions."""
    rng = np.random.default_rng(seed)
    if dist == "u11":
        x = rng.uniform(-1, 1, (n, c, h, w))
        wt = rng.uniform(-1, 1, (k, c, 3, 3))
    elif dist == "clip4":  # N(0,1) clipped to +-4, He filters
        x = np.clip(rng.standard_normal((n, c, h, w)), -4, 4)
        wt = rng.standard_normal((k, c, 3, 3)) * np.sqrt(2.0 / (9 * c))
    elif dist == "normal":
        x = rng.standard_normal((n, c, h, w))
        wt = rng.standard_normal((k, c, 3, 3)) * np.sqrt(2.0 / (9 * c))
    else:
        raise ValueError(dist)
    return x.astype(np.float16), wt.astype(np.float16)
