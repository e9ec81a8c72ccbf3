import os
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: full-size workload")


@pytest.fixture(scope="session")
def gpu():
    """One vp_ctx for the whole session (the HIP library; no fallback)."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    from voxplat_amd import GPU

    g = GPU(0)
    yield g
    torch.cuda.synchronize()
    g.close()


def pytest_runtest_setup(item):
    # name the case before a GPU step, so a lost box still says which one ran
    if "gpu" in item.keywords:
        sys.stdout.write(f"\n[gpu-case] {item.nodeid}\n")
        sys.stdout.flush()
