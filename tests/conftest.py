import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "oracle"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the native library")


@pytest.fixture(scope="session")
def ctx():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from untangling26_amd.runtime import Context

    c = Context(0)
    yield c
    c.synchronize()


def pytest_runtest_logstart(nodeid, location):
    # name the case before it runs, so that a run cut short still shows what was running
    print(f"\n[start] {nodeid}", flush=True)
