import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (REPO, os.path.join(REPO, "can-the-fly-play-f-zero-_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def pytest_report_header(config):
    # which native library the session loads (path, file hash, the source hash built into it and
    # whether it matches the sources in this tree)
    try:
        from flyhip import _lib

        return "flyhip library: " + ", ".join(f"{k}={v}" for k, v in _lib.library_info().items())
    except Exception as exc:  # no library built: the GPU tests fail loudly on their own
        return f"flyhip library: not loaded ({exc})"


@pytest.fixture(autouse=True)
def _name_the_case(request):
    # flushed before every test so a GPU failure names its case
    print(f"\n[case] {request.node.nodeid}", flush=True)
    yield


@pytest.fixture(scope="session")
def full_fly():
    """The full-size FlyWire-shaped connectome: 138,639 neurons, 15.1 M connections (seed 0)."""
    from flyhip.connectome import flywire_like

    return flywire_like()
