import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")
    config.addinivalue_line("markers", "slow: longer CPU test")


def pytest_report_header(config):
    """The host's C library and CPU beside every run: the oracle's glibc sin,
    cos and atan2 are what the reference's rounding depends on."""
    import platform
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return [f"glibc: {' '.join(platform.libc_ver())}; cpu: {model}"]


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (device count without initialising HIP)
        return torch.cuda.device_count() > 0
    except Exception:
        return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def ctx():
    """One HIP context for the whole GPU session (one process, one stream)."""
    from curvelet_construction_on_gpu_amd import Context
    c = Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def oracle():
    """The restatement with glibc math (bit-exact with the reference)."""
    from oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def oracle_cr():
    """The same restatement with the GPU path's correctly rounded math."""
    from oracle import Oracle
    return Oracle(cr=True)
