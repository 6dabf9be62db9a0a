import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "lbdemcoupling-public_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")
    config.addinivalue_line("markers", "slow: long-running test")


def _ensure_built():
    """Build the oracle / library in-tree if a fresh checkout lacks them."""
    oracle_so = os.path.join(ROOT, "oracle", "liboracle_psm.so")
    lib_so = os.path.join(PKG, "lib", "liblbdem_psm.so")
    if not os.path.exists(oracle_so):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True)
    if not os.path.exists(lib_so) and os.path.exists("/opt/rocm/bin/hipcc"):
        subprocess.run(["make", "-C", PKG, "-j8"], check=True)


_ensure_built()


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401
        return bool(torch.cuda.is_available())
    except Exception:
        return False


@pytest.fixture(scope="session")
def psm():
    import lbdem_psm
    lbdem_psm.load_library()
    return lbdem_psm


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.load()
    return oracle
