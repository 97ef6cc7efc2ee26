import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")


@pytest.fixture(scope="session")
def oracle():
    """The CPU restatement (built on demand; it ships prebuilt to the GPU box)."""
    from oracle import pyoracle
    if not pyoracle.ORACLE_SO.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "oracle"], check=True)
    return pyoracle.Oracle("oracle")


@pytest.fixture(scope="session")
def ref():
    """The reference compiled from its own headers; skipped where it was not built."""
    from oracle import pyoracle
    if not pyoracle.ref_available():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    return pyoracle.Oracle("ref")


@pytest.fixture(scope="session")
def ipf():
    """libipfg through its Python mirror; loading fails loudly if it is not built."""
    import imageprocessingflex_amd as m
    m._lib.load()
    m.init(0)
    return m
