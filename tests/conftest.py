import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs through libicephase.so")
    config.addinivalue_line("markers", "slow: larger CPU oracle cases")


@pytest.fixture(scope="session")
def oracle():
    import oracle_bind
    if not oracle_bind.ORACLE_SO.exists():
        oracle_bind.build_oracle()
    oracle_bind.lib()
    return oracle_bind


@pytest.fixture(scope="session")
def ctx():
    """One HIP context for the whole GPU session (tests run in one process)."""
    import ice_crystal_phase_amd as ip
    c = ip.Context(int(os.environ.get("ICEPHASE_DEVICE", "0")))
    yield c
    c.close()
