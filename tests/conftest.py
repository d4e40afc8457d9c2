import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    return oracle


@pytest.fixture(scope="session")
def engine():
    from cumetis_amd.engine import Engine
    eng = Engine(0)
    yield eng
    eng.close()


def say(msg):
    """Progress line naming the case, flushed before the GPU step runs."""
    print(msg, flush=True)
