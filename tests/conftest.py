"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on CPU."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI)")


@pytest.fixture(scope="session")
def orc():
    import pb_oracle
    pb_oracle.lib()
    return pb_oracle


@pytest.fixture(scope="session")
def ctx():
    import particlebase_amd as pb
    c = pb.Context(0)
    name, cus = c.info()
    print(f"\n[gpu] device {name} with {cus} CUs", flush=True)
    yield c
    c.close()


@pytest.fixture
def rng():
    return np.random.default_rng(12345)
