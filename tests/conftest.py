"""Shared fixtures: the product library, the oracle and (when built) the
reference library, plus the ``gpu`` marker."""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "quant-kernel_amd"))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def qk():
    import quantkernel_amd as qka

    return qka.QuantKernel()


@pytest.fixture(scope="session")
def orc():
    import oracle_lib

    return oracle_lib.Oracle()


@pytest.fixture(scope="session")
def ref():
    import oracle_lib

    r = oracle_lib.Reference.load()
    if r is None:
        pytest.skip("reference build oracle/_ref/libqkref.so not present")
    return r


@pytest.fixture(autouse=True)
def _philox_default(request):
    """Tests that change the RNG mode restore the default afterwards."""
    yield
    if "qk" in request.fixturenames:
        request.getfixturevalue("qk").set_rng_mode(0)
