import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture
def oracle():
    import _oracle

    _oracle.System.clear()
    _oracle.set_clearzero_default_eps(1e-14)
    yield _oracle
    _oracle.System.clear()


@pytest.fixture
def ps():
    import sparqsim_amd

    sparqsim_amd.System.clear()
    sparqsim_amd.set_clearzero_default_eps(1e-14)
    yield sparqsim_amd
    sparqsim_amd.System.clear()
