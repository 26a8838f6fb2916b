"""Shared pytest configuration: the ``gpu`` marker and import paths for the package and oracle."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "am1-rs_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libam1g.so on cuda:0)")


@pytest.fixture(scope="session")
def engine():
    import am1g

    eng = am1g.Engine(0)
    yield eng
    eng.close()
