import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


@pytest.fixture(scope="session")
def _session_ctx():
    from moquimc_amd import _abi
    c = _abi.Context(0)
    yield c
    c.close()


@pytest.fixture
def ctx(_session_ctx):
    """The session's context with the nominal scenario only, no beamline and
    DIRECT scoring."""
    _session_ctx.set_scenarios()
    _session_ctx.set_beamline()
    _session_ctx.set_roi(0)
    _session_ctx.set_roi(1)
    return _session_ctx
