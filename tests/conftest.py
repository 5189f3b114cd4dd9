import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs the HIP path)")


@pytest.fixture(scope="session")
def engine():
    import faz_amd

    e = faz_amd.Engine(0)
    yield e
    e.close()


def _engine_with(kernel, var="FAZ_INTERP_KERNEL"):
    import faz_amd

    old = os.environ.get(var)
    os.environ[var] = kernel
    try:
        return faz_amd.Engine(0)
    finally:
        if old is None:
            del os.environ[var]
        else:
            os.environ[var] = old


@pytest.fixture(scope="session")
def engine_pass():
    """A context forced onto the per-pass launches (no fused 3-D level kernel)."""
    e = _engine_with("pass")
    yield e
    e.close()


@pytest.fixture(scope="session")
def engine_level():
    """A context on the per-block level kernel (no whole-tile fast path)."""
    e = _engine_with("level")
    yield e
    e.close()


@pytest.fixture(scope="session")
def engine_onepass():
    """A context on the one-pass (decoupled look-back) Huffman encoder."""
    e = _engine_with("onepass", "FAZ_ENCODE")
    yield e
    e.close()


@pytest.fixture(scope="session")
def engine_tile():
    """A context that runs every 3-D level through the level kernels whatever
    its block count, so small fields exercise the whole-tile kernel."""
    e = _engine_with("tile")
    yield e
    e.close()


@pytest.fixture(scope="session")
def engine_norun():
    """A context whose decoder treats the anchors' opening run of zero codes
    like the rest of the bitstream (no run check, no shortcut)."""
    e = _engine_with("off", "FAZ_DECODE_RUN")
    yield e
    e.close()


@pytest.fixture(scope="session")
def engine_hostbounds():
    """A context whose REL compressions read the value range back and set the
    level bounds on the host."""
    e = _engine_with("off", "FAZ_DEVICE_BOUNDS")
    yield e
    e.close()
