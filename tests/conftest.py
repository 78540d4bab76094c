import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
TESTS = Path(__file__).resolve().parent
if str(TESTS) not in sys.path:
    sys.path.insert(0, str(TESTS))
# runtime-compiled kernels are cached on disk; the test session keeps its own cache instead of the user's
if "SC_KERNEL_CACHE_DIR" not in os.environ:
    import atexit
    import shutil
    import tempfile

    os.environ["SC_KERNEL_CACHE_DIR"] = tempfile.mkdtemp(prefix="sc_kernel_cache_")
    atexit.register(shutil.rmtree, os.environ["SC_KERNEL_CACHE_DIR"], True)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def hip_ctx():
    from statecraft_amd import capi

    # The shared context runs the precompiled sweep kernel (jit 0, 4 low qubits, in-place relabelling): the
    # parity suite's many circuit structures then need no run-time compilation.  The library's defaults
    # (runtime-compiled passes, geometry chosen per circuit) are exercised by the tests that create their own
    # contexts and by the engine API tests.
    ctx = capi.Context(0)
    ctx.set_option(capi.OPT_JIT, 0)
    ctx.set_option(capi.OPT_LOW_QUBITS, 4)
    ctx.set_option(capi.OPT_REMAP, 1)
    ctx.set_option(capi.OPT_CHUNK_BITS, 12)
    yield ctx
    ctx.close()


@pytest.fixture
def rng():
    return np.random.default_rng(0x51D5)
