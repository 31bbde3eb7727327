import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def gpu_ctx():
    """A context on device 0. On a machine without a gfx950 device this FAILS (GPU tests are
    never skipped silently); the native library must be built."""
    import portablert_amd as prt
    n, name = prt.probe()
    assert n > 0, "no gfx950 device visible: GPU tests need an MI355X"
    print(f"device: {name}", flush=True)
    ctx = prt.Context(0)
    yield ctx
    ctx.close()
