import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def ctx():
    """One HIP context for the GPU tests (fails loudly without a device)."""
    import neon_amd

    c = neon_amd.Context(0)
    yield c
    c.close()
