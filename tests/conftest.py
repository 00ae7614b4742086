import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def lux_ctx():
    """One HIP context for the GPU tests (fails loudly without the library)."""
    from bevy_penumbra_amd import lux
    ctx = lux.Context(0)
    yield ctx
    ctx.close()
