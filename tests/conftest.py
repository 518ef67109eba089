import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libmatrixgraph_amd.so on the device)")


@pytest.fixture(scope="session")
def ctx():
    """One GPU context for the whole GPU session (fails loudly without the HIP library)."""
    import matrixgraph_amd as mg

    c = mg.Context(0)
    yield c
    c.close()
