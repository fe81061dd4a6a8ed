import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _built(path):
    return os.path.exists(os.path.join(ROOT, path))


@pytest.fixture(scope="session", autouse=True)
def native_libs():
    """Build the native libraries in-tree if (and only if) they are missing."""
    if not (_built("radrhd_solver_amd/libradrhd_mesh.so") and _built("oracle/liboracle_rhllc.so")):
        from radrhd_solver_amd import build
        import oracle

        build.build_mesh()
        oracle.build(ref=False)
    yield
