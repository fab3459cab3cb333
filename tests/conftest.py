import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

_LIBS = [
    os.path.join(ROOT, "final184_amd", "libvgi.so"),
    os.path.join(ROOT, "final184_amd", "libvgi_gen.so"),
    os.path.join(ROOT, "oracle", "_build", "libvgi_oracle.so"),
]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: longer CPU test")
    # Build only when nothing is built (a fresh checkout); otherwise use what is there.
    if not all(os.path.exists(p) for p in _LIBS):
        from final184_amd import build as b
        b.build_all()


@pytest.fixture(scope="session")
def ctx():
    from final184_amd import vgi
    c = vgi.Context(0)
    yield c
    c.close()
