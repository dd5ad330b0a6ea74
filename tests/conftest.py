import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "t-blaeq_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_available():
    try:
        import torch  # noqa: F401  (device discovery only)
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(autouse=True)
def _close_test_handlers():
    """Handlers a test makes through batch_checks.make are destroyed when the test ends,
    passed or failed (module fixtures take theirs off the list themselves)."""
    import batch_checks

    n = len(batch_checks._OPEN)
    yield
    batch_checks.close_new(n)
