import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; runs the HIP kernels")


def _cuda_ok() -> bool:
    try:
        import torch

        return bool(torch.cuda.is_available())
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """The GPU tests need the HIP library and a visible device; they never fall back to the CPU."""
    if not _cuda_ok():
        pytest.skip("no GPU visible")
    from gonzales_amd import abi

    lib = abi.load()
    assert lib.gz_gpu_available(), "a GPU is visible but it is not a gfx950 device"
    import torch

    torch.cuda.init()
    return torch.device("cuda:0")
