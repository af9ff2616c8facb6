import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def gpu():
    """GPU tests fail (not skip) without a device: they only run with -m gpu
    on the GPU box, where a missing device or library is a real failure."""
    import torch

    from tool_godot_curvenet_amd import _lib
    assert torch.cuda.is_available(), "no HIP device visible"
    lib = _lib.load()
    assert lib.cn_device_check() == 0
    return lib
