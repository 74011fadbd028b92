import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "qiskit-blackwell_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built library")
    # the C-ABI library and the C oracle are built in-tree; build them once if
    # this checkout has not (hipcc cross-compiles without a GPU)
    lib = os.path.join(PKG, "blackwell_kernels", "libqsim_gfx950.so")
    if not os.path.exists(lib):
        import importlib
        importlib.import_module("build").build()
    from oracle import build_ref
    if not os.path.exists(build_ref.OUT):
        build_ref.build()


def _gpu_ok():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _gpu_ok():
        return
    skip = pytest.mark.skip(reason="no GPU visible")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def bk():
    import blackwell_kernels
    blackwell_kernels.load_library()
    return blackwell_kernels


@pytest.fixture(scope="session")
def dev():
    import torch
    return torch.device("cuda", 0)
