import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
for p in (str(ROOT), str(ROOT / "tests"), str(ROOT / "tests" / "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 device and libcupami_hip.so")


def pytest_collection_modifyitems(config, items):
    """PyTorch bundles its own HIP runtime.  Tests that hand torch device
    memory to the library need torch's runtime initialised before the
    library's, so a session that runs GPU tests initialises it first."""
    if not any(item.get_closest_marker("gpu") for item in items):
        return
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.is_available():
        torch.cuda.init()
