import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "movement-sim_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests of the HIP path")
    config.addinivalue_line("markers", "slow: longer CPU test")


def pytest_collection_modifyitems(session, config, items):
    """With GPU tests selected, torch's device runtime starts before any test
    loads libmsim, the order the whole suite runs in (torch's first device
    use has failed to find the GPU when it came after libmsim's)."""
    if not any(item.get_closest_marker("gpu") for item in items):
        return
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.is_available():
        torch.cuda.init()
