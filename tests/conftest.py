import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libpfgpu on cuda:0)")
    # torch and libpfgpu both need libamdhip64.so.7 and the first one loaded
    # serves the process; torch only initialises its GPU runtime against its
    # own copy.  A session that runs GPU tests therefore brings torch's
    # runtime up before any test loads libpfgpu.
    markexpr = (config.option.markexpr or "").replace(" ", "")
    if "notgpu" in markexpr:
        return
    try:
        import torch
    except ImportError:
        return
    torch.cuda.is_available()
