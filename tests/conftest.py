import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (HERE, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def golden():
    import _formats as fm
    g = os.path.join(HERE, "golden")
    x = fm.read_raw(os.path.join(g, "data.bin"))
    w, mu, var, _ = fm.read_gmm_txt(os.path.join(g, "myModel.gmm"))
    ll = fm.read_savemat_txt(os.path.join(g, "data.logL"))[0]
    return dict(x=x, w=w, mu=mu, var=var, ll=ll)


@pytest.fixture(scope="session", autouse=True)
def _torch_device_first(request):
    """Tests that use torch device tensors next to the library: torch's own
    HIP runtime must initialise the device before the library's runtime
    does, as bench.py does (torch first, then the library's handles)."""
    if any(item.get_closest_marker("gpu") for item in request.session.items):
        try:
            import torch
            if torch.cuda.is_available():
                torch.empty(1, device="cuda")
        except ImportError:
            pass
    yield
