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
