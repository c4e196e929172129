import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session", autouse=True)
def built():
    """The in-tree library (built only when absent) and the CPU oracle."""
    import ctbangbang_amd
    if not os.path.exists(ctbangbang_amd.library_path()):
        from ctbangbang_amd import build
        build.build()
    import oracle
    if not os.path.exists(os.path.join(oracle.HERE, "liboracle64.so")):
        oracle.build()
    return True


@pytest.fixture(scope="session")
def mini_dir(tmp_path_factory):
    from ctbangbang_amd import scans
    d = str(tmp_path_factory.mktemp("mini_install"))
    scans.mini_install(d)
    return d


def has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def ctx():
    import ctbangbang_amd as cb
    if not has_gpu():
        pytest.skip("no GPU")
    c = cb.Context(0)
    yield c
    c.close()
