import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")
    config.addinivalue_line("markers", "slow: long CPU run, opt in with LS_SLOW=1")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("LS_SLOW") == "1":
        return
    skip = pytest.mark.skip(reason="slow CPU test; set LS_SLOW=1")
    for it in items:
        if "slow" in it.keywords and "gpu" not in it.keywords:
            it.add_marker(skip)


def load_lsamd():
    if "lsamd" in sys.modules:
        return sys.modules["lsamd"]
    spec = importlib.util.spec_from_file_location(
        "lsamd", os.path.join(ROOT, "level-set-method-code_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["lsamd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def ls():
    return load_lsamd()


@pytest.fixture(scope="session")
def oracle():
    from oracle.pyoracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    torch.cuda.set_device(0)
    return torch
