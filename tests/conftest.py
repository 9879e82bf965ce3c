import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def ct():
    import ctbb_icd_amd as m
    m.set_verbose(False)
    return m


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def prm_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("prm")


@pytest.fixture(scope="session")
def make_scan(ct, prm_dir):
    from ctbb_icd_amd import configs
    cache = {}

    def _make(name, **over):
        key = (name, tuple(sorted(over.items())))
        if key not in cache:
            tag = f"{name}_{abs(hash(key)) % 10**8}"
            p = configs.write_prm(name, prm_dir / f"{tag}.yaml", **over)
            cache[key] = ct.Scan(p)
        return cache[key]
    return _make
