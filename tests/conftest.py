import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")
    # Build the product library and the oracle if a fresh checkout lacks them
    # (make is incremental; normally __graft_entry__.build() has done this).
    import subprocess
    for sub, lib in (("rabbitbin_amd/csrc", "rabbitbin_amd/librbg.so"), ("oracle", "oracle/liboracle.so")):
        if not os.path.exists(os.path.join(REPO, lib)):
            subprocess.run(["make", "-C", os.path.join(REPO, sub), "-j8"], check=True,
                           stdout=subprocess.DEVNULL)


@pytest.fixture(scope="session")
def ctx():
    from rabbitbin_amd import api
    c = api.Context(0)
    yield c
    c.close()
