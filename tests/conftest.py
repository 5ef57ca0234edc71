import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "beagle-lib_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

ORACLE_LIB = os.path.join(REPO, "oracle", "liboracle.so")
REF_LIB = os.path.join(REPO, "oracle", "_ref", "libhmsbeagle.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on cuda:0)")


@pytest.fixture(scope="session")
def oracle():
    from beagle_amd import api
    if not os.path.exists(ORACLE_LIB):
        pytest.fail("oracle/liboracle.so missing: run __graft_entry__.build()")
    return api.Library(ORACLE_LIB, prefix="orc")


@pytest.fixture(scope="session")
def ref():
    from beagle_amd import api
    if not os.path.exists(REF_LIB):
        pytest.skip("reference CPU build (oracle/_ref) not present")
    return api.Library(REF_LIB, prefix="beagle")


@pytest.fixture(scope="session")
def amd():
    import beagle_amd
    return beagle_amd.load()
