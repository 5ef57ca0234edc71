"""The C-ABI library loads and exports every entry point the header
declares; without a GPU it refuses to create instances (no fallback)."""
import os
import re

import pytest

from conftest import REPO


def declared(header="beagle.h"):
    txt = open(os.path.join(REPO, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(beagle[A-Z]\w*)\s*\(", txt)))


def test_header_declares_reference_api():
    names = declared()
    assert len(names) == 54
    assert "beagleUpdatePartials" in names and "beagleCalculateEdgeDerivatives" in names


def test_library_exports_every_declared_symbol(amd):
    names = declared() + declared("beagle_amd_ext.h")
    assert len(names) > 54
    for n in names:
        assert hasattr(amd.lib, n), n


def test_resource_zero_is_host(amd):
    res = amd.resources()
    assert res[0]["name"].startswith("CPU")
    for r in res[1:]:
        assert r["requiredFlags"] & (1 << 32)


def test_create_without_gpu_fails_loudly(amd):
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from beagle_amd import api
    with pytest.raises(api.BeagleError) as e:
        amd.create(4, 3, 4, 4, 10, 1, 6, 1, 0)
    assert e.value.code == api.ERROR_NO_RESOURCE


def test_bad_instance_number(amd):
    from beagle_amd import api
    import ctypes
    assert amd.call("FinalizeInstance", 12345) == api.ERROR_UNINITIALIZED_INSTANCE
    assert amd.call("FinalizeInstance", -1) == api.ERROR_UNINITIALIZED_INSTANCE
