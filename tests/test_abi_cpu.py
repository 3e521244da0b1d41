"""The C-ABI libraries load and export every function their headers declare."""
import ctypes
import os
import re

import pytest

from fastercap_amd import _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:fcg|fch)_[a-z_0-9]+)\s*\(", text)))


@pytest.mark.parametrize("header,loader,listed", [
    ("fastercap_amd.h", _ffi.device_lib, _ffi.DEVICE_SYMBOLS),
    ("fastercap_host.h", _ffi.host_lib, _ffi.HOST_SYMBOLS),
])
def test_exports(header, loader, listed):
    names = declared(header)
    assert names, header
    assert sorted(names) == sorted(listed)
    lib = loader()
    for n in names:
        assert hasattr(lib, n), n


def test_node_struct_size():
    assert ctypes.sizeof(_ffi.NodeGeom) == 96
