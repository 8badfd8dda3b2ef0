"""The C-ABI library loads and exports every entry point declared in
include/sw4hip.h; host-only set-up calls validate their arguments."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from common import ROOT
from sw4lite_amd import lib as L
from sw4lite_amd import ew


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "sw4hip.h")).read()
    return sorted(set(re.findall(r"\b(sw4hip_[a-z0-9_]+)\s*\(", txt)))


def test_header_symbols_exported():
    lib = C.CDLL(L.LIBPATH)
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_header():
    assert set(declared_functions()) == set(L._SIGS)


def test_version_and_error_string():
    lib = L.load()
    assert lib.sw4hip_version() >= 1
    d = L.grid_desc(5, 40, 40, 0.1, [2] * 6)   # nx too small
    wind = np.zeros(36, dtype=np.int32)
    npts = np.zeros(6, dtype=np.int32)
    rc = lib.sw4hip_host_boundary_windows(C.byref(d), wind.ctypes.data_as(C.c_void_p),
                                          npts.ctypes.data_as(C.c_void_p))
    assert rc == 1
    assert b"12" in lib.sw4hip_last_error()


def test_bad_descriptor_rejected():
    d = L.grid_desc(20, 20, 20, 0.1, [2, 2, 2, 2, 7, 2])
    with pytest.raises(L.Sw4HipError):
        ew.host_boundary_windows(d)
    d = L.grid_desc(20, 20, 20, -0.1, [2] * 6)
    with pytest.raises(L.Sw4HipError):
        ew.host_supergrid(d)


def test_kernel_entry_rejects_bad_layout():
    lib = L.load()
    rc = lib.sw4hip_predfort(2, -1, 10, -1, 10, -1, 10, None, None, None, None, None, None, 1.0, None)
    assert rc == 1
    assert b"corder" in lib.sw4hip_last_error()


def test_field_checks_of_the_ew_bindings():
    """EW.set_material/set_solution/gradient copy n float64 values per array:
    other dtypes, sizes, layouts or mixed host/device inputs are refused
    before any pointer reaches the library."""
    import torch
    n = 60
    ok = np.zeros(n)
    assert ew._fields(n, ok, ok.copy()) is False
    assert ew._fields(n, torch.zeros(n, dtype=torch.float64)) is False
    with pytest.raises(ValueError, match="float64"):
        ew._fields(n, np.zeros(n, dtype=np.float32))
    with pytest.raises(ValueError, match="float64"):
        ew._fields(n, torch.zeros(n))                      # torch's default float32
    with pytest.raises(ValueError, match="elements"):
        ew._fields(n, np.zeros(n - 1))
    with pytest.raises(ValueError, match="contiguous"):
        ew._fields(n, np.zeros(2 * n)[::2])

    class FakeDev:  # a device tensor as far as the checks look
        dtype = torch.float64
        device = "cuda:0"
        is_cuda = True

        def data_ptr(self):
            return 0

        def numel(self):
            return n

        def is_contiguous(self):
            return True

    with pytest.raises(ValueError, match="all be device tensors or all host"):
        ew._fields(n, FakeDev(), ok)
    assert ew._fields(n, FakeDev(), FakeDev()) is True
