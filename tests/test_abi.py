"""The C ABI library loads (no GPU needed) and exports every declared symbol."""
import ctypes
import os
import re

from conftest import ROOT


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "tnsgpu.h")).read()
    return sorted(set(re.findall(r"\b(tns_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from tnsp_amd import _lib

    L = _lib.lib()
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the ctypes layer binds exactly the declared surface
    assert sorted(_lib.exported_symbols()) == names


def test_error_path_reports_message():
    from tnsp_amd import _lib

    L = _lib.lib()
    h = ctypes.c_void_p()
    names = _lib.arr32([0, 0])
    nseg = _lib.arr32([1, 1])
    syms = _lib.arr32([0, 0, 0, 0])
    dims = _lib.arr64([2, 2])
    rc = L.tns_shape_create(3, 2, names, nseg, syms, dims, None, ctypes.byref(h))
    assert rc != 0
    assert b"duplicated" in L.tns_last_error()
    rc = L.tns_shape_create(42, 0, names, nseg, syms, dims, None, ctypes.byref(h))
    assert rc != 0 and b"kind" in L.tns_last_error()


def test_svd_sweep_limit_range_checked():
    from tnsp_amd import _lib

    L = _lib.lib()
    assert L.tns_set_svd_max_sweeps(0) != 0 and b"sweep" in L.tns_last_error()
    assert L.tns_set_svd_max_sweeps(60) == 0


def test_noalloc_rejects_before_any_device_call():
    """tns_plan_run_noalloc refuses plans beyond its static workspace and svd
    plans with a clear error, before touching a device (runs without one)."""
    import pytest

    from tnsp_amd import _lib
    from tnsp_amd import plan as P
    from tnsp_amd.shape import KINDS, Shape

    no = KINDS["No"]
    a = Shape.create(no, ["i", "k"], [3000, 3000])
    b = Shape.create(no, ["k", "j"], [3000, 3000])
    big = P.contract(a, b, [("k", "k")])
    assert big.workspace("float64") > 64 << 20
    with pytest.raises(RuntimeError, match="static maximum"):
        big.run_noalloc("float64", [0, 0], [0], 0)
    s = P.svd(Shape.create(no, ["i", "j"], [4, 4]), ["i"], "U", "V", "SU", "SV")
    with pytest.raises(RuntimeError, match="svd plans"):
        s.run_noalloc("float64", [0], [0, 0, 0], 0)
    assert _lib.lib().tns_plan_run_noalloc(big.handle, 7, None, None, None) != 0
