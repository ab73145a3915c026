"""CPU-side checks of the C ABI: the library loads, exports every entry point
the headers declare, and rejects bad arguments before touching a device."""
import ctypes as C
import os
import re

import pytest

from seisfat2d_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    inc = os.path.join(ROOT, "include")
    for f in sorted(os.listdir(inc)):
        if f.endswith(".h"):
            txt = open(os.path.join(inc, f)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names += re.findall(r"^\s*(?:const\s+)?[\w\s\*]*?\b(seisfat_\w+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_binding_list():
    assert set(declared_functions()) == set(capi.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(capi.EIK_LIB)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_invalid_arguments_rejected_without_device():
    L = capi.lib()
    h = C.c_void_p()
    assert L.seisfat_eik_create(1, 10, 5.0, 5.0, C.byref(h)) == capi.EINVAL
    assert b"2x2" in L.seisfat_last_error()
    assert L.seisfat_eik_create(10, 10, -5.0, 5.0, C.byref(h)) == capi.EINVAL
    assert L.seisfat_eik_create(10, 10, 5.0, float("nan"), C.byref(h)) == capi.EINVAL
    assert L.seisfat_eik_create(10, 10, 5.0, 5.0, None) == capi.EINVAL
    assert L.seisfat_eik_set_slowness_device(None, None, None) == capi.EINVAL
    assert L.seisfat_batch_run(None, None) == capi.EINVAL
    assert L.seisfat_eik_destroy(None) == capi.OK
    assert L.seisfat_batch_destroy(None) == capi.OK


def test_error_is_raised_as_exception():
    with pytest.raises(capi.SeisfatError) as e:
        capi.check(capi.lib().seisfat_eik_create(0, 0, 1.0, 1.0, C.byref(C.c_void_p())))
    assert e.value.code == capi.EINVAL
