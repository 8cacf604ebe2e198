"""The C-ABI library loads and exports every symbol include/ndsp.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import ndspmhd_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "ndsp.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(ndsp_\w+)\s*\(", txt, re.M)))


def test_header_and_python_agree():
    assert set(declared_symbols()) == set(ndspmhd_amd.SYMBOLS)


def test_library_exports_all_symbols():
    lib = ctypes.CDLL(ndspmhd_amd.LIBPATH)
    for s in declared_symbols():
        assert hasattr(lib, s), s


def test_bad_option_rejected_without_gpu_work():
    """ndsp_create validates options before touching memory; with no device it
    reports NO_DEVICE, with a device BAD_OPTION -- either way non-zero."""
    import numpy as np
    L = ndspmhd_amd.load_library()
    o = ndspmhd_amd.Options().replace(iav=1)
    iopt = o.iopt().astype(np.int32)
    dopt = o.dopt()
    h = ctypes.c_void_p()
    rc = L.ndsp_create(iopt.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                       dopt.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 10, 0, ctypes.byref(h))
    assert rc != 0
