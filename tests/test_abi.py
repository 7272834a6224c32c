"""The C-ABI libraries load and export every symbol their headers declare
(no GPU call is made here)."""
import os
import re

import cpsd_amd


def declared(header):
    txt = open(header).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = re.sub(r"^\s*#.*$", "", txt, flags=re.M)      # preprocessor lines (macros)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt)
    skip = {"if", "defined"}
    return sorted(set(n for n in names if n not in skip))


def test_hip_library_exports_header_symbols():
    lib = cpsd_amd.hip_lib()
    names = declared(os.path.join(cpsd_amd.ROOT, "include", "cpsd_abi.h"))
    assert set(names) == set(cpsd_amd.ABI_SYMBOLS)
    for n in names:
        assert hasattr(lib, n), n


def test_host_library_exports_header_symbols():
    lib = cpsd_amd.host_lib()
    names = declared(os.path.join(cpsd_amd.ROOT, "include", "cpsd_host.h"))
    assert set(names) == set(cpsd_amd.HOST_SYMBOLS)
    for n in names:
        assert hasattr(lib, n), n


def test_dropin_signature_is_the_reference_one():
    # cpsd takes the 42 arguments of ccpsdv3/cpsd.c:33-71
    txt = open(os.path.join(cpsd_amd.ROOT, "include", "cpsd_abi.h")).read()
    m = re.search(r"void cpsd\((.*?)\);", txt, flags=re.S)
    assert m and len(m.group(1).split(",")) == 42


def test_resident_api_rejects_a_null_handle():
    """every cpsd_dev_* call on a null handle returns an error with a message
    and touches no device (host logic only)"""
    import ctypes as C
    import numpy as np
    L = cpsd_amd.hip_lib()
    d = np.zeros(16)
    i = np.zeros(4, dtype=np.int32)
    z = C.c_double()
    k = C.c_int()
    calls = [
        lambda: L.cpsd_dev_set_orbitals(None, d),
        lambda: L.cpsd_dev_run(None, 1, 5.8, 4e5, d, C.byref(z), C.byref(z), C.byref(z)),
        lambda: L.cpsd_dev_set_ions(None, 0, d, d),
        lambda: L.cpsd_dev_get_ions(None, None, None, None),
        lambda: L.cpsd_dev_get(None, None, None, None, None, None, None),
        lambda: L.cpsd_dev_epilogue(None, None, None, None),
        lambda: L.cpsd_dev_set_precision(None, 1),
        lambda: L.cpsd_dev_set_graphs(None, 1),
        lambda: L.cpsd_dev_lagrange_info(None, C.byref(k), C.byref(z)),
        lambda: L.cpsd_dev_sizes(None, i),
        lambda: L.cpsd_dev_profile(None, 1),
        lambda: L.cpsd_dev_profile_read(None, C.create_string_buffer(64), 64, d, i, 4),
        lambda: L.cpsd_dev_debug_fetch(None, 0, d, 16),
        lambda: L.cpsd_dev_run_start(None, 1, 5.8, 4e5, d),
        lambda: L.cpsd_dev_run_finish(None, d, C.byref(z), C.byref(z), C.byref(z)),
    ]
    for f in calls:
        assert f() == -1
        assert "null" in cpsd_amd._err()
    L.cpsd_dev_destroy(None)
