"""The C-ABI libraries load and export every symbol their headers declare
(no GPU call is made here)."""
import os
import re

import cpsd_amd


def declared(header):
    txt = open(header).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
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
