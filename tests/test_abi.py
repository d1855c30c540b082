"""The drop-in boundary without a GPU: both libraries load, libsemper_c.so
exports exactly the reference's 11 symbols (tests/c/abi_symbols.txt), every
function the headers declare resolves, and the frozen layouts keep their
sizes (tests/c/contract.c:30-47). No call here reaches the GPU."""
import ctypes
import os
import re
import subprocess

import pytest

from helpers import PKG

ROOT = os.path.dirname(PKG)
REFERENCE_EXPORTS = {
    "semper_cancel", "semper_create", "semper_destroy", "semper_incremental_cancel",
    "semper_incremental_create", "semper_incremental_destroy", "semper_incremental_init",
    "semper_incremental_process", "semper_run", "semper_set_reference", "semper_version",
}


def declared(header, macro):
    text = open(os.path.join(ROOT, "include", header)).read()
    return set(re.findall(r"^" + macro + r"\s+[^;(]*?\b(\w+)\s*\(", text, re.M))


def exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], check=True, capture_output=True, text=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if ln.split() and ln.split()[-2] in ("T", "t", "W")}


@pytest.fixture(scope="module")
def libs():
    c = os.path.join(PKG, "libsemper_c.so")
    d = os.path.join(PKG, "libsemper_dev.so")
    for p in (c, d):
        if not os.path.exists(p):
            pytest.fail(f"{p} is not built; run __graft_entry__.build()")
    return c, d


def test_c_library_exports_exactly_the_reference_abi(libs):
    c, _ = libs
    assert declared("semper_c.h", "SEMPER_C_API") == REFERENCE_EXPORTS
    assert exported(c) == REFERENCE_EXPORTS
    lib = ctypes.CDLL(c)
    for name in REFERENCE_EXPORTS:
        assert getattr(lib, name) is not None


def test_dev_library_exports_its_header(libs):
    _, d = libs
    sdk = declared("semper_dev.h", "SDK_API")
    assert sdk and all(n.startswith("sdk_") for n in sdk)
    ex = exported(d)
    assert ex == sdk | REFERENCE_EXPORTS, sorted(ex ^ (sdk | REFERENCE_EXPORTS))
    lib = ctypes.CDLL(d)
    for name in sdk:
        assert getattr(lib, name) is not None


def test_frozen_layouts_and_version(libs):
    import importlib.util
    spec = importlib.util.spec_from_file_location("semper_amd", os.path.join(PKG, "semper_amd.py"))
    sa = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(sa)
    assert ctypes.sizeof(sa.Params) == 32 and sa.Params.use_6x6_interpolator.offset == 28
    assert ctypes.sizeof(sa.IncInfo) == 24 and sa.IncInfo.lost_this_frame.offset == 20
    assert set(sa.EXPORTS) == REFERENCE_EXPORTS
    text = open(os.path.join(ROOT, "include", "semper_c.h")).read()
    for name, val in (("SEMPER_FLOATS_PER_POINT", 8), ("SEMPER_METRICS_LEN", 24), ("SEMPER_ERR_ROI", -2),
                      ("SEMPER_ERR_INIT", -3), ("SEMPER_ERR_STRAIN_WINDOW", -4), ("SEMPER_ERR_SEED", -5),
                      ("SEMPER_ERR_CANCELLED", -99)):
        m = re.search(r"#define\s+" + name + r"\s+\(?(-?\d+)\)?", text)
        assert m and int(m.group(1)) == val, name
    lib = sa.load_c()
    v = lib.semper_version().decode()
    assert re.fullmatch(r"\d+\.\d+\.\d+", v) and v == "0.6.1"
