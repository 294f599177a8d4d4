"""The C-ABI library loads and exports every entry point include/glx.h
declares; the ctypes prototypes cover exactly those. No device calls."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "glx.h")


def header_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(glx_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = header_symbols()
    assert "glx_td_run" in syms and "glx_free_energy" in syms and len(syms) >= 30


def test_library_exports_every_symbol():
    from svirl_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("libglx.so not built (run `python -m svirl_amd._build`)")
    lib = ctypes.CDLL(_native.LIB_PATH)
    missing = [s for s in header_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_ctypes_prototypes_match_header():
    from svirl_amd import _native
    assert sorted(_native.EXPORTS) == header_symbols()


def test_product_does_not_import_oracle():
    """The shipped package never imports the test oracle."""
    pkg = os.path.join(ROOT, "svirl_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith(".py"):
                src = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in re.findall(r"^\s*(?:from|import)\s+(\w+)", src, flags=re.M), f


def test_build_script_flags():
    """FP64 kernels are built without FMA contraction; all HIP units with
    packed FP32 ops disabled."""
    from svirl_amd import _build
    units = dict(_build.UNITS)
    assert "-ffp-contract=off" in units["glx_f64.hip"]
    assert "-packed-fp32-ops" in _build.COMMON
