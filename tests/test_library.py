"""The C-ABI library loads and exports every symbol its headers declare (no
GPU needed: nothing here launches a kernel)."""
import ctypes
import os
import re
import subprocess

import pytest

from jsi_cad_next_versions_amd import runtime

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "jsi_mi355x.h")

# C++ symbols of the VERSION5 API (include/jsi_cad.hpp), mangled as the
# reference's own objects export them, so drivers relink unchanged.
LEGACY = ["_Z8init_cadv", "_Z10deinit_cadv", "_Z23allocate_from_workspacem",
          "_Z16minimum_distanceRK20EvalAndConstructTaskS1_", "_Z12intersection20EvalAndConstructTaskS_",
          "_Z12evalWithCudaPiPfiffffiS0_", "_Z9cadMallocPPvm", "_Z7cadFreePv",
          "_Z9cadMemcpyPvPKvm13CadMemcpyKind", "_Z9cadMemsetPvim", "_Z15cadIsAclEnabledv",
          "_Z12PrefixSumAddPKiPim", "_Z9compute_KPfS_S_i"]


def header_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(jsi_\w+)\s*\(", text, re.M)))


def test_header_and_binding_agree():
    assert header_functions() == sorted(runtime.EXPORTS)


def test_library_exports_all_symbols():
    if not os.path.exists(runtime.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(runtime.LIB_PATH)
    for name in header_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", runtime.LIB_PATH], capture_output=True, text=True).stdout
    for sym in LEGACY + runtime.OP_EXPORTS:
        assert re.search(rf"\bT {sym}$", out, re.M), sym


def test_params_default_matches_version5():
    if not os.path.exists(runtime.LIB_PATH):
        pytest.skip("library not built")
    L = runtime.load_library()
    p = runtime.Params()
    L.jsi_params_default(ctypes.byref(p))
    d = runtime.default_params()
    for f, _ in runtime.Params._fields_:
        assert getattr(p, f) == getattr(d, f), f
    assert L.jsi_status_string(-4) == b"capacity exceeded"
