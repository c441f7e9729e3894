"""The C-ABI library loads without a GPU and exports every declared entry point."""

import ctypes
import os
import re

from nrpy_amd import runtime

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    with open(os.path.join(ROOT, "include", "bhah_hip.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(bhah_hip_\w+)\s*\(", text)))


def test_header_declares_the_reference_counterparts():
    names = declared_functions()
    for ref in ["rhs_eval", "Ricci_eval", "MoL_step_forward_in_time", "apply_bcs_outerradiation_and_inner",
                "apply_bcs_outerextrap_and_inner", "apply_bcs_inner_only",
                "enforce_detgammabar_equals_detgammahat", "diagnostics_constraints",
                "MoL_malloc_intermediate_stage_gfs"]:
        assert f"bhah_hip_{ref}" in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(runtime.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert set(declared_functions()) == set(runtime.SIGNATURES)


def test_abi_version_call_needs_no_device():
    lib = runtime.load_library()
    assert lib.bhah_hip_abi_version() == 2


def test_null_arguments_are_rejected_without_touching_a_device():
    lib = runtime.load_library()
    assert lib.bhah_hip_patch_create(None, None) == -1
    assert b"NULL" in lib.bhah_hip_last_error()
    assert lib.bhah_hip_rhs_eval(None, None, None, None) == -1
