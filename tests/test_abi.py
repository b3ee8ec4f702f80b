"""The C-ABI libraries load and export every entry point the headers
declare (no device calls: this runs without a GPU)."""
import ctypes
import os
import re

import gamera_op as go

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^\s*(?:int|size_t|void|double|long long)\s+\**(gk_\w+)\s*\(", text, re.M)))


def test_device_library_exports_header_symbols():
    names = declared("gamera_op.h")
    assert "gk_create" in names and "gk_step" in names
    lib = ctypes.CDLL(go.LIB_OP)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_host_facade_exports_reference_entry_points():
    h, _ = go._load()
    for name in ["initialize_solver", "solve", "finalize_solver", "getweights",
                 "set_background_field", "set_geometry_arrays", "AinitB",
                 "GaussianLineIntegral", "read_config", "initialize_config",
                 "initialize_logging", "get_dt", "get_derived_variables",
                 "gas_bc_symmetric_i_low", "gem_bc_symmetric_k_high",
                 "set_problem_config", "problem_init", "boundary_conditions"]:
        assert hasattr(h, name), name
    for var in ["gas", "gem", "geo", "rec", "x1", "x1c", "x1ctr", "dx1", "config", "is", "ie",
                "CA", "problem_config"]:
        assert hasattr(h, var), var
