"""The C-ABI library loads and exports every entry point its header declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import subprocess

import pytest

import kevernals_amd as K


def test_header_declares_entry_points():
    syms = K.declared_symbols()
    for s in ("kv_create", "kv_set_geometry", "kv_forward_project", "kv_back_project", "kv_mlem_iteration",
              "kv_sart_iteration", "kv_shift_and_add", "kv_log_likelihood", "kv_trace_rays"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    lib = K.load_library()
    missing = [s for s in K.declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_exported_symbols_are_declared():
    out = subprocess.run(["nm", "-D", "--defined-only", K.LIB_PATH], capture_output=True, text=True, check=True)
    exported = {ln.split()[-1] for ln in out.stdout.splitlines() if " T kv_" in ln}
    assert exported == set(K.declared_symbols())


def test_version_and_error_string_without_gpu():
    lib = K.load_library()
    assert b"gfx950" in lib.kv_version()
    assert lib.kv_last_error() == b""


def test_gfx950_code_object_present():
    out = subprocess.run(["/opt/rocm/bin/roc-obj-ls", K.LIB_PATH], capture_output=True, text=True)
    if out.returncode != 0:
        pytest.skip("roc-obj-ls unavailable")
    assert "gfx950" in out.stdout


def test_no_cpu_fallback_in_product_package():
    """The product package never imports the oracle."""
    pkg = os.path.dirname(K.__file__)
    for root, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".h")):
                text = open(os.path.join(root, f), errors="ignore").read()
                assert "import oracle" not in text and "from oracle" not in text, f
                assert "liboracle" not in text, f


def test_dropin_library_exports_what_its_header_declares():
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(K.__file__)))
    header = open(os.path.join(root, "include", "kevernals_dropin.h")).read()
    declared = set(re.findall(r"^\s*(?:int|void|const float|kvd_image)\s*\*?\s*(kvd_\w+)\s*\(", header, re.M))
    lib = os.path.join(os.path.dirname(K.__file__), "lib", "libkevernals_dropin.so")
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True)
    exported = {ln.split()[-1] for ln in out.stdout.splitlines() if " T kvd_" in ln}
    assert declared and exported == declared
