"""The C ABI library loads and exports every symbol include/qsim_hip.h declares
(no GPU needed; no compute is called)."""

import os
import re

import pytest

import qsim_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qsim_hip_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = header_functions(os.path.join(ROOT, "include", "qsim_hip.h"))
    assert len(names) > 30
    dll = qsim_amd.load_library()
    missing = [n for n in names if not hasattr(dll, n)]
    assert not missing, missing


def test_python_signatures_cover_header():
    names = header_functions(os.path.join(ROOT, "include", "qsim_hip.h"))
    declared = {s[0] for s in qsim_amd.SIGNATURES}
    assert set(names) == declared


def test_min_size_needs_no_gpu():
    dll = qsim_amd.load_library()
    assert dll.qsim_hip_min_size(0) == 64
    assert dll.qsim_hip_min_size(20) == 2 << 20
    assert dll.qsim_hip_num_partial_norms(30) == 1024


def test_library_is_gfx950_code():
    # The shared object carries a gfx950 code object (offload bundle).
    data = open(qsim_amd.LIB_PATH, "rb").read()
    assert b"gfx950" in data
